// The optimizer loop on the device: plm_hip_bspline_optimize_lbfgs.
//
// The algorithm is the one liblbfgs 1.9 runs for the reference
// (libs/liblbfgs-1.9/lib/lbfgs.c:256-653, 823-1330; driven by
// src/register/bspline_optimize_liblbfgs.cxx:83-109): L-BFGS two-loop
// recursion over the last m (s, y) pairs with H0 = (ys/yy) I, and the
// More-Thuente line search with every evaluation reported to the progress
// rule.  Its vectors live in device memory (lbfgs_kernels.hip); this file
// holds the scalar logic -- the line search and the small dense solve of
// the direction.  Per evaluation the host reads back only a few dozen
// totals (the objective, g.d, x.x, g.g, the new pair's products), which a
// one-workgroup pass writes into mapped host memory.
//
// The first trial of each line search is formed by the pass that forms
// the direction, before the host knows the initial slope g.d: the slope
// comes back with that trial's results.  Should the slope be
// positive (liblbfgs returns LBFGSERR_INCREASEGRADIENT before evaluating
// anything), the trial is dropped uncounted -- its coefficients went to the
// spare buffer -- so the run's record matches liblbfgs's.

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "plm_hip_bspline.h"
#include "bspline_internal.h"
#include "lbfgs_kernels.h"


namespace {

constexpr int MAX_PAST = 64;

// One step of the line search: step, value, slope along the direction.
struct Pt {
    double st, f, g;
};

double
lmax (double a, double b)
{
    return a >= b ? a : b;
}

double
lmin (double a, double b)
{
    return a <= b ? a : b;
}

// Minimizer of the cubic interpolating (u, fu, du) and (v, fv, dv)
// (More and Thuente, ACM TOMS 20(3), 1994, section 4; lbfgs.c:1043-1057).
double
cubic_min (double u, double fu, double du, double v, double fv, double dv)
{
    const double span = v - u;
    const double theta = (fu - fv) * 3 / span + du + dv;
    const double s = lmax (lmax (std::fabs (theta), std::fabs (du)), std::fabs (dv));
    const double a = theta / s;
    double gamma = s * std::sqrt (a * a - (du / s) * (dv / s));
    if (v < u) gamma = -gamma;
    const double p = gamma - du + theta;
    const double q = gamma - du + gamma + dv;
    return u + p / q * span;
}

// The cubic step that falls back to tmin / tmax when the cubic has no
// minimizer beyond v (lbfgs.c:1071-1091).
double
cubic_min_safe (double u, double fu, double du, double v, double fv, double dv, double tmin,
    double tmax)
{
    const double span = v - u;
    const double theta = (fu - fv) * 3 / span + du + dv;
    const double s = lmax (lmax (std::fabs (theta), std::fabs (du)), std::fabs (dv));
    const double a = theta / s;
    double gamma = s * std::sqrt (lmax (0, a * a - (du / s) * (dv / s)));
    if (u < v) gamma = -gamma;
    const double p = gamma - dv + theta;
    const double q = gamma - dv + gamma + du;
    const double r = p / q;
    if (r < 0. && gamma != 0.) return v - r * span;
    return a < 0 ? tmax : tmin;
}

// Quadratic through (u, fu, du) and (v, fv); secant through (u, du), (v, dv)
// (lbfgs.c:1103-1118).
double
quad_min (double u, double fu, double du, double v, double fv)
{
    const double a = v - u;
    return u + du / ((fu - fv) / a + du) / 2 * a;
}

double
secant_min (double u, double du, double v, double dv)
{
    const double a = u - v;
    return v + dv / (dv - du) * a;
}

// Next trial step and the new interval of uncertainty [bx, by]
// (update_trial_interval, lbfgs.c:1159-1326).
int
next_trial (Pt& bx, Pt& by, Pt& t, double tmin, double tmax, int& brackt)
{
    const bool opposite = t.g * (bx.g / std::fabs (bx.g)) < 0.;
    bool bound;
    double next;
    if (brackt) {
        if (t.st <= lmin (bx.st, by.st) || lmax (bx.st, by.st) <= t.st)
            return PLM_LBFGSERR_OUTOFINTERVAL;
        if (0. <= bx.g * (t.st - bx.st)) return PLM_LBFGSERR_INCREASEGRADIENT;
        if (tmax < tmin) return PLM_LBFGSERR_INCORRECT_TMINMAX;
    }
    if (bx.f < t.f) {
        brackt = 1;
        bound = true;
        const double mc = cubic_min (bx.st, bx.f, bx.g, t.st, t.f, t.g);
        const double mq = quad_min (bx.st, bx.f, bx.g, t.st, t.f);
        next = std::fabs (mc - bx.st) < std::fabs (mq - bx.st) ? mc : mc + 0.5 * (mq - mc);
    } else if (opposite) {
        brackt = 1;
        bound = false;
        const double mc = cubic_min (bx.st, bx.f, bx.g, t.st, t.f, t.g);
        const double mq = secant_min (bx.st, bx.g, t.st, t.g);
        next = std::fabs (mc - t.st) > std::fabs (mq - t.st) ? mc : mq;
    } else if (std::fabs (t.g) < std::fabs (bx.g)) {
        bound = true;
        const double mc = cubic_min_safe (bx.st, bx.f, bx.g, t.st, t.f, t.g, tmin, tmax);
        const double mq = secant_min (bx.st, bx.g, t.st, t.g);
        if (brackt) next = std::fabs (t.st - mc) < std::fabs (t.st - mq) ? mc : mq;
        else next = std::fabs (t.st - mc) > std::fabs (t.st - mq) ? mc : mq;
    } else {
        bound = false;
        if (brackt) next = cubic_min (t.st, t.f, t.g, by.st, by.f, by.g);
        else next = bx.st < t.st ? tmax : tmin;
    }
    if (bx.f < t.f) {
        by = t;
    } else {
        if (opposite) by = bx;
        bx = t;
    }
    if (tmax < next) next = tmax;
    if (next < tmin) next = tmin;
    if (brackt && bound) {
        const double mq = bx.st + 0.66 * (by.st - bx.st);
        if (bx.st < by.st) {
            if (mq < next) next = mq;
        } else {
            if (next < mq) next = mq;
        }
    }
    t.st = next;
    return 0;
}

// Device and pinned-host state of one run.
struct Run {
    plm_hip_ctx* c = nullptr;
    hipStream_t st = nullptr;
    int64_t n = 0;
    int nb = 0;
    int metric = 0, grad_mode = 0;
    const plm_reg_parms* reg = nullptr;
    plm_lbfgs_parms p {};
    plm_lbfgs_progress_fn progress = nullptr;
    void* inst = nullptr;

    double* X[2] = {nullptr, nullptr};      // x and xp, by index xi / xpi
    double* G[2] = {nullptr, nullptr};      // g and gp, by index gi / gpi
    float* C[2] = {nullptr, nullptr};       // coefficients of the evaluations
    double* d = nullptr;
    double* S = nullptr;                    // history slot j at S + j n
    double* Y = nullptr;
    double* score = nullptr;                // device: score, reg score, num_vox
    double* part = nullptr;                 // post partials, component c at part + c nb
    double* slope_part = nullptr;           // combine's partials of g.d
    double* hh = nullptr;                   // pinned, written by k_lb_finish: score, reg score,
                                            // num_vox, slope, post component totals
    double* hh_dev = nullptr;               // its device address
    int xi = 0, xpi = 1, gi = 0, gpi = 1;
    int c_last = -1;        // C buffer of the last counted evaluation
    int c_trial = 0;        // C buffer of the trial in flight
    int feval = 0, it = 0;
    int err = PLM_HIP_OK;   // first C-ABI error; the run stops on it
    double f = 0., dg = 0., xx = 0., gg = 0.;   // the last trial's values
    // what the trials of the current search do: the first one forms the
    // direction from dir_b pairs (newest in slot dir_e); each one writes
    // its would-be pair into slot post_e with post_bp earlier pairs
    int dir_b = 0, dir_e = 0, post_e = -1, post_bp = 0;
    int ncomp = 0;                          // post components held: 7 + 5 (m - 1)
    // Gram entries of the history by slot: SY[i m + j] = s_i.y_j, YY = y_i.y_j
    std::vector<double> SY, YY;
    plmhip::LbCombineArgs dir {};           // the next direction's coefficients

    ~Run ()
    {
        void* bufs[] = {X[0], X[1], G[0], G[1], C[0], C[1], d, S, Y, score, part, slope_part};
        for (void* b : bufs) if (b) (void) hipFree (b);
        if (hh) (void) hipHostFree (hh);
    }

    size_t hh_doubles () const { return 4 + (size_t) ncomp; }

    int hip (hipError_t e, const char* what)
    {
        if (e != hipSuccess && err == PLM_HIP_OK) err = plmhip::abi_hip_fail (e, what);
        return err;
    }

    template <typename T>
    int alloc (T** ptr, size_t count, const char* what)
    {
        return hip (hipMalloc ((void**) ptr, sizeof (T) * (count ? count : 1)), what);
    }

    // One evaluation: its point (kind 0: the start point d_x0; 1: the
    // search's first trial, formed with the direction; 2: a later trial
    // xp + stp d), the objective and gradient there, the post pass; wait
    // and read back.  Not yet counted.
    int trial (int kind, double stp, const double* x0 = nullptr)
    {
        c_trial = c_last == 0 ? 1 : 0;
        float* cf = C[c_trial];
        hipError_t e;
        if (kind == 0) {
            e = plmhip::lb_launch_trial (x0, nullptr, 0., X[xi], cf, n, st);
        } else if (kind == 1) {
            plmhip::LbCombineArgs a = dir;
            a.b = dir_b;
            a.S = S;
            a.Y = Y;
            a.g = G[gpi];
            a.xp = X[xpi];
            a.stp = stp;
            a.d = d;
            a.x = X[xi];
            a.coeff = cf;
            a.slope = slope_part;
            e = plmhip::lb_launch_combine (a, n, st);
        } else {
            e = plmhip::lb_launch_trial (X[xpi], d, stp, X[xi], cf, n, st);
        }
        if (hip (e, "L-BFGS trial point")) return err;
        int rc = plm_hip_bspline_score_device (c, metric, grad_mode, cf, score, G[gi],
            reinterpret_cast<int64_t*> (score + 2), nullptr, st);
        if (rc < 0) return err = rc;
        if (reg) {
            rc = plm_hip_bspline_regularize (c, reg, cf, score + 1, G[gi], 1, st);
            if (rc < 0) return err = rc;
        }
        plmhip::LbPostArgs pa {X[xi], X[xpi], G[gi], G[gpi], kind == 0 ? nullptr : d, S, Y,
            kind == 0 ? -1 : post_e, post_bp, p.m, part};
        if (hip (plmhip::lb_launch_post (pa, n, st), "L-BFGS post pass")) return err;
        const int nc = kind == 0 ? 3 : plmhip::lb_post_components (post_bp);
        if (hip (plmhip::lb_launch_finish (part, nc, nb, slope_part, score, hh_dev, st), "L-BFGS totals"))
            return err;
        if (hip (hipStreamSynchronize (st), "L-BFGS evaluation")) return err;
        f = reg ? hh[0] + hh[1] : hh[0];
        dg = hh[4];
        xx = hh[5];
        gg = hh[6];
        return PLM_HIP_OK;
    }

    void count ()
    {
        c_last = c_trial;
        ++feval;
    }

    // The progress report: the caller's callback, then the reference's
    // max_its / max_feval rule.
    int report (double fx, double xnorm, double gnorm, double step, int k, int ls)
    {
        it = k;
        int ret = 0;
        if (progress)
            ret = progress (inst, X[xi], G[gi], fx, xnorm, gnorm, step, (int) n, k, ls);
        if (!ret && ((p.max_its >= 0 && it > p.max_its) || (p.max_feval >= 0 && feval > p.max_feval)))
            ret = 1;
        return ret;
    }

    // The coefficients of the next direction from the accepted trial's post
    // partials: the newest pair (slot e) and the b - 1 before it.  In age
    // order (position 0 the oldest), R the upper triangle of S'Y, D its
    // diagonal, gamma = ys / yy of the newest pair, p = S'g, q = Y'g:
    //   u = R^-1 p,  a = R^-T ((D + gamma Y'Y) u - gamma q),
    //   H g = gamma g + S a - gamma Y u     (the compact form; its d = -H g
    // is the two-loop recursion's).  The newest pair's Gram row and column
    // are stored for the later iterations.
    void solve_direction (int b, int e)
    {
        const int m = p.m;
        dir_b = b;
        dir_e = e;
        // post components: 3 s_e.y_e, 4 y_e.y_e, 5 s_e.g, 6 y_e.g; for the
        // pair k older than the newest, 7 + 5 (k - 1) + {s_e.y_j, s_j.y_e,
        // y_e.y_j, s_j.g, y_j.g}
        auto red = [&] (int c) { return hh[4 + c]; };
        SY[e * m + e] = red (3);
        YY[e * m + e] = red (4);
        double P[plmhip::LB_MAX_M], Q[plmhip::LB_MAX_M];
        int o[plmhip::LB_MAX_M];
        for (int i = 0; i < b; ++i) o[i] = (e - (b - 1) + i + m) % m;
        P[b - 1] = red (5);
        Q[b - 1] = red (6);
        for (int k = 1; k < b; ++k) {
            const int j = (e - k + m) % m;
            const int c = 7 + 5 * (k - 1);
            SY[e * m + j] = red (c);
            SY[j * m + e] = red (c + 1);
            YY[e * m + j] = YY[j * m + e] = red (c + 2);
            P[b - 1 - k] = red (c + 3);
            Q[b - 1 - k] = red (c + 4);
        }
        auto sy = [&] (int i, int j) { return SY[o[i] * m + o[j]]; };
        auto yy = [&] (int i, int j) { return YY[o[i] * m + o[j]]; };
        const double gamma = sy (b - 1, b - 1) / yy (b - 1, b - 1);
        double u[plmhip::LB_MAX_M], t[plmhip::LB_MAX_M];
        for (int i = b - 1; i >= 0; --i) {
            double acc = P[i];
            for (int j = i + 1; j < b; ++j) acc = acc - sy (i, j) * u[j];
            u[i] = acc / sy (i, i);
        }
        for (int i = 0; i < b; ++i) {
            double w = 0.0;
            for (int j = 0; j < b; ++j) w = w + yy (i, j) * u[j];
            t[i] = sy (i, i) * u[i] + gamma * w - gamma * Q[i];
        }
        for (int i = 0; i < b; ++i) {
            double acc = t[i];
            for (int j = 0; j < i; ++j) acc = acc - sy (j, i) * dir.ca[j];
            dir.ca[i] = acc / sy (i, i);
            dir.cu[i] = u[i];
            dir.slot[i] = o[i];
        }
        dir.gamma = gamma;
    }

    // The More-Thuente search along d from (xp, gp), the current point
    // being already moved to the xp / gp buffers; d itself is formed by the
    // first trial.  Returns the evaluations it took or a negative liblbfgs
    // code; err set on a C-ABI failure.  The initial slope g.d arrives with
    // the first trial's results.
    int search (double* fx, double* stp, int k)
    {
        if (*stp <= 0.) return PLM_LBFGSERR_INVALIDPARAMETERS;
        int count_ = 0, brackt = 0, stage1 = 1, uinfo = 0;
        const double finit = *fx;
        double dginit = 0., dgtest = 0.;
        double width = p.max_step - p.min_step;
        double prev_width = 2.0 * width;
        Pt bx = {0., finit, 0.}, by = {0., finit, 0.};
        double stmin, stmax;
        for (;;) {
            if (brackt) {
                stmin = lmin (bx.st, by.st);
                stmax = lmax (bx.st, by.st);
            } else {
                stmin = bx.st;
                stmax = *stp + 4.0 * (*stp - bx.st);
            }
            if (*stp < p.min_step) *stp = p.min_step;
            if (p.max_step < *stp) *stp = p.max_step;
            if ((brackt && ((*stp <= stmin || stmax <= *stp) || p.max_linesearch <= count_ + 1
                     || uinfo != 0))
                || (brackt && (stmax - stmin <= p.xtol * stmax)))
                *stp = bx.st;

            if (trial (count_ == 0 ? 1 : 2, *stp)) return PLM_LBFGSERR_INVALIDPARAMETERS;
            if (count_ == 0) {
                dginit = hh[3];
                if (0 < dginit) return PLM_LBFGSERR_INCREASEGRADIENT;    // trial dropped
                dgtest = p.ftol * dginit;
                bx.g = by.g = dginit;
            }
            count ();
            *fx = f;
            if (report (bx.f, std::sqrt (xx), std::sqrt (gg), *stp, k, 0))
                return PLM_LBFGSERR_MAXIMUMLINESEARCH;
            const double ftest1 = finit + *stp * dgtest;
            ++count_;

            if (brackt && ((*stp <= stmin || stmax <= *stp) || uinfo != 0))
                return PLM_LBFGSERR_ROUNDING_ERROR;
            if (*stp == p.max_step && f <= ftest1 && dg <= dgtest) return PLM_LBFGSERR_MAXIMUMSTEP;
            if (*stp == p.min_step && (ftest1 < f || dgtest <= dg)) return PLM_LBFGSERR_MINIMUMSTEP;
            if (brackt && (stmax - stmin) <= p.xtol * stmax) return PLM_LBFGSERR_WIDTHTOOSMALL;
            if (p.max_linesearch <= count_) return PLM_LBFGSERR_MAXIMUMLINESEARCH;
            if (f <= ftest1 && std::fabs (dg) <= p.gtol * (-dginit)) return count_;

            if (stage1 && f <= ftest1 && lmin (p.ftol, p.gtol) * dginit <= dg) stage1 = 0;

            Pt t = {*stp, f, dg};
            if (stage1 && ftest1 < f && f <= bx.f) {
                Pt mx = {bx.st, bx.f - bx.st * dgtest, bx.g - dgtest};
                Pt my = {by.st, by.f - by.st * dgtest, by.g - dgtest};
                Pt mt = {t.st, t.f - t.st * dgtest, t.g - dgtest};
                uinfo = next_trial (mx, my, mt, stmin, stmax, brackt);
                bx = {mx.st, mx.f + mx.st * dgtest, mx.g + dgtest};
                by = {my.st, my.f + my.st * dgtest, my.g + dgtest};
                *stp = mt.st;
            } else {
                uinfo = next_trial (bx, by, t, stmin, stmax, brackt);
                *stp = t.st;
            }
            if (brackt) {
                if (0.66 * prev_width <= std::fabs (by.st - bx.st)) *stp = bx.st + 0.5 * (by.st - bx.st);
                prev_width = width;
                width = std::fabs (by.st - bx.st);
            }
        }
    }
};

int
check_parms (const plm_lbfgs_parms& p)
{
    if (p.m < 1 || p.m > plmhip::LB_MAX_M)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "L-BFGS m must be 1..16");
    if (p.past < 0 || p.past > MAX_PAST)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "L-BFGS past must be 0..64");
    if (!(p.epsilon >= 0.) || !(p.delta >= 0.) || !(p.min_step >= 0.) || !(p.max_step >= p.min_step)
        || !(p.ftol >= 0.) || !(p.gtol >= 0.) || !(p.xtol >= 0.) || p.max_linesearch <= 0)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "invalid L-BFGS parameter (as lbfgs() rejects it)");
    return PLM_HIP_OK;
}

} // namespace

extern "C" {

void
plm_hip_lbfgs_parms_init (plm_lbfgs_parms* p)
{
    if (!p) return;
    p->m = 6;
    p->epsilon = 1e-5;
    p->past = 0;
    p->delta = 1e-5;
    p->max_iterations = 0;
    p->max_linesearch = 40;
    p->min_step = 1e-20;
    p->max_step = 1e20;
    p->ftol = 1e-4;
    p->gtol = 0.9;
    p->xtol = 1.0e-16;
    p->max_its = 10;
    p->max_feval = 10;
}

int
plm_hip_bspline_optimize_lbfgs (plm_hip_ctx* c, int metric, int grad_mode,
    const plm_reg_parms* reg, const plm_lbfgs_parms* parms, double* d_x, float* d_coeff,
    plm_lbfgs_progress_fn progress, void* instance, plm_lbfgs_result* result)
{
    if (!c || !d_x || !result) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "NULL argument");
    if (metric == PLM_METRIC_MSE && grad_mode == PLM_GRAD_NONE)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "the optimizer needs a gradient");
    Run r;
    int device = 0;
    int64_t knots = 0;
    int rc = plmhip::ctx_info (c, &device, &r.st, &knots);
    if (rc) return rc;
    if (parms) r.p = *parms;
    else plm_hip_lbfgs_parms_init (&r.p);
    rc = check_parms (r.p);
    if (rc) return rc;
    if (r.hip (hipSetDevice (device), "hipSetDevice")) return r.err;
    r.c = c;
    r.n = 3 * knots;
    r.nb = plmhip::lb_blocks (r.n);
    r.ncomp = plmhip::lb_post_components (r.p.m - 1);
    r.SY.assign ((size_t) r.p.m * r.p.m, 0.0);
    r.YY.assign ((size_t) r.p.m * r.p.m, 0.0);
    r.metric = metric;
    r.grad_mode = grad_mode;
    r.reg = reg;
    r.progress = progress;
    r.inst = instance;
    const int m = r.p.m;
    const size_t n = (size_t) r.n;
    if (r.alloc (&r.X[0], n, "hipMalloc(L-BFGS x)") || r.alloc (&r.X[1], n, "hipMalloc(L-BFGS x)")
        || r.alloc (&r.G[0], n, "hipMalloc(L-BFGS g)") || r.alloc (&r.G[1], n, "hipMalloc(L-BFGS g)")
        || r.alloc (&r.C[0], n, "hipMalloc(L-BFGS coeff)") || r.alloc (&r.C[1], n, "hipMalloc(L-BFGS coeff)")
        || r.alloc (&r.d, n, "hipMalloc(L-BFGS d)") || r.alloc (&r.S, n * m, "hipMalloc(L-BFGS s)")
        || r.alloc (&r.Y, n * m, "hipMalloc(L-BFGS y)") || r.alloc (&r.score, 3, "hipMalloc(L-BFGS score)")
        || r.alloc (&r.part, (size_t) r.ncomp * r.nb, "hipMalloc(L-BFGS partials)")
        || r.alloc (&r.slope_part, (size_t) r.nb, "hipMalloc(L-BFGS slope partials)"))
        return r.err;
    if (r.hip (hipHostMalloc ((void**) &r.hh, sizeof (double) * r.hh_doubles (), hipHostMallocMapped),
            "hipHostMalloc(L-BFGS readback)")
        || r.hip (hipHostGetDevicePointer ((void**) &r.hh_dev, r.hh, 0), "hipHostGetDevicePointer"))
        return r.err;
    // the regularizer adds into the gradient and leaves its score; without
    // it the reg score slot stays zero
    if (r.hip (hipMemsetAsync (r.score, 0, sizeof (double) * 3, r.st), "L-BFGS score clear")
        || r.hip (hipMemsetAsync (r.slope_part, 0, sizeof (double) * r.nb, r.st), "L-BFGS slope clear"))
        return r.err;

    std::vector<double> pf ((size_t) (r.p.past > 0 ? r.p.past : 1));
    int ret = 0;
    double fx = 0.;
    double step = 0.;
    int k = 0, end = 0;

    // the start point, evaluated
    if (r.trial (0, 0., d_x)) return r.err;
    r.count ();
    fx = r.f;
    pf[0] = fx;
    {
        double xnorm = std::sqrt (r.xx);
        const double gnorm = std::sqrt (r.gg);
        if (xnorm < 1.0) xnorm = 1.0;
        if (gnorm / xnorm <= r.p.epsilon) {
            ret = PLM_LBFGS_ALREADY_MINIMIZED;
            goto done;
        }
    }
    step = 1.0 / std::sqrt (r.gg);      // 1 / |d| with d = -g
    k = 1;
    r.dir_b = 0;                        // d = -g
    for (;;) {
        std::swap (r.xi, r.xpi);
        std::swap (r.gi, r.gpi);
        r.post_e = end;
        r.post_bp = (m <= k ? m : k) - 1;
        const int ls = r.search (&fx, &step, k);
        if (r.err) return r.err;
        if (ls < 0) {
            std::swap (r.xi, r.xpi);
            std::swap (r.gi, r.gpi);
            ret = ls;
            goto done;
        }
        double xnorm = std::sqrt (r.xx);
        const double gnorm = std::sqrt (r.gg);
        if ((ret = r.report (fx, xnorm, gnorm, step, k, ls))) goto done;
        if (xnorm < 1.0) xnorm = 1.0;
        if (gnorm / xnorm <= r.p.epsilon) {
            ret = PLM_LBFGS_SUCCESS;
            break;
        }
        if (r.p.past > 0) {
            if (r.p.past <= k) {
                const double rate = (pf[(size_t) (k % r.p.past)] - fx) / fx;
                if (rate < r.p.delta) {
                    ret = PLM_LBFGS_STOP;
                    break;
                }
            }
            pf[(size_t) (k % r.p.past)] = fx;
        }
        if (r.p.max_iterations != 0 && r.p.max_iterations < k + 1) {
            ret = PLM_LBFGSERR_MAXIMUMITERATION;
            break;
        }
        // the accepted trial's pair (slot end) is the newest of the next
        // direction, with the bound - 1 before it
        r.solve_direction (m <= k ? m : k, end);
        ++k;
        end = (end + 1) % m;
        step = 1.0;
    }

done:
    if (r.hip (hipMemcpyAsync (d_x, r.X[r.xi], sizeof (double) * n, hipMemcpyDeviceToDevice, r.st),
            "L-BFGS x out"))
        return r.err;
    if (d_coeff && r.c_last >= 0
        && r.hip (hipMemcpyAsync (d_coeff, r.C[r.c_last], sizeof (float) * n, hipMemcpyDeviceToDevice, r.st),
               "L-BFGS coefficients out"))
        return r.err;
    if (r.hip (hipStreamSynchronize (r.st), "L-BFGS end")) return r.err;
    result->status = ret;
    result->iterations = r.it;
    result->fevals = r.feval;
    result->fx = fx;
    return PLM_HIP_OK;
}

} // extern "C"
