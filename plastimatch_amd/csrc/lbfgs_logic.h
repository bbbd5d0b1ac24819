// The scalar logic of the device-resident L-BFGS (internal): liblbfgs's
// lbfgs() main loop and More-Thuente line search
// (libs/liblbfgs-1.9/lib/lbfgs.c:256-653, 823-1326) as the reference runs
// them (src/register/bspline_optimize_liblbfgs.cxx:83-109, every
// line-search evaluation reported to progress), restated as a state machine
// that advances once per evaluation: LbMachine::advance takes the totals of
// the evaluation just made and leaves in LbMachine::a what the next one
// does (or that the run has ended).  The same code runs on the host, when
// a progress callback needs each report as it comes, and on one device
// thread after every evaluation, when the whole loop stays on the GPU.
// Only IEEE double operations that round the same on both (+ - * /, sqrt,
// fabs; no contraction) are used, so both give the same bits.
#ifndef PLM_LBFGS_LOGIC_H
#define PLM_LBFGS_LOGIC_H

#include <hip/hip_runtime.h>
#include <math.h>

#include "plm_hip_bspline.h"
#include "lbfgs_kernels.h"

namespace plmhip {

constexpr int LB_MAX_PAST = 64;
constexpr int LB_CONTINUE = -2000000;   // the search goes on (not a liblbfgs code)

// What the next evaluation chain does; every vector kernel of the chain
// reads it (from device memory).
struct LbAct {
    int done;               // the run has ended: the kernels of a chain return at once
    int kind;               // 0 the start point, 1 a new direction and its first trial, 2 a trial
    double stp;             // trial step along d
    int xi, xpi, gi, gpi;   // roles of the two x and two g buffers
    int post_e;             // history slot of the trial's would-be pair (-1: none)
    int post_bp;            // earlier pairs the post pass forms products with
    int ncomp;              // post components the totals hold
    // the direction of kind 1: d = -gamma g - sum_i ca_i s_i + gamma sum_i
    // cu_i y_i over the dir_b pairs in slots slot[0..dir_b) (oldest first);
    // dir_b = 0: d = -g
    int dir_b;
    int slot[LB_MAX_M];
    double ca[LB_MAX_M];
    double cu[LB_MAX_M];
    double gamma;
};

// One progress report (lbfgs.c:903-918 and the main loop's).
struct LbReport {
    double fx, xnorm, gnorm, step;
    int k, ls;
};

// One step of the line search: step, value, slope along the direction.
struct LbPt {
    double st, f, g;
};

__host__ __device__ inline double
lb_max (double a, double b)
{
    return a >= b ? a : b;
}

__host__ __device__ inline double
lb_min (double a, double b)
{
    return a <= b ? a : b;
}

// Minimizer of the cubic interpolating (u, fu, du) and (v, fv, dv)
// (More and Thuente, ACM TOMS 20(3), 1994, section 4; lbfgs.c:1043-1057).
__host__ __device__ inline double
lb_cubic_min (double u, double fu, double du, double v, double fv, double dv)
{
    const double span = v - u;
    const double theta = (fu - fv) * 3 / span + du + dv;
    const double s = lb_max (lb_max (fabs (theta), fabs (du)), fabs (dv));
    const double a = theta / s;
    double gamma = s * sqrt (a * a - (du / s) * (dv / s));
    if (v < u) gamma = -gamma;
    const double p = gamma - du + theta;
    const double q = gamma - du + gamma + dv;
    return u + p / q * span;
}

// The cubic step that falls back to tmin / tmax when the cubic has no
// minimizer beyond v (lbfgs.c:1071-1091).
__host__ __device__ inline double
lb_cubic_min_safe (double u, double fu, double du, double v, double fv, double dv, double tmin,
    double tmax)
{
    const double span = v - u;
    const double theta = (fu - fv) * 3 / span + du + dv;
    const double s = lb_max (lb_max (fabs (theta), fabs (du)), fabs (dv));
    const double a = theta / s;
    double gamma = s * sqrt (lb_max (0, a * a - (du / s) * (dv / s)));
    if (u < v) gamma = -gamma;
    const double p = gamma - dv + theta;
    const double q = gamma - dv + gamma + du;
    const double r = p / q;
    if (r < 0. && gamma != 0.) return v - r * span;
    return a < 0 ? tmax : tmin;
}

// Quadratic through (u, fu, du) and (v, fv); secant through (u, du), (v, dv)
// (lbfgs.c:1103-1118).
__host__ __device__ inline double
lb_quad_min (double u, double fu, double du, double v, double fv)
{
    const double a = v - u;
    return u + du / ((fu - fv) / a + du) / 2 * a;
}

__host__ __device__ inline double
lb_secant_min (double u, double du, double v, double dv)
{
    const double a = u - v;
    return v + dv / (dv - du) * a;
}

// Next trial step and the new interval of uncertainty [bx, by]
// (update_trial_interval, lbfgs.c:1159-1326).
__host__ __device__ inline int
lb_next_trial (LbPt& bx, LbPt& by, LbPt& t, double tmin, double tmax, int& brackt)
{
    const bool opposite = t.g * (bx.g / fabs (bx.g)) < 0.;
    bool bound;
    double next;
    if (brackt) {
        if (t.st <= lb_min (bx.st, by.st) || lb_max (bx.st, by.st) <= t.st)
            return PLM_LBFGSERR_OUTOFINTERVAL;
        if (0. <= bx.g * (t.st - bx.st)) return PLM_LBFGSERR_INCREASEGRADIENT;
        if (tmax < tmin) return PLM_LBFGSERR_INCORRECT_TMINMAX;
    }
    if (bx.f < t.f) {
        brackt = 1;
        bound = true;
        const double mc = lb_cubic_min (bx.st, bx.f, bx.g, t.st, t.f, t.g);
        const double mq = lb_quad_min (bx.st, bx.f, bx.g, t.st, t.f);
        next = fabs (mc - bx.st) < fabs (mq - bx.st) ? mc : mc + 0.5 * (mq - mc);
    } else if (opposite) {
        brackt = 1;
        bound = false;
        const double mc = lb_cubic_min (bx.st, bx.f, bx.g, t.st, t.f, t.g);
        const double mq = lb_secant_min (bx.st, bx.g, t.st, t.g);
        next = fabs (mc - t.st) > fabs (mq - t.st) ? mc : mq;
    } else if (fabs (t.g) < fabs (bx.g)) {
        bound = true;
        const double mc = lb_cubic_min_safe (bx.st, bx.f, bx.g, t.st, t.f, t.g, tmin, tmax);
        const double mq = lb_secant_min (bx.st, bx.g, t.st, t.g);
        if (brackt) next = fabs (t.st - mc) < fabs (t.st - mq) ? mc : mq;
        else next = fabs (t.st - mc) > fabs (t.st - mq) ? mc : mq;
    } else {
        bound = false;
        if (brackt) next = lb_cubic_min (t.st, t.f, t.g, by.st, by.f, by.g);
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

// The run's state.  Totals of an evaluation (hh): 0 score, 1 regularization
// score, 2 overlap count, 3 the slope of a new direction's first trial,
// 4 + c post component c (0 g.d, 1 x.x, 2 g.g, 3 s_e.y_e, 4 y_e.y_e,
// 5 s_e.g, 6 y_e.g, then five per earlier pair).
struct LbMachine {
    plm_lbfgs_parms p;
    int use_reg;            // the objective adds the regularization score
    LbAct a;
    int phase;              // 0: the start point is being evaluated; 1: a search trial; 2: done
    int ret;                // the run's status once done
    int k, end, it, feval;
    int evals;              // evaluations advanced over, dropped trials included
    int c_buf;              // x buffer of the last counted evaluation (-1: none)
    double fx, step;
    double f, dg, xx, gg;   // the last evaluation's values
    double pf[LB_MAX_PAST];
    // the line search in progress
    int count_, brackt, stage1, uinfo;
    double finit, dginit, dgtest, width, prev_width, stmin, stmax;
    LbPt bx, by;
    // Gram entries of the history by slot: SY[i m + j] = s_i.y_j, YY = y_i.y_j
    double SY[LB_MAX_M * LB_MAX_M];
    double YY[LB_MAX_M * LB_MAX_M];
    // progress reports: every one is counted in nrep, the first cap kept
    LbReport* rep;
    int cap, nrep;
    // host runs only: the caller's progress callback, asked first at every
    // report (nonzero stops)
    plm_lbfgs_progress_fn progress;
    void* inst;
    const double* cb_x[2];  // the x and g buffers handed to it
    const double* cb_g[2];
    int n;

    // The run before its first evaluation: the start point x0, to be
    // copied into buffer 0.
    __host__ __device__ void
    init (const plm_lbfgs_parms& parms, int reg, int nn)
    {
        p = parms;
        use_reg = reg;
        n = nn;
        a = LbAct {};
        a.kind = 0;
        a.xi = 0;
        a.xpi = 1;
        a.gi = 0;
        a.gpi = 1;
        a.post_e = -1;
        a.ncomp = 3;
        phase = 0;
        ret = 0;
        k = end = it = feval = evals = 0;
        c_buf = -1;
        fx = step = f = dg = xx = gg = 0.;
        count_ = brackt = stage1 = uinfo = 0;
        nrep = 0;
        for (int i = 0; i < LB_MAX_M * LB_MAX_M; i++) SY[i] = YY[i] = 0.;
    }

    __host__ __device__ void
    finish (int r)
    {
        phase = 2;
        ret = r;
        a.done = 1;
    }

    __host__ __device__ void
    count ()
    {
        c_buf = a.xi;
        ++feval;
    }

    // A progress report: the callback (host runs), then the reference's
    // max_its / max_feval rule (bspline_optimize_liblbfgs.cxx:55-80).
    __host__ __device__ int
    report (double rfx, double xnorm, double gnorm, double rstep, int rk, int ls)
    {
        it = rk;
        if (rep && nrep < cap) rep[nrep] = LbReport {rfx, xnorm, gnorm, rstep, rk, ls};
        ++nrep;
        int r = 0;
        if (progress) r = progress (inst, cb_x[a.xi], cb_g[a.gi], rfx, xnorm, gnorm, rstep, n, rk, ls);
        if (!r && ((p.max_its >= 0 && it > p.max_its) || (p.max_feval >= 0 && feval > p.max_feval))) r = 1;
        return r;
    }

    // The coefficients of the next direction from the accepted trial's post
    // components: the newest pair (slot e) and the b - 1 before it.  In age
    // order (position 0 the oldest), R the upper triangle of S'Y, D its
    // diagonal, gamma = ys / yy of the newest pair, P = S'g, Q = Y'g:
    //   u = R^-1 P,  ca = R^-T ((D + gamma Y'Y) u - gamma Q),
    //   H g = gamma g + S ca - gamma Y u
    // (the compact form of Byrd, Nocedal and Schnabel, Math. Prog. 63,
    // 1994; its d = -H g is liblbfgs's two-loop recursion, lbfgs.c:555-609).
    // The newest pair's Gram row and column are kept for later iterations.
    // B = b is a template argument, so the slot order and the work arrays
    // are registers and every Gram entry's address is known before the
    // arithmetic starts (on the device the machine lives in LDS, where a
    // loop of dependent index-then-value reads is a chain of LDS latencies).
    template <int B>
    __host__ __device__ void
    solve_b (const double* red, int e)
    {
        const int m = p.m;
        int o[B];
        double P[B], Q[B];
        // o[i] = (e - (B - 1) + i) mod m, the slots in age order
        o[0] = (e - (B - 1) + m) % m;
#pragma unroll
        for (int i = 1; i < B; ++i) o[i] = o[i - 1] + 1 == m ? 0 : o[i - 1] + 1;
        SY[e * m + e] = red[3];
        YY[e * m + e] = red[4];
        P[B - 1] = red[5];
        Q[B - 1] = red[6];
#pragma unroll
        for (int kk = 1; kk < B; ++kk) {
            const int j = o[B - 1 - kk];
            const int c = 7 + 5 * (kk - 1);
            SY[e * m + j] = red[c];
            SY[j * m + e] = red[c + 1];
            YY[e * m + j] = YY[j * m + e] = red[c + 2];
            P[B - 1 - kk] = red[c + 3];
            Q[B - 1 - kk] = red[c + 4];
        }
        // s_o(i).y_o(j) and y_o(i).y_o(j) in age order: no store falls
        // between these reads, so they are free to be issued early
        const double* sy = SY;
        const double* yy = YY;
        const double gamma = sy[o[B - 1] * m + o[B - 1]] / yy[o[B - 1] * m + o[B - 1]];
        double u[B], t[B], ca[B];
#pragma unroll
        for (int i = B - 1; i >= 0; --i) {
            double acc = P[i];
#pragma unroll
            for (int j = i + 1; j < B; ++j) acc = acc - sy[o[i] * m + o[j]] * u[j];
            u[i] = acc / sy[o[i] * m + o[i]];
        }
#pragma unroll
        for (int i = 0; i < B; ++i) {
            double w = 0.0;
#pragma unroll
            for (int j = 0; j < B; ++j) w = w + yy[o[i] * m + o[j]] * u[j];
            t[i] = sy[o[i] * m + o[i]] * u[i] + gamma * w - gamma * Q[i];
        }
#pragma unroll
        for (int i = 0; i < B; ++i) {
            double acc = t[i];
#pragma unroll
            for (int j = 0; j < i; ++j) acc = acc - sy[o[j] * m + o[i]] * ca[j];
            ca[i] = acc / sy[o[i] * m + o[i]];
        }
#pragma unroll
        for (int i = 0; i < B; ++i) {
            a.ca[i] = ca[i];
            a.cu[i] = u[i];
            a.slot[i] = o[i];
        }
        a.gamma = gamma;
    }

    // The same solve for any b, in the same order of operations: u in
    // a.cu, t then ca in place in a.ca (ca[i] is formed from t[i] and the
    // ca[j], j < i, already written), P and Q read from the components.
    __host__ __device__ void
    solve_any (const double* red, int b, int e)
    {
        const int m = p.m;
        const int o0 = (e - (b - 1) + m) % m;
        auto o = [&] (int i) { const int s = o0 + i; return s >= m ? s - m : s; };
        auto P = [&] (int i) { return i == b - 1 ? red[5] : red[7 + 5 * (b - 2 - i) + 3]; };
        auto Q = [&] (int i) { return i == b - 1 ? red[6] : red[7 + 5 * (b - 2 - i) + 4]; };
        SY[e * m + e] = red[3];
        YY[e * m + e] = red[4];
        for (int kk = 1; kk < b; ++kk) {
            const int j = o (b - 1 - kk);
            const int c = 7 + 5 * (kk - 1);
            SY[e * m + j] = red[c];
            SY[j * m + e] = red[c + 1];
            YY[e * m + j] = YY[j * m + e] = red[c + 2];
        }
        const double gamma = SY[e * m + e] / YY[e * m + e];
        for (int i = b - 1; i >= 0; --i) {
            double acc = P (i);
            for (int j = i + 1; j < b; ++j) acc = acc - SY[o (i) * m + o (j)] * a.cu[j];
            a.cu[i] = acc / SY[o (i) * m + o (i)];
        }
        for (int i = 0; i < b; ++i) {
            double w = 0.0;
            for (int j = 0; j < b; ++j) w = w + YY[o (i) * m + o (j)] * a.cu[j];
            a.ca[i] = SY[o (i) * m + o (i)] * a.cu[i] + gamma * w - gamma * Q (i);
        }
        for (int i = 0; i < b; ++i) {
            double acc = a.ca[i];
            for (int j = 0; j < i; ++j) acc = acc - SY[o (j) * m + o (i)] * a.ca[j];
            a.ca[i] = acc / SY[o (i) * m + o (i)];
            a.slot[i] = o (i);
        }
        a.gamma = gamma;
    }

    // b <= 6 (liblbfgs's default m is 6) in registers; the register form
    // of longer histories spills
    __host__ __device__ void
    solve_direction (const double* red, int b, int e)
    {
        a.dir_b = b;
        switch (b) {
        case 1: solve_b<1> (red, e); break;
        case 2: solve_b<2> (red, e); break;
        case 3: solve_b<3> (red, e); break;
        case 4: solve_b<4> (red, e); break;
        case 5: solve_b<5> (red, e); break;
        case 6: solve_b<6> (red, e); break;
        default: solve_any (red, b, e); break;
        }
    }

    // The top of the search loop (lbfgs.c:960-985): the interval, the step
    // clamped to it, and the next trial.
    __host__ __device__ void
    search_top ()
    {
        if (brackt) {
            stmin = lb_min (bx.st, by.st);
            stmax = lb_max (bx.st, by.st);
        } else {
            stmin = bx.st;
            stmax = step + 4.0 * (step - bx.st);
        }
        if (step < p.min_step) step = p.min_step;
        if (p.max_step < step) step = p.max_step;
        if ((brackt && ((step <= stmin || stmax <= step) || p.max_linesearch <= count_ + 1 || uinfo != 0))
            || (brackt && (stmax - stmin <= p.xtol * stmax)))
            step = bx.st;
        a.kind = count_ == 0 ? 1 : 2;
        a.stp = step;
        a.ncomp = lb_post_components (a.post_bp);
    }

    // A new iteration: the current point moves to the xp / gp buffers, the
    // search starts from it along the direction its first trial forms
    // (LB_CONTINUE), or fails at once (a liblbfgs code).
    __host__ __device__ int
    begin_iteration ()
    {
        int s = a.xi; a.xi = a.xpi; a.xpi = s;
        s = a.gi; a.gi = a.gpi; a.gpi = s;
        a.post_e = end;
        a.post_bp = (p.m <= k ? p.m : k) - 1;
        if (step <= 0.) return PLM_LBFGSERR_INVALIDPARAMETERS;
        count_ = brackt = uinfo = 0;
        stage1 = 1;
        finit = fx;
        dginit = dgtest = 0.;
        width = p.max_step - p.min_step;
        prev_width = 2.0 * width;
        bx = LbPt {0., finit, 0.};
        by = LbPt {0., finit, 0.};
        phase = 1;
        search_top ();
        return LB_CONTINUE;
    }

    // The search loop's body after a trial (lbfgs.c:987-1035): the
    // search's result, or LB_CONTINUE with the next trial set up.
    __host__ __device__ int
    after_trial (const double* hh)
    {
        if (count_ == 0) {
            dginit = hh[3];
            if (0 < dginit) return PLM_LBFGSERR_INCREASEGRADIENT;   // the trial is dropped
            dgtest = p.ftol * dginit;
            bx.g = by.g = dginit;
        }
        count ();
        fx = f;
        if (report (bx.f, sqrt (xx), sqrt (gg), step, k, 0)) return PLM_LBFGSERR_MAXIMUMLINESEARCH;
        const double ftest1 = finit + step * dgtest;
        ++count_;
        if (brackt && ((step <= stmin || stmax <= step) || uinfo != 0)) return PLM_LBFGSERR_ROUNDING_ERROR;
        if (step == p.max_step && f <= ftest1 && dg <= dgtest) return PLM_LBFGSERR_MAXIMUMSTEP;
        if (step == p.min_step && (ftest1 < f || dgtest <= dg)) return PLM_LBFGSERR_MINIMUMSTEP;
        if (brackt && (stmax - stmin) <= p.xtol * stmax) return PLM_LBFGSERR_WIDTHTOOSMALL;
        if (p.max_linesearch <= count_) return PLM_LBFGSERR_MAXIMUMLINESEARCH;
        if (f <= ftest1 && fabs (dg) <= p.gtol * (-dginit)) return count_;
        if (stage1 && f <= ftest1 && lb_min (p.ftol, p.gtol) * dginit <= dg) stage1 = 0;
        LbPt t = {step, f, dg};
        if (stage1 && ftest1 < f && f <= bx.f) {
            LbPt mx = {bx.st, bx.f - bx.st * dgtest, bx.g - dgtest};
            LbPt my = {by.st, by.f - by.st * dgtest, by.g - dgtest};
            LbPt mt = {t.st, t.f - t.st * dgtest, t.g - dgtest};
            uinfo = lb_next_trial (mx, my, mt, stmin, stmax, brackt);
            bx = LbPt {mx.st, mx.f + mx.st * dgtest, mx.g + dgtest};
            by = LbPt {my.st, my.f + my.st * dgtest, my.g + dgtest};
            step = mt.st;
        } else {
            uinfo = lb_next_trial (bx, by, t, stmin, stmax, brackt);
            step = t.st;
        }
        if (brackt) {
            if (0.66 * prev_width <= fabs (by.st - bx.st)) step = bx.st + 0.5 * (by.st - bx.st);
            prev_width = width;
            width = fabs (by.st - bx.st);
        }
        search_top ();
        return LB_CONTINUE;
    }

    // The main loop after a search that returned ls (lbfgs.c:611-652): the
    // report, the stopping tests, and the next iteration -- LB_CONTINUE
    // once its first trial is set up or the run has ended, else the code
    // its search failed with at once.  red: the accepted trial's post
    // components.
    __host__ __device__ int
    after_search (int ls, const double* red)
    {
        if (ls < 0) {
            // the point reverts to the search's start
            int s = a.xi; a.xi = a.xpi; a.xpi = s;
            s = a.gi; a.gi = a.gpi; a.gpi = s;
            finish (ls);
            return LB_CONTINUE;
        }
        double xnorm = sqrt (xx);
        const double gnorm = sqrt (gg);
        const int r = report (fx, xnorm, gnorm, step, k, ls);
        if (r) {
            finish (r);
            return LB_CONTINUE;
        }
        if (xnorm < 1.0) xnorm = 1.0;
        if (gnorm / xnorm <= p.epsilon) {
            finish (PLM_LBFGS_SUCCESS);
            return LB_CONTINUE;
        }
        if (p.past > 0) {
            if (p.past <= k) {
                const double rate = (pf[k % p.past] - fx) / fx;
                if (rate < p.delta) {
                    finish (PLM_LBFGS_STOP);
                    return LB_CONTINUE;
                }
            }
            pf[k % p.past] = fx;
        }
        if (p.max_iterations != 0 && p.max_iterations < k + 1) {
            finish (PLM_LBFGSERR_MAXIMUMITERATION);
            return LB_CONTINUE;
        }
        // the accepted trial's pair (slot end) is the newest of the next
        // direction, with the bound - 1 before it
        solve_direction (red, p.m <= k ? p.m : k, end);
        ++k;
        end = (end + 1) % p.m;
        step = 1.0;
        return begin_iteration ();
    }

    // After an evaluation with totals hh.
    __host__ __device__ void
    advance (const double* hh)
    {
        if (phase == 2) return;
        ++evals;
        f = use_reg ? hh[0] + hh[1] : hh[0];
        dg = hh[4];
        xx = hh[5];
        gg = hh[6];
        if (phase == 0) {
            count ();
            fx = f;
            pf[0] = fx;
            double xnorm = sqrt (xx);
            const double gnorm = sqrt (gg);
            if (xnorm < 1.0) xnorm = 1.0;
            if (gnorm / xnorm <= p.epsilon) {
                finish (PLM_LBFGS_ALREADY_MINIMIZED);
                return;
            }
            step = 1.0 / sqrt (gg);     // 1 / |d| with d = -g
            k = 1;
            end = 0;
            a.dir_b = 0;                // d = -g
        }
        int ls = phase == 0 ? begin_iteration () : after_trial (hh);
        // a search that ends the iteration (and a next one that fails at once)
        while (ls != LB_CONTINUE) ls = after_search (ls, hh + 4);
    }
};

} // namespace plmhip

#endif
