/* L-BFGS with the More-Thuente line search: CPU restatement (TEST
 * INFRASTRUCTURE).
 *
 * The reference optimizes B-spline coefficients with the liblbfgs 1.9 it
 * vendors (libs/liblbfgs-1.9/lib/lbfgs.c), driven by
 * bspline_optimize_liblbfgs (src/register/bspline_optimize_liblbfgs.cxx:
 * 83-109) with the library's default parameters (lbfgs.c:113-118).  Its
 * copy carries one local change: the line search reports every evaluation
 * to the progress callback (lbfgs.c:903-918), which is how the reference
 * enforces max_feval.  This file restates that algorithm for the path the
 * reference runs -- orthantwise_c = 0, the More-Thuente line search -- so
 * the device optimizer can be checked against it.  Only the tests and the
 * CPU leg of bench.py load it.
 *
 * Arithmetic order: every vector reduction is the sequential sum
 * s = 0; s += a[i] * b[i] over i = 0..n-1, and every scalar expression is
 * evaluated in the order liblbfgs writes it, with -ffp-contract=off; the
 * trajectory is then the reference's bit for bit (checked against a build
 * of the vendored lbfgs.c in the tests).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "plm_oracle.h"

static double
lb_dot (const double* a, const double* b, int n)
{
    double s = 0.;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static double
lb_max (double a, double b)
{
    return a >= b ? a : b;
}

static double
lb_min (double a, double b)
{
    return a <= b ? a : b;
}

/* Minimizer of the cubic through (u, fu, du) and (v, fv, dv): the cubic
 * step of More and Thuente, "Line search algorithms with guaranteed
 * sufficient decrease", ACM TOMS 20(3), 1994, section 4 (lbfgs.c:1043-
 * 1057).  The scale s keeps theta^2 - du*dv from overflowing. */
static double
mt_cubic (double u, double fu, double du, double v, double fv, double dv)
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

/* The cubic step when the cubic may have no minimizer in the direction of
 * the step: then it is tmax or tmin by the sign of theta (lbfgs.c:1071-
 * 1091, the safeguarded form of the same paper). */
static double
mt_cubic_safe (double u, double fu, double du, double v, double fv, double dv,
    double tmin, double tmax)
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

/* Minimizer of the quadratic through (u, fu, du) and (v, fv)
 * (lbfgs.c:1103-1105). */
static double
mt_quad (double u, double fu, double du, double v, double fv)
{
    const double a = v - u;
    return u + du / ((fu - fv) / a + du) / 2 * a;
}

/* Secant step: the zero of the line through (u, du) and (v, dv)
 * (lbfgs.c:1116-1118). */
static double
mt_secant (double u, double du, double v, double dv)
{
    const double a = u - v;
    return v + dv / (dv - du) * a;
}

/* One point of the line search: step, value, derivative along the
 * direction. */
typedef struct {
    double st, f, g;
} lb_pt;

/* The trial-step update of More and Thuente (their four cases) and the
 * update of the interval of uncertainty [bx, by]; t is the current trial
 * and gets the next one.  lbfgs.c:1159-1326 (update_trial_interval). */
static int
mt_update (lb_pt* bx, lb_pt* by, lb_pt* t, double tmin, double tmax, int* brackt)
{
    /* the derivatives at t and bx differ in sign (lbfgs.c:1176 with the
       double form of fsigndiff, arithmetic_ansi.h:34) */
    const int opposite = t->g * (bx->g / fabs (bx->g)) < 0.;
    int bound;
    double next;

    if (*brackt) {
        if (t->st <= lb_min (bx->st, by->st) || lb_max (bx->st, by->st) <= t->st)
            return ORC_LBFGSERR_OUTOFINTERVAL;
        if (0. <= bx->g * (t->st - bx->st)) return ORC_LBFGSERR_INCREASEGRADIENT;
        if (tmax < tmin) return ORC_LBFGSERR_INCORRECT_TMINMAX;
    }

    if (bx->f < t->f) {
        /* higher value: the minimum is bracketed; the cubic step if it is
           the nearer to bx, else halfway to the quadratic one */
        *brackt = 1;
        bound = 1;
        const double mc = mt_cubic (bx->st, bx->f, bx->g, t->st, t->f, t->g);
        const double mq = mt_quad (bx->st, bx->f, bx->g, t->st, t->f);
        next = fabs (mc - bx->st) < fabs (mq - bx->st) ? mc : mc + 0.5 * (mq - mc);
    } else if (opposite) {
        /* lower value, derivatives of opposite sign: bracketed; the step
           farther from t of the cubic and the secant */
        *brackt = 1;
        bound = 0;
        const double mc = mt_cubic (bx->st, bx->f, bx->g, t->st, t->f, t->g);
        const double mq = mt_secant (bx->st, bx->g, t->st, t->g);
        next = fabs (mc - t->st) > fabs (mq - t->st) ? mc : mq;
    } else if (fabs (t->g) < fabs (bx->g)) {
        /* lower value, same sign, smaller derivative */
        bound = 1;
        const double mc = mt_cubic_safe (bx->st, bx->f, bx->g, t->st, t->f, t->g, tmin, tmax);
        const double mq = mt_secant (bx->st, bx->g, t->st, t->g);
        if (*brackt) next = fabs (t->st - mc) < fabs (t->st - mq) ? mc : mq;
        else next = fabs (t->st - mc) > fabs (t->st - mq) ? mc : mq;
    } else {
        /* lower value, same sign, derivative not smaller */
        bound = 0;
        if (*brackt) next = mt_cubic (t->st, t->f, t->g, by->st, by->f, by->g);
        else next = bx->st < t->st ? tmax : tmin;
    }

    /* the interval: [bx, t] after a higher value, else t becomes the best
       point and bx the other end when the derivatives differ in sign */
    if (bx->f < t->f) {
        *by = *t;
    } else {
        if (opposite) *by = *bx;
        *bx = *t;
    }

    if (tmax < next) next = tmax;
    if (next < tmin) next = tmin;
    if (*brackt && bound) {
        const double mq = bx->st + 0.66 * (by->st - bx->st);
        if (bx->st < by->st) {
            if (mq < next) next = mq;
        } else {
            if (next < mq) next = mq;
        }
    }
    t->st = next;
    return 0;
}

typedef struct {
    int n;
    orc_lbfgs_eval_fn eval;
    orc_lbfgs_progress_fn progress;
    void* inst;
    const orc_lbfgs_parms* p;
    int k;               /* iteration number, as the reference's global k */
} lb_run;

/* The More-Thuente line search along s from (xp, gp): x, f, g end at the
 * accepted point; returns the number of evaluations or a negative code.
 * lbfgs.c:823-1034 with the per-evaluation progress report (:903-918). */
static int
mt_search (lb_run* r, double* x, double* f, double* g, const double* s, double* stp,
    const double* xp)
{
    const int n = r->n;
    const orc_lbfgs_parms* p = r->p;
    if (*stp <= 0.) return ORC_LBFGSERR_INVALIDPARAMETERS;
    const double dginit = lb_dot (g, s, n);
    if (0 < dginit) return ORC_LBFGSERR_INCREASEGRADIENT;

    int count = 0, brackt = 0, stage1 = 1, uinfo = 0;
    const double finit = *f;
    const double dgtest = p->ftol * dginit;
    double width = p->max_step - p->min_step;
    double prev_width = 2.0 * width;
    lb_pt bx = {0., finit, dginit}, by = {0., finit, dginit};
    double stmin, stmax;

    for (;;) {
        if (brackt) {
            stmin = lb_min (bx.st, by.st);
            stmax = lb_max (bx.st, by.st);
        } else {
            stmin = bx.st;
            stmax = *stp + 4.0 * (*stp - bx.st);
        }
        if (*stp < p->min_step) *stp = p->min_step;
        if (p->max_step < *stp) *stp = p->max_step;
        /* an unusual end: fall back to the best step so far */
        if ((brackt && ((*stp <= stmin || stmax <= *stp) || p->max_linesearch <= count + 1
                 || uinfo != 0))
            || (brackt && (stmax - stmin <= p->xtol * stmax)))
            *stp = bx.st;

        for (int i = 0; i < n; ++i) x[i] = xp[i];
        for (int i = 0; i < n; ++i) x[i] += *stp * s[i];
        *f = r->eval (r->inst, x, g, n, *stp);
        const double dg = lb_dot (g, s, n);
        if (r->progress) {
            const double xnorm = sqrt (lb_dot (x, x, n));
            const double gnorm = sqrt (lb_dot (g, g, n));
            if (r->progress (r->inst, x, g, bx.f, xnorm, gnorm, *stp, n, r->k, 0))
                return ORC_LBFGSERR_MAXIMUMLINESEARCH;
        }
        const double ftest1 = finit + *stp * dgtest;
        ++count;

        if (brackt && ((*stp <= stmin || stmax <= *stp) || uinfo != 0))
            return ORC_LBFGSERR_ROUNDING_ERROR;
        if (*stp == p->max_step && *f <= ftest1 && dg <= dgtest) return ORC_LBFGSERR_MAXIMUMSTEP;
        if (*stp == p->min_step && (ftest1 < *f || dgtest <= dg)) return ORC_LBFGSERR_MINIMUMSTEP;
        if (brackt && (stmax - stmin) <= p->xtol * stmax) return ORC_LBFGSERR_WIDTHTOOSMALL;
        if (p->max_linesearch <= count) return ORC_LBFGSERR_MAXIMUMLINESEARCH;
        if (*f <= ftest1 && fabs (dg) <= p->gtol * (-dginit)) return count;

        if (stage1 && *f <= ftest1 && lb_min (p->ftol, p->gtol) * dginit <= dg) stage1 = 0;

        lb_pt t = {*stp, *f, dg};
        if (stage1 && ftest1 < *f && *f <= bx.f) {
            /* the modified function f(a) - a * dgtest while no step has met
               the sufficient decrease with a nonnegative slope */
            lb_pt mx = {bx.st, bx.f - bx.st * dgtest, bx.g - dgtest};
            lb_pt my = {by.st, by.f - by.st * dgtest, by.g - dgtest};
            lb_pt mt = {t.st, t.f - t.st * dgtest, t.g - dgtest};
            uinfo = mt_update (&mx, &my, &mt, stmin, stmax, &brackt);
            bx.st = mx.st; bx.f = mx.f + mx.st * dgtest; bx.g = mx.g + dgtest;
            by.st = my.st; by.f = my.f + my.st * dgtest; by.g = my.g + dgtest;
            *stp = mt.st;
        } else {
            uinfo = mt_update (&bx, &by, &t, stmin, stmax, &brackt);
            *stp = t.st;
        }

        if (brackt) {
            if (0.66 * prev_width <= fabs (by.st - bx.st)) *stp = bx.st + 0.5 * (by.st - bx.st);
            prev_width = width;
            width = fabs (by.st - bx.st);
        }
    }
}

void
orc_lbfgs_parms_init (orc_lbfgs_parms* p)
{
    /* lbfgs.c:113-118 (_defparam) */
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
}

/* lbfgs() for orthantwise_c = 0 and the More-Thuente search
 * (lbfgs.c:256-653): two-loop recursion of Nocedal (Math. Comp. 35, 1980,
 * p. 779) over the last m (s, y) pairs, H0 scaled by ys/yy. */
int
orc_lbfgs (int n, double* x, double* ptr_fx, orc_lbfgs_eval_fn eval,
    orc_lbfgs_progress_fn progress, void* inst, const orc_lbfgs_parms* pin)
{
    orc_lbfgs_parms p;
    if (pin) p = *pin; else orc_lbfgs_parms_init (&p);
    if (n <= 0) return ORC_LBFGSERR_INVALID_N;
    if (p.epsilon < 0.) return ORC_LBFGSERR_INVALID_EPSILON;
    if (p.past < 0) return ORC_LBFGSERR_INVALID_TESTPERIOD;
    if (p.delta < 0.) return ORC_LBFGSERR_INVALID_DELTA;
    if (p.min_step < 0.) return ORC_LBFGSERR_INVALID_MINSTEP;
    if (p.max_step < p.min_step) return ORC_LBFGSERR_INVALID_MAXSTEP;
    if (p.ftol < 0.) return ORC_LBFGSERR_INVALID_FTOL;
    if (p.gtol < 0.) return ORC_LBFGSERR_INVALID_GTOL;
    if (p.xtol < 0.) return ORC_LBFGSERR_INVALID_XTOL;
    if (p.max_linesearch <= 0) return ORC_LBFGSERR_INVALID_MAXLINESEARCH;

    const int m = p.m;
    const size_t vb = sizeof (double) * (size_t) n;
    double* xp = calloc (1, vb);
    double* g = calloc (1, vb);
    double* gp = calloc (1, vb);
    double* d = calloc (1, vb);
    double* S = calloc ((size_t) m, vb);        /* s of history slot j at S + j n */
    double* Y = calloc ((size_t) m, vb);
    double* ys_h = calloc ((size_t) m, sizeof (double));
    double* alpha = calloc ((size_t) m, sizeof (double));
    double* pf = p.past > 0 ? calloc ((size_t) p.past, sizeof (double)) : NULL;
    int ret = 0;
    double fx = 0.;
    lb_run r = {n, eval, progress, inst, &p, 0};

    if (!xp || !g || !gp || !d || !S || !Y || !ys_h || !alpha || (p.past > 0 && !pf)) {
        ret = ORC_LBFGSERR_OUTOFMEMORY;
        goto done;
    }

    fx = eval (inst, x, g, n, 0);
    if (pf) pf[0] = fx;
    for (int i = 0; i < n; ++i) d[i] = -g[i];
    {
        double xnorm = sqrt (lb_dot (x, x, n));
        const double gnorm = sqrt (lb_dot (g, g, n));
        if (xnorm < 1.0) xnorm = 1.0;
        if (gnorm / xnorm <= p.epsilon) {
            ret = ORC_LBFGS_ALREADY_MINIMIZED;
            goto done;
        }
    }
    double step = 1.0 / sqrt (lb_dot (d, d, n));

    r.k = 1;
    int end = 0;
    for (;;) {
        memcpy (xp, x, vb);
        memcpy (gp, g, vb);
        const int ls = mt_search (&r, x, &fx, g, d, &step, xp);
        if (ls < 0) {
            memcpy (x, xp, vb);
            memcpy (g, gp, vb);
            ret = ls;
            goto done;
        }
        double xnorm = sqrt (lb_dot (x, x, n));
        const double gnorm = sqrt (lb_dot (g, g, n));
        if (progress && (ret = progress (inst, x, g, fx, xnorm, gnorm, step, n, r.k, ls)))
            goto done;
        if (xnorm < 1.0) xnorm = 1.0;
        if (gnorm / xnorm <= p.epsilon) {
            ret = ORC_LBFGS_SUCCESS;
            break;
        }
        if (pf) {
            if (p.past <= r.k) {
                const double rate = (pf[r.k % p.past] - fx) / fx;
                if (rate < p.delta) {
                    ret = ORC_LBFGS_STOP;
                    break;
                }
            }
            pf[r.k % p.past] = fx;
        }
        if (p.max_iterations != 0 && p.max_iterations < r.k + 1) {
            ret = ORC_LBFGSERR_MAXIMUMITERATION;
            break;
        }

        /* newest pair into slot `end` */
        double* sn = S + (size_t) end * n;
        double* yn = Y + (size_t) end * n;
        for (int i = 0; i < n; ++i) sn[i] = x[i] - xp[i];
        for (int i = 0; i < n; ++i) yn[i] = g[i] - gp[i];
        const double ys = lb_dot (yn, sn, n);
        const double yy = lb_dot (yn, yn, n);
        ys_h[end] = ys;

        const int bound = m <= r.k ? m : r.k;
        ++r.k;
        end = (end + 1) % m;

        /* d = -H g: newest to oldest, scale, oldest to newest */
        for (int i = 0; i < n; ++i) d[i] = -g[i];
        int j = end;
        for (int i = 0; i < bound; ++i) {
            j = (j + m - 1) % m;
            const double* sj = S + (size_t) j * n;
            const double* yj = Y + (size_t) j * n;
            alpha[j] = lb_dot (sj, d, n);
            alpha[j] /= ys_h[j];
            const double c = -alpha[j];
            for (int e = 0; e < n; ++e) d[e] += c * yj[e];
        }
        const double scale = ys / yy;
        for (int e = 0; e < n; ++e) d[e] *= scale;
        for (int i = 0; i < bound; ++i) {
            const double* sj = S + (size_t) j * n;
            const double* yj = Y + (size_t) j * n;
            double beta = lb_dot (yj, d, n);
            beta /= ys_h[j];
            const double c = alpha[j] - beta;
            for (int e = 0; e < n; ++e) d[e] += c * sj[e];
            j = (j + 1) % m;
        }
        step = 1.0;
    }

done:
    if (ptr_fx) *ptr_fx = fx;
    free (pf);
    free (alpha);
    free (ys_h);
    free (Y);
    free (S);
    free (d);
    free (gp);
    free (g);
    free (xp);
    return ret;
}
