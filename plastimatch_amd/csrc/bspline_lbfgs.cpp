// The optimizer loop on the device: plm_hip_bspline_optimize_lbfgs and
// plm_hip_bspline_optimize_lbfgs_resident.
//
// The algorithm is the one liblbfgs 1.9 runs for the reference
// (libs/liblbfgs-1.9/lib/lbfgs.c:256-653, 823-1330; driven by
// src/register/bspline_optimize_liblbfgs.cxx:83-109): L-BFGS two-loop
// recursion over the last m (s, y) pairs with H0 = (ys/yy) I, and the
// More-Thuente line search with every evaluation reported to the progress
// rule.  Its vectors live in device memory (lbfgs_kernels.hip); its scalar
// logic is the state machine of lbfgs_logic.h, advanced once per
// evaluation.  plm_hip_bspline_optimize_lbfgs runs the machine here, on
// the host: after each evaluation chain the host reads a few dozen totals
// from mapped memory, advances (calling the progress callback at every
// report), and passes the next action to the chain's kernels by value.
// plm_hip_bspline_optimize_lbfgs_resident runs it on the device
// (k_lb_finish): the chain, whose kernels then read the action from device
// memory, is captured once in a hipGraph and replayed back to back, a few
// replays ahead, while the host only watches for the end.
//
// The first trial of each line search is formed by the pass that forms
// the direction, before the initial slope g.d is known: the slope comes
// back with that trial's results.  Should the slope be positive (liblbfgs
// returns LBFGSERR_INCREASEGRADIENT before evaluating anything), the trial
// is dropped uncounted, so the run's record matches liblbfgs's; the
// coefficients of the last counted evaluation are re-formed from its x at
// the end.

#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "plm_hip_bspline.h"
#include "bspline_internal.h"
#include "bspline_kernels.h"
#include "lbfgs_kernels.h"
#include "lbfgs_logic.h"

using plmhip::LbAct;
using plmhip::LbMachine;

static_assert (sizeof (plmhip::LbReport) == sizeof (plm_lbfgs_report), "report layout");

namespace plmhip {

// The device buffers of a run, kept by the context for the next run with
// the same size and history length.
struct LbWorkspace {
    int64_t n = 0;
    int m = 0;
    int cap = 0;                            // reports the log holds
    LbBufs b {};
    LbMachine* d_mach = nullptr;            // the machine (device-resident runs)
    plm_lbfgs_report* d_rep = nullptr;      // their report log
    double* hh = nullptr;                   // pinned totals (host-driven runs)
    double* hh_dev = nullptr;
    LbFlag* flag = nullptr;                 // pinned progress note (device-resident runs)
    LbFlag* flag_dev = nullptr;

    ~LbWorkspace ()
    {
        void* bufs[] = {b.X[0], b.X[1], b.G[0], b.G[1], b.geval, b.coeff, b.d, b.S, b.Y, b.slope, b.part,
            b.score, d_mach, d_rep};
        for (void* q : bufs) if (q) (void) hipFree (q);
        void* hbufs[] = {hh, flag};
        for (void* q : hbufs) if (q) (void) hipHostFree (q);
    }
};

void
lb_workspace_free (LbWorkspace* w)
{
    delete w;
}

} // namespace plmhip

namespace {

int
check_parms (const plm_lbfgs_parms& p)
{
    if (p.m < 1 || p.m > plmhip::LB_MAX_M)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "L-BFGS m must be 1..16");
    if (p.past < 0 || p.past > plmhip::LB_MAX_PAST)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "L-BFGS past must be 0..64");
    if (!(p.epsilon >= 0.) || !(p.delta >= 0.) || !(p.min_step >= 0.) || !(p.max_step >= p.min_step)
        || !(p.ftol >= 0.) || !(p.gtol >= 0.) || !(p.xtol >= 0.) || p.max_linesearch <= 0)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "invalid L-BFGS parameter (as lbfgs() rejects it)");
    return PLM_HIP_OK;
}

// A device-resident run replays a graph of RESIDENT_CHAINS evaluation
// chains (a replay's boundary costs the queue more than the kernel
// boundaries inside it), one replay ahead of the chains finished.
constexpr int RESIDENT_CHAINS = 4;

// One run: its stream, its objective, its workspace.
struct Run {
    plm_hip_ctx* c = nullptr;
    hipStream_t st = nullptr;
    int metric = 0, grad_mode = 0;
    const plm_reg_parms* reg = nullptr;
    plm_lbfgs_parms p {};
    plmhip::LbWorkspace* w = nullptr;       // owned by the context
    plmhip::LbBufs b {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int err = PLM_HIP_OK;   // first C-ABI error; the run stops on it

    ~Run ()
    {
        if (exec) (void) hipGraphExecDestroy (exec);
        if (graph) (void) hipGraphDestroy (graph);
    }

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

    template <typename T>
    int pinned (T** host, T** dev, size_t count, const char* what)
    {
        if (hip (hipHostMalloc ((void**) host, sizeof (T) * count, hipHostMallocMapped | hipHostMallocCoherent),
                what))
            return err;
        return hip (hipHostGetDevicePointer ((void**) dev, *host, 0), "hipHostGetDevicePointer");
    }

    int setup (plm_hip_ctx* ctx, int met, int gm, const plm_reg_parms* rg, const plm_lbfgs_parms* parms,
        const double* d_x0)
    {
        int device = 0;
        int64_t knots = 0;
        int rc = plmhip::ctx_info (ctx, &device, &st, &knots);
        if (rc) return err = rc;
        if (parms) p = *parms;
        else plm_hip_lbfgs_parms_init (&p);
        rc = check_parms (p);
        if (rc) return err = rc;
        if (hip (hipSetDevice (device), "hipSetDevice")) return err;
        c = ctx;
        metric = met;
        grad_mode = gm;
        reg = rg;
        const size_t n = (size_t) (3 * knots);
        plmhip::LbWorkspace** slot = plmhip::ctx_lb_workspace (ctx);
        if (*slot && ((*slot)->n != (int64_t) n || (*slot)->m != p.m)) {
            plmhip::lb_workspace_free (*slot);
            *slot = nullptr;
        }
        if (!*slot) {
            // (filled by the allocations below; kept by the context from here)
            w = *slot = new plmhip::LbWorkspace ();
            w->n = (int64_t) n;
            w->m = p.m;
            plmhip::LbBufs& wb = w->b;
            wb.n = (int64_t) n;
            wb.nb = plmhip::lb_blocks (wb.n);
            wb.m = p.m;
            const int ncomp = plmhip::lb_post_components (p.m - 1);
            if (alloc (&wb.X[0], n, "hipMalloc(L-BFGS x)") || alloc (&wb.X[1], n, "hipMalloc(L-BFGS x)")
                || alloc (&wb.G[0], n, "hipMalloc(L-BFGS g)") || alloc (&wb.G[1], n, "hipMalloc(L-BFGS g)")
                || alloc (&wb.geval, n, "hipMalloc(L-BFGS g)") || alloc (&wb.coeff, n, "hipMalloc(L-BFGS coeff)")
                || alloc (&wb.d, n, "hipMalloc(L-BFGS d)") || alloc (&wb.S, n * p.m, "hipMalloc(L-BFGS s)")
                || alloc (&wb.Y, n * p.m, "hipMalloc(L-BFGS y)")
                || alloc (&wb.score, 3, "hipMalloc(L-BFGS score)")
                || alloc (&wb.part, (size_t) ncomp * wb.nb, "hipMalloc(L-BFGS partials)")
                || alloc (&wb.slope, (size_t) wb.nb, "hipMalloc(L-BFGS slope partials)")
                || pinned (&w->hh, &w->hh_dev, 4 + (size_t) ncomp, "hipHostMalloc(L-BFGS readback)")
                || alloc (&w->d_mach, 1, "hipMalloc(L-BFGS state)")
                || pinned (&w->flag, &w->flag_dev, 1, "hipHostMalloc(L-BFGS note)")) {
                plmhip::lb_workspace_free (w);
                *slot = w = nullptr;
                return err;
            }
        }
        w = *slot;
        b = w->b;
        b.x0 = d_x0;
        // the regularizer adds into the gradient and leaves its score;
        // without it the reg score slot stays zero
        if (hip (hipMemsetAsync (b.score, 0, sizeof (double) * 3, st), "L-BFGS score clear")
            || hip (hipMemsetAsync (b.slope, 0, sizeof (double) * b.nb, st), "L-BFGS slope clear"))
            return err;
        if (reg && (rc = plmhip::ctx_reg_prepare (c))) return err = rc;
        return PLM_HIP_OK;
    }

    // The gradient reduce of an MSE evaluation with the post pass folded in
    // (lb_launch_reduce_post), launched by the evaluation in place of its
    // own reduce.
    struct Fused {
        const LbAct* act;
        const LbAct* av;
        const plmhip::LbBufs* b;
        bool launched;
    };
    static hipError_t
    reduce_post (void* user, const plmhip::Plan& plan, int nparts, int tile_scores, const plmhip::EvalOut& out0,
        hipStream_t s)
    {
        Fused* f = static_cast<Fused*> (user);
        f->launched = true;
        return plmhip::lb_launch_reduce_post (plan.d_pairs, plan.part_score, plan.part_count, nparts,
            tile_scores, out0, f->act, *f->av, *f->b, s);
    }

    // One evaluation chain reading its action from act (device memory),
    // or with act null from av.  An MSE objective without regularization
    // (m <= 6, n within one element a thread of the post grid) takes the
    // fused reduce and post pass: the same bits, one launch fewer.
    int chain (const LbAct* act, const LbAct& av, LbMachine* mach)
    {
        const int* skip = act ? &act->done : nullptr;
        if (hip (plmhip::lb_launch_next (act, av, b, st), "L-BFGS trial point")) return err;
        const bool fuse = metric == PLM_METRIC_MSE && !reg && grad_mode != PLM_GRAD_NONE && b.m <= 6
            && (int64_t) b.nb * plmhip::LB_THREADS >= b.n;
        Fused f {act, &av, &b, false};
        const plmhip::ReduceHook hook {&reduce_post, &f};
        int rc = plmhip::ctx_eval (c, metric, grad_mode, b.coeff, b.score, b.geval,
            reinterpret_cast<int64_t*> (b.score + 2), st, skip, fuse ? &hook : nullptr);
        if (rc < 0) return err = rc;
        if (reg && (rc = plmhip::ctx_regularize (c, reg, b.coeff, b.score + 1, b.geval, st, skip)) < 0)
            return err = rc;
        if ((!f.launched && hip (plmhip::lb_launch_post (act, av, b, st), "L-BFGS post pass"))
            || hip (plmhip::lb_launch_finish (act, av, mach, b, w->hh_dev, w->flag_dev, st), "L-BFGS totals"))
            return err;
        return PLM_HIP_OK;
    }

    // The machine on the host; the callback sees every report as it comes.
    int run_host (LbMachine& M, plm_lbfgs_progress_fn progress, void* inst)
    {
        M.progress = progress;
        M.inst = inst;
        M.cb_x[0] = b.X[0];
        M.cb_x[1] = b.X[1];
        M.cb_g[0] = b.G[0];
        M.cb_g[1] = b.G[1];
        while (M.phase != 2) {
            if (chain (nullptr, M.a, nullptr) || hip (hipStreamSynchronize (st), "L-BFGS evaluation"))
                return err;
            M.advance (w->hh);
        }
        return PLM_HIP_OK;
    }

    // The machine on the device; RESIDENT_CHAINS chains captured once and
    // replayed until the run ends.
    int run_device (LbMachine& M, plm_lbfgs_report* reports, int max_reports, int* n_reports)
    {
        const int cap = max_reports > 0 ? max_reports : 0;
        if (cap > w->cap) {
            if (w->d_rep) (void) hipFree (w->d_rep);
            w->d_rep = nullptr;
            w->cap = 0;
            if (alloc (&w->d_rep, (size_t) cap, "hipMalloc(L-BFGS log)")) return err;
            w->cap = cap;
        }
        M.rep = reinterpret_cast<plmhip::LbReport*> (w->d_rep);
        M.cap = cap;
        M.progress = nullptr;
        std::memset ((void*) w->flag, 0, sizeof (*w->flag));
        if (hip (hipMemcpyAsync (w->d_mach, &M, sizeof (LbMachine), hipMemcpyHostToDevice, st), "L-BFGS state")
            || hip (hipStreamSynchronize (st), "L-BFGS state"))
            return err;
        if (hip (hipStreamBeginCapture (st, hipStreamCaptureModeRelaxed), "L-BFGS capture")) return err;
        int rc = PLM_HIP_OK;
        for (int i = 0; i < RESIDENT_CHAINS && !rc; i++) rc = chain (&w->d_mach->a, M.a, w->d_mach);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture (st, &g);
        graph = g;
        if (rc) return rc;
        if (hip (ec, "L-BFGS capture") || hip (hipGraphInstantiate (&exec, graph, nullptr, nullptr, 0),
                "L-BFGS graph"))
            return err;
        const volatile plmhip::LbFlag* f = w->flag;
        int launched = 0;
        while (!f->done) {
            if (launched - f->evals <= RESIDENT_CHAINS) {
                if (hip (hipGraphLaunch (exec, st), "L-BFGS replay")) return err;
                launched += RESIDENT_CHAINS;
                continue;
            }
            // every chain launched so far may have run without ending the
            // run: more are due once the stream has drained
            const hipError_t q = hipStreamQuery (st);
            if (q != hipSuccess && q != hipErrorNotReady) return hip (q, "L-BFGS evaluation");
        }
        if (hip (hipStreamSynchronize (st), "L-BFGS end")
            || hip (hipMemcpy (&M, w->d_mach, sizeof (LbMachine), hipMemcpyDeviceToHost), "L-BFGS state"))
            return err;
        const int nr = M.nrep < cap ? M.nrep : cap;
        if (reports && nr > 0
            && hip (hipMemcpy (reports, w->d_rep, sizeof (plm_lbfgs_report) * nr, hipMemcpyDeviceToHost),
                   "L-BFGS log"))
            return err;
        if (n_reports) *n_reports = M.nrep;
        return PLM_HIP_OK;
    }

    // x and the last counted evaluation's coefficients out.
    int finish (const LbMachine& M, double* d_x, float* d_coeff)
    {
        if (hip (hipMemcpyAsync (d_x, b.X[M.a.xi], sizeof (double) * b.n, hipMemcpyDeviceToDevice, st),
                "L-BFGS x out"))
            return err;
        if (d_coeff && M.c_buf >= 0
            && hip (plmhip::lb_launch_convert (b.X[M.c_buf], d_coeff, b.n, st), "L-BFGS coefficients out"))
            return err;
        return hip (hipStreamSynchronize (st), "L-BFGS end");
    }
};

int
optimize (plm_hip_ctx* c, int metric, int grad_mode, const plm_reg_parms* reg, const plm_lbfgs_parms* parms,
    double* d_x, float* d_coeff, plm_lbfgs_progress_fn progress, void* instance, bool resident,
    plm_lbfgs_report* reports, int max_reports, int* n_reports, plm_lbfgs_result* result)
{
    if (!c || !d_x || !result) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "NULL argument");
    if (metric == PLM_METRIC_MSE && grad_mode == PLM_GRAD_NONE)
        return plmhip::abi_fail (PLM_HIP_ERR_ARG, "the optimizer needs a gradient");
    if (max_reports < 0) return plmhip::abi_fail (PLM_HIP_ERR_ARG, "max_reports must be >= 0");
    Run r;
    if (r.setup (c, metric, grad_mode, reg, parms, d_x)) return r.err;
    LbMachine M;
    M.init (r.p, reg ? 1 : 0, (int) r.b.n);
    M.rep = nullptr;
    M.cap = 0;
    if (resident ? r.run_device (M, reports, max_reports, n_reports) : r.run_host (M, progress, instance))
        return r.err;
    if (r.finish (M, d_x, d_coeff)) return r.err;
    result->status = M.ret;
    result->iterations = M.it;
    result->fevals = M.feval;
    result->fx = M.fx;
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
    // (the host-driven loop: its per-evaluation round trip costs less than
    // the scalar logic on one device thread)
    return optimize (c, metric, grad_mode, reg, parms, d_x, d_coeff, progress, instance, false, nullptr, 0,
        nullptr, result);
}

int
plm_hip_bspline_optimize_lbfgs_resident (plm_hip_ctx* c, int metric, int grad_mode,
    const plm_reg_parms* reg, const plm_lbfgs_parms* parms, double* d_x, float* d_coeff,
    plm_lbfgs_report* reports, int max_reports, int* n_reports, plm_lbfgs_result* result)
{
    return optimize (c, metric, grad_mode, reg, parms, d_x, d_coeff, nullptr, nullptr, true, reports,
        max_reports, n_reports, result);
}

} // extern "C"
