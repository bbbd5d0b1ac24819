// Vector kernels of the device-resident L-BFGS (internal).
//
// Every vector of the optimizer (x, g, the direction d, the m (s, y)
// pairs) stays in device memory in double.  One evaluation is a chain of
// kernels on one stream: k_lb_next forms the point (the start point, a new
// direction d = -H g and its first trial, or a later trial xp + stp d) and
// its float coefficients; the score (and regularization) kernels evaluate
// them into a fixed gradient buffer; k_lb_post keeps that gradient in its
// role's buffer and forms, as per-workgroup partials, every dot product
// the optimizer needs next -- the line search's g.d, x.x, g.g and, should
// the trial be accepted, the new pair s = x - xp, y = g - gp with its
// products against the pairs kept in the history and against g; the
// one-workgroup k_lb_finish sums them in a fixed order and either hands the
// totals to the host or runs the optimizer's scalar logic (LbMachine,
// lbfgs_logic.h) itself on one thread, leaving the next chain's action in
// device memory.  Every kernel of the chain reads that action (LbAct), so
// the chain has no per-evaluation argument: it is captured once in a
// hipGraph and replayed, and once the run has ended every kernel of a
// replay returns at once.  The direction is the compact form of the L-BFGS
// matrix (Byrd, Nocedal and Schnabel, Math. Prog. 63, 1994, section 3),
// liblbfgs's two-loop recursion (libs/liblbfgs-1.9/lib/lbfgs.c:555-609)
// regrouped, with its small solve done by the scalar logic.  Every sum has
// a fixed order, so a run gives the same bits every time.
#ifndef PLM_LBFGS_KERNELS_H
#define PLM_LBFGS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plmhip {

constexpr int LB_THREADS = 256;
constexpr int LB_MAX_BLOCKS = 256;
constexpr int LB_MAX_M = 16;            // history pairs

// Partial components of k_lb_post with bp earlier pairs: 0 g.d, 1 x.x,
// 2 g.g; then for the new pair e: 3 s_e.y_e, 4 y_e.y_e, 5 s_e.g, 6 y_e.g;
// then for earlier pair i = 1..bp (slot e - i): s_e.y_j, s_j.y_e, y_e.y_j,
// s_j.g, y_j.g.
__host__ __device__ inline int
lb_post_components (int bp)
{
    return 7 + 5 * bp;
}

// Workgroups of every L-BFGS vector kernel over n elements.
inline int
lb_blocks (int64_t n)
{
    const int64_t b = (n + LB_THREADS - 1) / LB_THREADS;
    return (int) (b < 1 ? 1 : b > LB_MAX_BLOCKS ? LB_MAX_BLOCKS : b);
}

struct LbAct;
struct LbMachine;

// The run's device buffers.
struct LbBufs {
    double* X[2];           // x and xp, by the roles LbAct::xi / xpi
    double* G[2];           // g and gp, by LbAct::gi / gpi
    double* geval;          // the gradient the evaluation writes
    float* coeff;           // the coefficients it reads, (float) x
    double* d;              // search direction
    double* S;              // history, slot j at S + j n
    double* Y;
    double* slope;          // k_lb_next's partials of g.d (a new direction)
    double* part;           // k_lb_post's partials, component c at part + c nb
    double* score;          // score, regularization score, overlap count (int64 bits)
    const double* x0;       // the start point
    int64_t n;
    int nb;                 // lb_blocks (n)
    int m;                  // history pairs of the run
};

// What k_lb_finish tells a host that polls a device-resident run (mapped
// host memory): evaluations finished, reports made, and whether it ended.
struct LbFlag {
    int evals;
    int done;
    int nrep;
    int pad;
};

// The chain's kernels read the action from act in device memory (a
// device-resident run) or, with act null, from av passed by value (a
// host-driven run, which hands each chain its action).
// The point of the next evaluation, its coefficients, and for a new
// direction d itself and the partials of g.d.
hipError_t lb_launch_next (const LbAct* act, const LbAct& av, const LbBufs& b, hipStream_t st);
// The post-evaluation pass (instantiated for the run's m).
hipError_t lb_launch_post (const LbAct* act, const LbAct& av, const LbBufs& b, hipStream_t st);
// The MSE gradient reduce of a single-pair evaluation and the post pass in
// one launch (m <= 6, one element a thread: n <= nb * 256; out0.score and
// out0.nvox set): the gradient goes straight into the trial's g buffer.
struct PairDev;
struct EvalOut;
hipError_t lb_launch_reduce_post (const PairDev* pairs, const double* part_score, const int* part_count,
    int nparts, int tile_scores, const EvalOut& out0, const LbAct* act, const LbAct& av, const LbBufs& b,
    hipStream_t st);
// The totals: with machine null, written to host (mapped host memory:
// score, reg score, count, slope, then the components); else the machine
// advances on the device (act must be &machine->a) and flag is updated.
hipError_t lb_launch_finish (const LbAct* act, const LbAct& av, LbMachine* machine, const LbBufs& b,
    double* host, LbFlag* flag, hipStream_t st);
// coeff = (float) x.
hipError_t lb_launch_convert (const double* x, float* coeff, int64_t n, hipStream_t st);

} // namespace plmhip

#endif
