// Vector kernels of the device-resident L-BFGS (internal).
//
// Every vector of the optimizer (x, g, the direction d, the m (s, y)
// pairs) stays in device memory in double; the host runs only scalar
// work: the More-Thuente line search and the small dense algebra of the
// direction.  Per evaluation one pass (k_lb_post) forms every dot product
// the optimizer needs next -- the line search's g.d, x.x, g.g and, should
// the trial be accepted, the new pair s = x - xp, y = g - gp with its
// products against the pairs kept in the history and against g -- as
// per-workgroup partials; a one-workgroup pass (k_lb_finish) sums them and
// writes the totals into mapped host memory.  The host keeps the history's
// Gram entries and solves the compact
// form of the L-BFGS matrix (Byrd, Nocedal and Schnabel, Math. Prog. 63,
// 1994, section 3), which is liblbfgs's two-loop recursion
// (libs/liblbfgs-1.9/lib/lbfgs.c:555-609) regrouped; per iteration one pass
// (k_lb_combine) then forms d = -H g, the first trial point and the
// partials of the initial slope.  Every sum has a fixed order, so a run
// gives the same bits every time.
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
inline int
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

// What k_lb_post reads and writes.
struct LbPostArgs {
    const double* x;        // the trial point
    const double* xp;       // the current point
    const double* g;        // gradient at x
    const double* gp;       // gradient at xp
    const double* d;        // search direction (null: g, for the start point)
    double* S;              // history, slot j at S + j n
    double* Y;
    int e;                  // slot of the new pair; -1: no pair (start point)
    int bp;                 // earlier pairs kept with it, 0 .. m - 1
    int m;
    double* part;           // component c's partials at part + c nb
};

// x = xp + stp * d (x = xp when d is null: the start point); coeff =
// (float) x.
hipError_t lb_launch_trial (const double* xp, const double* d, double stp, double* x,
    float* coeff, int64_t n, hipStream_t st);
// The post-evaluation pass.
hipError_t lb_launch_post (const LbPostArgs& a, int64_t n, hipStream_t st);
// Totals for the host: host[0..2] = score[0..2]; host[3] = the slope
// (nb partials at slope_part); host[4 + c] = component c (nb partials at
// part + c nb), c < ncomp.  One workgroup, fixed order.
hipError_t lb_launch_finish (const double* part, int ncomp, int nb, const double* slope_part,
    const double* score, double* host, hipStream_t st);

// What k_lb_combine reads and writes: d = -gamma g - sum_i ca_i s_i +
// gamma sum_i cu_i y_i over the b pairs in slots slot[0..b) (oldest
// first); b = 0: d = -g.
struct LbCombineArgs {
    int b;
    int slot[LB_MAX_M];
    double ca[LB_MAX_M];
    double cu[LB_MAX_M];
    double gamma;
    const double* S;
    const double* Y;
    const double* g;        // gradient at the current point
    const double* xp;       // the current point
    double stp;             // first trial step
    double* d;              // out: direction
    double* x;              // out: xp + stp d
    float* coeff;           // out: (float) x
    double* slope;          // out: partials of g.d
};

hipError_t lb_launch_combine (const LbCombineArgs& a, int64_t n, hipStream_t st);

} // namespace plmhip

#endif
