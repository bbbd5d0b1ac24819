// Vector kernels of the device-resident L-BFGS (internal).
//
// Every vector of the optimizer (x, g, the direction d, the m (s, y)
// pairs) stays in device memory in double; the host runs only the scalar
// logic of the More-Thuente line search.  Per evaluation one pass
// (k_lb_post) forms every dot product the optimizer needs next: the line
// search's g.d, x.x, g.g, and -- should the trial be accepted -- the new
// pair s = x - xp, y = g - gp, its products with the pairs kept in the
// history and with g.  Per iteration one pass (k_lb_combine) turns them
// into the search direction by the compact form of the L-BFGS matrix
// (Byrd, Nocedal and Schnabel, Math. Prog. 63, 1994, section 3), which is
// the two-loop recursion of liblbfgs (libs/liblbfgs-1.9/lib/lbfgs.c:
// 555-609) regrouped: d = -H g with H0 = (ys/yy) I of the newest pair.
// Dot products are reduced in two fixed-order levels -- per workgroup into
// partials, then the partials in lane-then-butterfly order -- so every run
// gives the same bits.
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

// Inner products of the history by slot: SY[i][j] = s_i.y_j, YY[i][j] =
// y_i.y_j, kept on the device from one iteration to the next.
struct LbGram {
    double SY[LB_MAX_M][LB_MAX_M];
    double YY[LB_MAX_M][LB_MAX_M];
};

// The host's reduction of nb partials, in the order the kernels use:
// lane l sums parts l, l + 64, ... in turn, then a butterfly over the 64
// lane sums (s_l + s_{l^o}, o = 32 .. 1).
double lb_reduce_parts_host (const double* parts, int nb);

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
    const double* score;    // objective parts written by the evaluation:
    double* host;           // copied to host[0..2]; line-search partials at host + 3
    double* part;           // pair partials: component c >= 3 at part + (c - 3) nb
};

// x = xp + stp * d (x = xp when d is null: the start point); coeff =
// (float) x.
hipError_t lb_launch_trial (const double* xp, const double* d, double stp, double* x,
    float* coeff, int64_t n, hipStream_t st);
// The post-evaluation pass.
hipError_t lb_launch_post (const LbPostArgs& a, int64_t n, hipStream_t st);

// What k_lb_combine reads and writes.
struct LbCombineArgs {
    const double* part;     // k_lb_post's pair partials (b > 0)
    int b;                  // pairs in the direction: 0 (d = -g) .. m
    int e;                  // newest slot
    int m;
    const double* S;
    const double* Y;
    LbGram* gram;
    const double* g;        // gradient at the current point
    const double* xp;       // the current point
    double stp;             // first trial step
    double* d;              // out: direction
    double* x;              // out: xp + stp d
    float* coeff;           // out: (float) x
    double* slope;          // out: partials of g.d (host memory)
};

hipError_t lb_launch_combine (const LbCombineArgs& a, int64_t n, hipStream_t st);

} // namespace plmhip

#endif
