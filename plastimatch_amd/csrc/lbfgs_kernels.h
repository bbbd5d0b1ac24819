// Vector kernels of the device-resident L-BFGS (internal).
//
// Every vector of the optimizer (x, g, the direction d, the m (s, y)
// pairs) stays in device memory in double; the host runs only the scalar
// logic of the More-Thuente line search.  Dot products are reduced in two
// fixed-order levels -- per workgroup into partials, then the partials by
// lb_reduce_parts's lane-then-butterfly order -- so every run gives the
// same bits.  A kernel that needs the dot of the kernel before it reduces
// that kernel's partials itself (every workgroup redundantly, in the same
// order), so the two-loop recursion runs as a chain of launches with no
// host round trip.
#ifndef PLM_LBFGS_KERNELS_H
#define PLM_LBFGS_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plmhip {

constexpr int LB_THREADS = 256;
constexpr int LB_MAX_BLOCKS = 256;
constexpr int LB_MAX_M = 32;            // history pairs

// Workgroups of every L-BFGS vector kernel over n elements.
inline int
lb_blocks (int64_t n)
{
    const int64_t b = (n + LB_THREADS - 1) / LB_THREADS;
    return (int) (b < 1 ? 1 : b > LB_MAX_BLOCKS ? LB_MAX_BLOCKS : b);
}

// Device scalars of the two-loop recursion: ys of every history slot, the
// alphas of the current direction, yy of the newest pair.
struct LbScalars {
    double ys[LB_MAX_M];
    double alpha[LB_MAX_M];
    double yy;
};

// The host's reduction of nb partials, in the order the kernels use:
// lane l sums parts l, l + 64, ... in turn, then a butterfly over the 64
// lane sums (s_l + s_{l^o}, o = 32 .. 1).
double lb_reduce_parts_host (const double* parts, int nb);

// x = xp + stp * d (x = xp when d is null); coeff = (float) x.
hipError_t lb_launch_trial (const double* xp, const double* d, double stp, double* x,
    float* coeff, int64_t n, hipStream_t st);
// d = -g.
hipError_t lb_launch_neg (const double* g, double* d, int64_t n, hipStream_t st);
// Partials of g.d, x.x, g.g: part[q * nb + block], q = 0, 1, 2.
hipError_t lb_launch_dots (const double* g, const double* d, const double* x, double* part,
    int64_t n, hipStream_t st);
// Newest pair and steepest direction: s = x - xp, y = g - gp, d = -g;
// partials of y.s, y.y, s.d.
hipError_t lb_launch_pair (const double* x, const double* xp, const double* g,
    const double* gp, double* s, double* y, double* d, double* part, int64_t n, hipStream_t st);
// First loop, one history slot j (newest first).  first: part_in holds
// the pair kernel's partials (ys, yy and s_j.d of the newest slot jn),
// else s_j.d alone.  alpha_j = s_j.d / ys_j; d += -alpha_j y_j; on the last
// step also d *= ys_jn / yy.  Writes partials of next.d (next: s of the
// following slot, or y_j on the last step, which the second loop starts
// with).
hipError_t lb_launch_loop1 (const double* part_in, int first, int last, int j, int jn,
    const double* yj, const double* next, double* d, LbScalars* sc, double* part_out,
    int64_t n, hipStream_t st);
// Second loop, one history slot j (oldest first): beta = y_j.d / ys_j;
// d += (alpha_j - beta) s_j; partials of next.d (y of the following slot,
// or g on the last step: the line search's initial slope).
hipError_t lb_launch_loop2 (const double* part_in, int j, const double* sj, const double* next,
    double* d, const LbScalars* sc, double* part_out, int64_t n, hipStream_t st);

} // namespace plmhip

#endif
