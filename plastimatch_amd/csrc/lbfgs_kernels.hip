// Vector kernels of the device-resident L-BFGS: the vector operations of
// liblbfgs's lbfgs() (libs/liblbfgs-1.9/lib/lbfgs.c:449-609) and of its
// More-Thuente line search (:903-906), in double, fused so that each
// kernel does one pass over its vectors.
//
// Built with -ffp-contract=off: each element update rounds its product and
// its sum separately, as liblbfgs's vecadd / vecdiff / vecscale do.

#include "lbfgs_kernels.h"

namespace plmhip {

namespace {

// Sum of nb partials in the fixed lane-then-butterfly order; every lane of
// the wave returns the same value.
__device__ __forceinline__ double
wave_reduce_parts (const double* __restrict__ p, int nb)
{
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int k = lane; k < nb; k += 64) s += p[k];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor (s, o, 64);
    return s;
}

// Workgroup sums of Q per-thread values into part[q * gridDim.x + block]:
// butterfly within each wave, then the four wave sums as (w0 + w1) +
// (w2 + w3).
template <int Q>
__device__ __forceinline__ void
block_partials (double (&v)[Q], double* __restrict__ part)
{
    __shared__ double red[Q][LB_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double s = v[q];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor (s, o, 64);
        if (lane == 0) red[q][w] = s;
    }
    __syncthreads ();
    if (threadIdx.x < Q) {
        const int q = threadIdx.x;
        part[q * gridDim.x + blockIdx.x] = (red[q][0] + red[q][1]) + (red[q][2] + red[q][3]);
    }
}

__global__ void __launch_bounds__ (LB_THREADS)
k_lb_trial (const double* __restrict__ xp, const double* __restrict__ d, double stp,
    double* __restrict__ x, float* __restrict__ coeff, int64_t n)
{
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        double v = xp[e];
        if (d) v = v + stp * d[e];
        x[e] = v;
        coeff[e] = (float) v;
    }
}

__global__ void __launch_bounds__ (LB_THREADS)
k_lb_neg (const double* __restrict__ g, double* __restrict__ d, int64_t n)
{
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride)
        d[e] = -g[e];
}

__global__ void __launch_bounds__ (LB_THREADS)
k_lb_dots (const double* __restrict__ g, const double* __restrict__ d,
    const double* __restrict__ x, double* __restrict__ part, int64_t n)
{
    double v[3] = {0.0, 0.0, 0.0};
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        const double ge = g[e], de = d[e], xe = x[e];
        v[0] += ge * de;
        v[1] += xe * xe;
        v[2] += ge * ge;
    }
    block_partials<3> (v, part);
}

__global__ void __launch_bounds__ (LB_THREADS)
k_lb_pair (const double* __restrict__ x, const double* __restrict__ xp,
    const double* __restrict__ g, const double* __restrict__ gp, double* __restrict__ s,
    double* __restrict__ y, double* __restrict__ d, double* __restrict__ part, int64_t n)
{
    double v[3] = {0.0, 0.0, 0.0};
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        const double se = x[e] - xp[e];
        const double ge = g[e];
        const double ye = ge - gp[e];
        const double de = -ge;
        s[e] = se;
        y[e] = ye;
        d[e] = de;
        v[0] += ye * se;
        v[1] += ye * ye;
        v[2] += se * de;
    }
    block_partials<3> (v, part);
}

__global__ void __launch_bounds__ (LB_THREADS)
k_lb_loop1 (const double* __restrict__ part_in, int nb_in, int first, int last, int j, int jn,
    const double* __restrict__ yj, const double* __restrict__ next, double* __restrict__ d,
    LbScalars* __restrict__ sc, double* __restrict__ part_out, int64_t n)
{
    double ys_j, ys_n = 0.0, yy = 1.0, dot;
    if (first) {
        ys_n = wave_reduce_parts (part_in, nb_in);
        yy = wave_reduce_parts (part_in + nb_in, nb_in);
        dot = wave_reduce_parts (part_in + 2 * nb_in, nb_in);
        ys_j = ys_n;
    } else {
        dot = wave_reduce_parts (part_in, nb_in);
        ys_j = sc->ys[j];
        if (last) {
            ys_n = sc->ys[jn];
            yy = sc->yy;
        }
    }
    const double alpha = dot / ys_j;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (first) {
            sc->ys[jn] = ys_n;
            sc->yy = yy;
        }
        sc->alpha[j] = alpha;
    }
    const double c = -alpha;
    const double scale = ys_n / yy;
    double v[1] = {0.0};
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        double de = d[e] + c * yj[e];
        if (last) de = de * scale;
        d[e] = de;
        v[0] += next[e] * de;
    }
    block_partials<1> (v, part_out);
}

__global__ void __launch_bounds__ (LB_THREADS)
k_lb_loop2 (const double* __restrict__ part_in, int nb_in, int j, const double* __restrict__ sj,
    const double* __restrict__ next, double* __restrict__ d, const LbScalars* __restrict__ sc,
    double* __restrict__ part_out, int64_t n)
{
    const double dot = wave_reduce_parts (part_in, nb_in);
    const double beta = dot / sc->ys[j];
    const double c = sc->alpha[j] - beta;
    double v[1] = {0.0};
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        const double de = d[e] + c * sj[e];
        d[e] = de;
        v[0] += next[e] * de;
    }
    block_partials<1> (v, part_out);
}

} // namespace

double
lb_reduce_parts_host (const double* parts, int nb)
{
    double s[64];
    for (int l = 0; l < 64; ++l) {
        s[l] = 0.0;
        for (int k = l; k < nb; k += 64) s[l] += parts[k];
    }
    for (int o = 32; o >= 1; o >>= 1) {
        double t[64];
        for (int l = 0; l < 64; ++l) t[l] = s[l] + s[l ^ o];
        for (int l = 0; l < 64; ++l) s[l] = t[l];
    }
    return s[0];
}

hipError_t
lb_launch_trial (const double* xp, const double* d, double stp, double* x, float* coeff,
    int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_trial, dim3 (lb_blocks (n)), dim3 (LB_THREADS), 0, st, xp, d, stp, x,
        coeff, n);
    return hipGetLastError ();
}

hipError_t
lb_launch_neg (const double* g, double* d, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_neg, dim3 (lb_blocks (n)), dim3 (LB_THREADS), 0, st, g, d, n);
    return hipGetLastError ();
}

hipError_t
lb_launch_dots (const double* g, const double* d, const double* x, double* part, int64_t n,
    hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_dots, dim3 (lb_blocks (n)), dim3 (LB_THREADS), 0, st, g, d, x, part, n);
    return hipGetLastError ();
}

hipError_t
lb_launch_pair (const double* x, const double* xp, const double* g, const double* gp, double* s,
    double* y, double* d, double* part, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_pair, dim3 (lb_blocks (n)), dim3 (LB_THREADS), 0, st, x, xp, g, gp, s,
        y, d, part, n);
    return hipGetLastError ();
}

hipError_t
lb_launch_loop1 (const double* part_in, int first, int last, int j, int jn, const double* yj,
    const double* next, double* d, LbScalars* sc, double* part_out, int64_t n, hipStream_t st)
{
    const int nb = lb_blocks (n);
    hipLaunchKernelGGL (k_lb_loop1, dim3 (nb), dim3 (LB_THREADS), 0, st, part_in, nb, first, last,
        j, jn, yj, next, d, sc, part_out, n);
    return hipGetLastError ();
}

hipError_t
lb_launch_loop2 (const double* part_in, int j, const double* sj, const double* next, double* d,
    const LbScalars* sc, double* part_out, int64_t n, hipStream_t st)
{
    const int nb = lb_blocks (n);
    hipLaunchKernelGGL (k_lb_loop2, dim3 (nb), dim3 (LB_THREADS), 0, st, part_in, nb, j, sj, next,
        d, sc, part_out, n);
    return hipGetLastError ();
}

} // namespace plmhip
