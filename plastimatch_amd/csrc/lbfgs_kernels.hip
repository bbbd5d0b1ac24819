// Vector kernels of the device-resident L-BFGS: the vector work of
// liblbfgs's lbfgs() (libs/liblbfgs-1.9/lib/lbfgs.c:449-609) and of its
// More-Thuente trials (:903-906), in double: one pass after each
// evaluation, one per iteration to form the direction.
//
// Built with -ffp-contract=off: each element update rounds its product and
// its sum separately, as liblbfgs's vecadd / vecdiff do.

#include "lbfgs_kernels.h"

namespace plmhip {

namespace {

// Workgroup sum of one per-thread value: a butterfly within each wave,
// then the four wave sums as (w0 + w1) + (w2 + w3); valid in thread 0.
__device__ __forceinline__ double
block_sum1 (double v)
{
    __shared__ double red[LB_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor (v, o, 64);
    if (lane == 0) red[w] = v;
    __syncthreads ();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Workgroup sums of NC per-thread values into out[c * nb + block]:
// components go through LDS sixteen at a time -- thread (c, l) adds
// threads l, l + 16, ..., l + 240 of component c in turn, then thread c
// adds its sixteen partial sums in turn.  Two LDS round trips per chunk
// instead of a shuffle butterfly per component.
template <int NC>
__device__ __forceinline__ void
block_partials (const double (&v)[NC], double* __restrict__ out, int nb)
{
    __shared__ double buf[16][LB_THREADS];
    __shared__ double mid[16][16];
    const int t = threadIdx.x;
#pragma unroll
    for (int c0 = 0; c0 < NC; c0 += 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (c0 + c < NC) buf[c][t] = v[c0 + c];
        __syncthreads ();
        {
            const int c = t >> 4, l = t & 15;
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < LB_THREADS / 16; ++q) s += buf[c][l + 16 * q];
            mid[c][l] = s;
        }
        __syncthreads ();
        if (t < 16 && c0 + t < NC) {
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < 16; ++l) s += mid[t][l];
            out[(c0 + t) * nb + blockIdx.x] = s;
        }
        __syncthreads ();
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

// BP < 0: the start point (line-search dots only); else the new pair into
// slot a.e with BP earlier pairs.
template <int BP>
__global__ void __launch_bounds__ (LB_THREADS)
k_lb_post (LbPostArgs a, int64_t n)
{
    constexpr int NC = BP < 0 ? 3 : 7 + 5 * BP;
    double v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = 0.0;
    int slot[BP > 0 ? BP : 1];
#pragma unroll
    for (int i = 1; i <= BP; ++i) slot[i - 1] = (a.e - i + a.m) % a.m;
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        const double xe = a.x[e], ge = a.g[e];
        const double de = a.d ? a.d[e] : ge;
        v[0] += ge * de;
        v[1] += xe * xe;
        v[2] += ge * ge;
        if constexpr (BP >= 0) {
            const double se = xe - a.xp[e];
            const double ye = ge - a.gp[e];
            a.S[(int64_t) a.e * n + e] = se;
            a.Y[(int64_t) a.e * n + e] = ye;
            v[3] += se * ye;
            v[4] += ye * ye;
            v[5] += se * ge;
            v[6] += ye * ge;
#pragma unroll
            for (int i = 0; i < BP; ++i) {
                const double sj = a.S[(int64_t) slot[i] * n + e];
                const double yj = a.Y[(int64_t) slot[i] * n + e];
                v[7 + 5 * i + 0] += se * yj;
                v[7 + 5 * i + 1] += sj * ye;
                v[7 + 5 * i + 2] += ye * yj;
                v[7 + 5 * i + 3] += sj * ge;
                v[7 + 5 * i + 4] += yj * ge;
            }
        }
    }
    block_partials<NC> (v, a.part, gridDim.x);
}

// One workgroup: the slope's total into host[3], component c's into
// host[4 + c]; each is the sum of its nb partials, thread (k, l) of a group
// of sixteen totals adding partials l, l + 16, ... in turn, then the
// sixteen in turn.
__global__ void __launch_bounds__ (LB_THREADS)
k_lb_finish (const double* __restrict__ part, int ncomp, int nb, const double* __restrict__ slope_part,
    const double* __restrict__ score, double* __restrict__ host)
{
    __shared__ double mid[16][16];
    const int t = threadIdx.x;
    const int total = ncomp + 1;
    for (int k0 = 0; k0 < total; k0 += 16) {
        const int k = k0 + (t >> 4), l = t & 15;
        // every partial of this thread loaded before any is added (nb <=
        // LB_MAX_BLOCKS), then added in q order; the absent ones add +0.0,
        // which leaves a sum that started at +0.0 unchanged
        const double* p = k == 0 ? slope_part : part + (int64_t) (k - 1) * nb;
        double pv[LB_MAX_BLOCKS / 16];
#pragma unroll
        for (int i = 0; i < LB_MAX_BLOCKS / 16; i++) {
            const int q = l + 16 * i;
            pv[i] = (k < total && q < nb) ? p[q] : 0.0;
        }
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < LB_MAX_BLOCKS / 16; i++) s += pv[i];
        mid[t >> 4][l] = s;
        __syncthreads ();
        if (t < 16 && k0 + t < total) {
            double u = 0.0;
#pragma unroll
            for (int j = 0; j < 16; ++j) u += mid[t][j];
            host[3 + k0 + t] = u;
        }
        __syncthreads ();
    }
    if (t == 0) {
        host[0] = score[0];
        host[1] = score[1];
        host[2] = score[2];
    }
}

// The direction d = -H g from the coefficients the host solved for, the
// first trial point xp + stp d, its coefficients, and the partials of g.d.
__global__ void __launch_bounds__ (LB_THREADS)
k_lb_combine (LbCombineArgs a, int64_t n)
{
    double v = 0.0;
    const int b = a.b;
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        const double ge = a.g[e];
        double de;
        if (b == 0) {
            de = -ge;
        } else {
            double t1 = 0.0, t2 = 0.0;
            for (int i = 0; i < b; ++i) {
                const int64_t off = (int64_t) a.slot[i] * n + e;
                t1 = t1 + a.ca[i] * a.S[off];
                t2 = t2 + a.cu[i] * a.Y[off];
            }
            de = -a.gamma * ge - t1 + a.gamma * t2;
        }
        a.d[e] = de;
        const double xe = a.xp[e] + a.stp * de;
        a.x[e] = xe;
        a.coeff[e] = (float) xe;
        v += ge * de;
    }
    const double s = block_sum1 (v);
    if (threadIdx.x == 0) a.slope[blockIdx.x] = s;
}

template <int BP>
hipError_t
launch_post_bp (const LbPostArgs& a, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_post<BP>, dim3 (lb_blocks (n)), dim3 (LB_THREADS), 0, st, a, n);
    return hipGetLastError ();
}

} // namespace

hipError_t
lb_launch_finish (const double* part, int ncomp, int nb, const double* slope_part, const double* score,
    double* host, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_finish, dim3 (1), dim3 (LB_THREADS), 0, st, part, ncomp, nb, slope_part, score,
        host);
    return hipGetLastError ();
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
lb_launch_post (const LbPostArgs& a, int64_t n, hipStream_t st)
{
    if (a.e < 0) return launch_post_bp<-1> (a, n, st);
    switch (a.bp) {
    case 0: return launch_post_bp<0> (a, n, st);
    case 1: return launch_post_bp<1> (a, n, st);
    case 2: return launch_post_bp<2> (a, n, st);
    case 3: return launch_post_bp<3> (a, n, st);
    case 4: return launch_post_bp<4> (a, n, st);
    case 5: return launch_post_bp<5> (a, n, st);
    case 6: return launch_post_bp<6> (a, n, st);
    case 7: return launch_post_bp<7> (a, n, st);
    case 8: return launch_post_bp<8> (a, n, st);
    case 9: return launch_post_bp<9> (a, n, st);
    case 10: return launch_post_bp<10> (a, n, st);
    case 11: return launch_post_bp<11> (a, n, st);
    case 12: return launch_post_bp<12> (a, n, st);
    case 13: return launch_post_bp<13> (a, n, st);
    case 14: return launch_post_bp<14> (a, n, st);
    case 15: return launch_post_bp<15> (a, n, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t
lb_launch_combine (const LbCombineArgs& a, int64_t n, hipStream_t st)
{
    if (a.b < 0 || a.b > LB_MAX_M) return hipErrorInvalidValue;
    hipLaunchKernelGGL (k_lb_combine, dim3 (lb_blocks (n)), dim3 (LB_THREADS), 0, st, a, n);
    return hipGetLastError ();
}

} // namespace plmhip
