// Vector kernels of the device-resident L-BFGS: the vector work of
// liblbfgs's lbfgs() (libs/liblbfgs-1.9/lib/lbfgs.c:449-609) and of its
// More-Thuente trials (:903-906), in double, as two passes per iteration.
//
// Built with -ffp-contract=off: each element update rounds its product and
// its sum separately, as liblbfgs's vecadd / vecdiff do.

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

// Workgroup sums of NC per-thread values: a butterfly within each wave,
// then the four wave sums as (w0 + w1) + (w2 + w3).  Returns them in
// out[c] for threads c < NC (undefined elsewhere).
template <int NC>
__device__ __forceinline__ double
block_sums (const double (&v)[NC])
{
    __shared__ double red[NC][LB_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        double s = v[c];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor (s, o, 64);
        if (lane == 0) red[c][w] = s;
    }
    __syncthreads ();
    const int c = threadIdx.x < NC ? threadIdx.x : 0;
    return (red[c][0] + red[c][1]) + (red[c][2] + red[c][3]);
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
    const double s = block_sums<NC> (v);
    const int nb = gridDim.x;
    const int c = threadIdx.x;
    if (c < 3) a.host[3 + c * nb + blockIdx.x] = s;
    else if (c < NC) a.part[(c - 3) * nb + blockIdx.x] = s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.host[0] = a.score[0];
        a.host[1] = a.score[1];
        a.host[2] = a.score[2];
    }
}

// The direction d = -H g in the compact form, the first trial point
// xp + stp d, its coefficients, and the partials of g.d.
//
// With the b kept pairs in age order (position i = 0 oldest .. b - 1 the
// newest, slot e), R the upper triangle of S'Y, D its diagonal,
// gamma = s_e.y_e / y_e.y_e, p = S'g and q = Y'g:
//   u = R^-1 p,   a = R^-T ((D + gamma Y'Y) u - gamma q),
//   H g = gamma g + S a - gamma Y u.
// The products of the newest pair come from k_lb_post's partials, those
// among the earlier pairs from the Gram record; workgroup 0 stores the
// newest pair's row and column there for the next iterations (no other
// workgroup reads them from it).
__global__ void __launch_bounds__ (LB_THREADS)
k_lb_combine (LbCombineArgs a, int64_t n)
{
    __shared__ double red[4 + 5 * (LB_MAX_M - 1)];
    __shared__ double SYm[LB_MAX_M][LB_MAX_M], YYm[LB_MAX_M][LB_MAX_M];
    __shared__ double P[LB_MAX_M], Q[LB_MAX_M];
    __shared__ double ca[LB_MAX_M], cu[LB_MAX_M];
    __shared__ double s_gamma;
    __shared__ int s_slot[LB_MAX_M];
    const int nb = gridDim.x;
    const int b = a.b;
    if (b > 0) {
        const int m = a.m, e = a.e;
        const int nc = 4 + 5 * (b - 1);
        for (int c = threadIdx.x >> 6; c < nc; c += LB_THREADS / 64) {
            const double s = wave_reduce_parts (a.part + (int64_t) c * nb, nb);
            if ((threadIdx.x & 63) == 0) red[c] = s;
        }
        if (threadIdx.x < b) s_slot[threadIdx.x] = (e - (b - 1) + (int) threadIdx.x + m) % m;
        __syncthreads ();
        // products by age position (0 oldest .. b - 1 newest): those with
        // the newest pair from the partials (position i is b - 1 - i pairs
        // older than it), the rest from the Gram record
        if (threadIdx.x < b * b) {
            const int i = threadIdx.x / b, j = threadIdx.x % b;
            double sy, yy;
            if (i == b - 1 && j == b - 1) {
                sy = red[0];
                yy = red[1];
            } else if (i == b - 1) {
                sy = red[4 + 5 * (b - 2 - j) + 0];
                yy = red[4 + 5 * (b - 2 - j) + 2];
            } else if (j == b - 1) {
                sy = red[4 + 5 * (b - 2 - i) + 1];
                yy = red[4 + 5 * (b - 2 - i) + 2];
            } else {
                sy = a.gram->SY[s_slot[i]][s_slot[j]];
                yy = a.gram->YY[s_slot[i]][s_slot[j]];
            }
            SYm[i][j] = sy;
            YYm[i][j] = yy;
        }
        if (threadIdx.x < b) {
            const int i = threadIdx.x;
            P[i] = i == b - 1 ? red[2] : red[4 + 5 * (b - 2 - i) + 3];
            Q[i] = i == b - 1 ? red[3] : red[4 + 5 * (b - 2 - i) + 4];
        }
        __syncthreads ();
        if (threadIdx.x < 64) {
            // one wave: lane l holds component l of u, then of a; every
            // inner sum is a butterfly over the lanes (fixed order)
            const int l = threadIdx.x;
            const int lc = l < b ? l : 0;           // lanes past b read row / column 0
            const double gamma = SYm[b - 1][b - 1] / YYm[b - 1][b - 1];
            double ul = 0.0;
            for (int i = b - 1; i >= 0; --i) {
                double c = (l > i && l < b) ? SYm[i][lc] * ul : 0.0;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor (c, o, 64);
                const double ui = (P[i] - c) / SYm[i][i];
                if (l == i) ul = ui;
            }
            double w = 0.0;
            for (int j = 0; j < b; ++j) w = w + YYm[lc][j] * __shfl (ul, j, 64);
            const double tl = l < b ? SYm[lc][lc] * ul + gamma * w - gamma * Q[lc] : 0.0;
            double al = 0.0;
            for (int i = 0; i < b; ++i) {
                double c = l < i ? SYm[lc][i] * al : 0.0;
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor (c, o, 64);
                const double ti = __shfl (tl, i, 64);
                const double ai = (ti - c) / SYm[i][i];
                if (l == i) al = ai;
            }
            if (l < b) {
                ca[l] = al;
                cu[l] = ul;
            }
            if (l == 0) s_gamma = gamma;
        }
        if (blockIdx.x == 0) {
            // the newest pair's row and column into the Gram record
            if (threadIdx.x == 0) {
                a.gram->SY[e][e] = red[0];
                a.gram->YY[e][e] = red[1];
            } else if (threadIdx.x < b) {
                const int k = threadIdx.x;
                const int j = (e - k + m) % m;
                a.gram->SY[e][j] = red[4 + 5 * (k - 1) + 0];
                a.gram->SY[j][e] = red[4 + 5 * (k - 1) + 1];
                a.gram->YY[e][j] = red[4 + 5 * (k - 1) + 2];
                a.gram->YY[j][e] = red[4 + 5 * (k - 1) + 2];
            }
        }
        __syncthreads ();
    }
    double v[1] = {0.0};
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        const double ge = a.g[e];
        double de;
        if (b == 0) {
            de = -ge;
        } else {
            double t1 = 0.0, t2 = 0.0;
            for (int i = 0; i < b; ++i) {
                const int64_t off = (int64_t) s_slot[i] * n + e;
                t1 = t1 + ca[i] * a.S[off];
                t2 = t2 + cu[i] * a.Y[off];
            }
            de = -s_gamma * ge - t1 + s_gamma * t2;
        }
        a.d[e] = de;
        const double xe = a.xp[e] + a.stp * de;
        a.x[e] = xe;
        a.coeff[e] = (float) xe;
        v[0] += ge * de;
    }
    const double s = block_sums<1> (v);
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
    if (a.b < 0 || a.b > LB_MAX_M || a.b > a.m) return hipErrorInvalidValue;
    hipLaunchKernelGGL (k_lb_combine, dim3 (lb_blocks (n)), dim3 (LB_THREADS), 0, st, a, n);
    return hipGetLastError ();
}

} // namespace plmhip
