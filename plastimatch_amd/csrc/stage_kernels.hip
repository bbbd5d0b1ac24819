// Stage preprocessing kernels: the float resample of
// volume_resample_float_li / _nn (src/base/volume_resample.cxx:26-135) over
// host-made axis tables, and the min/max scan of
// bspline_initialize_mi_hist_eqsp (src/register/joint_histogram.cxx:
// 115-139).  Built with -ffp-contract=off: the eight weighted corners are
// formed and summed exactly as LI_VALUE writes them
// (src/base/interpolate_macros.h:23-38), so the result is the reference's
// bit for bit.

#include <float.h>

#include "stage_kernels.h"

namespace plmhip {

namespace {

constexpr int RS_THREADS = 256;

// One workgroup per (x segment, row j, slice k) of the output.
template <bool NEAREST>
__global__ void __launch_bounds__ (RS_THREADS)
k_resample (const float* __restrict__ in, int id0, int id1, float* __restrict__ out, int od0,
    int od1, const ResampleAxis* __restrict__ ax, const ResampleAxis* __restrict__ ay,
    const ResampleAxis* __restrict__ az)
{
    const int i = blockIdx.x * RS_THREADS + threadIdx.x;
    if (i >= od0) return;
    const int j = blockIdx.y, k = blockIdx.z;
    const ResampleAxis X = ax[i], Y = ay[j], Z = az[k];
    const int64_t v = i + (int64_t) od0 * (j + (int64_t) od1 * k);
    if (!(X.inside & Y.inside & Z.inside)) {
        out[v] = 0.0f;
        return;
    }
    const int64_t d01 = (int64_t) id0 * id1;
    const int64_t f = X.idx + (int64_t) id0 * Y.idx + d01 * Z.idx;
    if (NEAREST) {
        out[v] = in[f];
        return;
    }
    const float a = X.w1 * Y.w1 * Z.w1 * in[f];
    const float b = X.w2 * Y.w1 * Z.w1 * in[f + 1];
    const float c = X.w1 * Y.w2 * Z.w1 * in[f + id0];
    const float e = X.w2 * Y.w2 * Z.w1 * in[f + id0 + 1];
    const float g = X.w1 * Y.w1 * Z.w2 * in[f + d01];
    const float h = X.w2 * Y.w1 * Z.w2 * in[f + d01 + 1];
    const float p = X.w1 * Y.w2 * Z.w2 * in[f + d01 + id0];
    const float q = X.w2 * Y.w2 * Z.w2 * in[f + d01 + id0 + 1];
    out[v] = a + b + c + e + g + h + p + q;
}

__global__ void __launch_bounds__ (256)
k_minmax (const float* __restrict__ img, int64_t n, float* __restrict__ part)
{
    __shared__ float smin[4], smax[4];
    float mn = img[0], mx = img[0];
    const int64_t stride = (int64_t) gridDim.x * 256;
    for (int64_t e = (int64_t) blockIdx.x * 256 + threadIdx.x; e < n; e += stride) {
        const float v = img[e];
        // a NaN fails both tests, as in the reference's scan
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const float a = __shfl_xor (mn, o, 64), b = __shfl_xor (mx, o, 64);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if ((threadIdx.x & 63) == 0) {
        smin[threadIdx.x >> 6] = mn;
        smax[threadIdx.x >> 6] = mx;
    }
    __syncthreads ();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            mn = smin[w] < mn ? smin[w] : mn;
            mx = smax[w] > mx ? smax[w] : mx;
        }
        part[2 * blockIdx.x] = mn;
        part[2 * blockIdx.x + 1] = mx;
    }
}

// Key counts and sums of the V-optimal histogram: per workgroup in LDS
// (integer counts, double sums), then added to the device totals.
__global__ void __launch_bounds__ (256)
k_vopt_keys (const float* __restrict__ img, int64_t n, float offset, float delta,
    unsigned long long* __restrict__ count, double* __restrict__ sum)
{
    __shared__ unsigned int c[VOPT_KEYS];
    __shared__ double s[VOPT_KEYS];
    for (int k = threadIdx.x; k < VOPT_KEYS; k += 256) {
        c[k] = 0u;
        s[k] = 0.0;
    }
    __syncthreads ();
    const int64_t stride = (int64_t) gridDim.x * 256;
    for (int64_t e = (int64_t) blockIdx.x * 256 + threadIdx.x; e < n; e += stride) {
        // (v - offset) and its quotient in float, as the reference bins it
        const float x = img[e] - offset;
        int k = (int) floorf (x / delta);
        k = k < 0 ? 0 : (k >= VOPT_KEYS ? VOPT_KEYS - 1 : k);
        atomicAdd (&c[k], 1u);
        atomicAdd (&s[k], (double) x);
    }
    __syncthreads ();
    for (int k = threadIdx.x; k < VOPT_KEYS; k += 256) {
        if (c[k]) {
            atomicAdd (&count[k], (unsigned long long) c[k]);
            atomicAdd (&sum[k], s[k]);
        }
    }
}

// vopt_bin_error (src/register/joint_histogram.cxx:62-82) on the prefix
// tables: the keys start + 1 .. end as one bin, DBL_MAX below one voxel.
__device__ __forceinline__ double
vopt_err (int start, int end, const double* s, const double* ssq, const double* cnt)
{
    const double sq = ssq[end] - ssq[start];
    const double d = s[end] - s[start];
    const double nk = (double) end - (double) start + 1.0;
    const double nv = cnt[end] - cnt[start];
    const double v = sq - (d * d) / nk;
    return nv < 1.0 ? DBL_MAX : v;
}

// Row j of the programme (bspline_initialize_mi_hist_vopt,
// src/register/joint_histogram.cxx:277-301): for every key i, the least
// prev[k] + vopt_err(k + 1, i) over k < i, a tie going to the later k, with
// DBL_MAX and split 0 where no candidate is <= DBL_MAX; row 0 (prev null)
// is vopt_err(0, i).  One wave per key: lane l takes k = l, l + 64, ... in
// increasing order with the sequential rule, then the lanes' (value, k)
// pairs are combined by the least value, equal values by the larger k --
// the pair the sequential loop ends with, since every candidate is formed
// by the same double operations.
__global__ void __launch_bounds__ (256)
k_vopt_row (const double* __restrict__ s, const double* __restrict__ ssq,
    const double* __restrict__ cnt, const double* __restrict__ prev, double* __restrict__ cur,
    int* __restrict__ split)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= VOPT_KEYS) return;
    if (!prev) {
        if (lane == 0) cur[i] = vopt_err (0, i, s, ssq, cnt);
        return;
    }
    double best = DBL_MAX;
    int arg = -1;
    for (int k = lane; k < i; k += 64) {
        const double c = prev[k] + vopt_err (k + 1, i, s, ssq, cnt);
        if (c <= best) {
            best = c;
            arg = k;
        }
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double ob = __shfl_xor (best, o);
        const int oa = __shfl_xor (arg, o);
        if (ob < best || (ob == best && oa > arg)) {
            best = ob;
            arg = oa;
        }
    }
    if (lane == 0) {
        cur[i] = best;
        split[i] = arg < 0 ? 0 : arg;
    }
}

} // namespace

hipError_t
launch_resample (const float* in, const int id[3], float* out, const int od[3],
    const ResampleAxis* ax, const ResampleAxis* ay, const ResampleAxis* az, int nearest,
    hipStream_t st)
{
    const dim3 grid ((od[0] + RS_THREADS - 1) / RS_THREADS, od[1], od[2]);
    if (nearest)
        hipLaunchKernelGGL (k_resample<true>, grid, dim3 (RS_THREADS), 0, st, in, id[0], id[1], out,
            od[0], od[1], ax, ay, az);
    else
        hipLaunchKernelGGL (k_resample<false>, grid, dim3 (RS_THREADS), 0, st, in, id[0], id[1], out,
            od[0], od[1], ax, ay, az);
    return hipGetLastError ();
}

hipError_t
launch_minmax (const float* img, int64_t n, float* part, int* blocks, hipStream_t st)
{
    int64_t b = (n + 255) / 256;
    if (b > MINMAX_BLOCKS) b = MINMAX_BLOCKS;
    if (b < 1) b = 1;
    *blocks = (int) b;
    hipLaunchKernelGGL (k_minmax, dim3 ((unsigned) b), dim3 (256), 0, st, img, n, part);
    return hipGetLastError ();
}

hipError_t
launch_vopt_keys (const float* img, int64_t n, float offset, float delta, unsigned long long* count,
    double* sum, hipStream_t st)
{
    int64_t b = (n + 255) / 256;
    if (b > 1024) b = 1024;
    if (b < 1) b = 1;
    hipLaunchKernelGGL (k_vopt_keys, dim3 ((unsigned) b), dim3 (256), 0, st, img, n, offset, delta,
        count, sum);
    return hipGetLastError ();
}

hipError_t
launch_vopt_dp (const double* s, const double* ssq, const double* cnt, int bins, double* err,
    int* split, hipStream_t st)
{
    const dim3 grid ((VOPT_KEYS + 3) / 4);
    hipLaunchKernelGGL (k_vopt_row, grid, dim3 (256), 0, st, s, ssq, cnt, nullptr, err, split);
    for (int j = 1; j < bins; j++) {
        const double* prev = err + (size_t) ((j - 1) & 1) * VOPT_KEYS;
        double* cur = err + (size_t) (j & 1) * VOPT_KEYS;
        hipLaunchKernelGGL (k_vopt_row, grid, dim3 (256), 0, st, s, ssq, cnt, prev, cur,
            split + (size_t) j * VOPT_KEYS);
    }
    return hipGetLastError ();
}

} // namespace plmhip
