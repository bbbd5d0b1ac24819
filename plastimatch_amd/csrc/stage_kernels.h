// Kernels of the stage preprocessing (internal): subsampling onto a new
// grid and the min/max scan of the equal-spacing histogram.
#ifndef PLM_STAGE_KERNELS_H
#define PLM_STAGE_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plmhip {

// One output index along one axis, precomputed on the host: whether its
// rounded input coordinate lies inside the input, and the input index and
// weights the voxel uses -- the floor index with li_clamp's fractions
// (trilinear) or the rounded index (nearest).
struct ResampleAxis {
    int idx;
    float w1;               // weight of idx      (li_frac_1)
    float w2;               // weight of idx + 1  (li_frac_2)
    int inside;
};

// out (od0 x od1 x od2, x fastest) from in (id0 x id1 x id2); axis tables
// ax[od0], ay[od1], az[od2].
hipError_t launch_resample (const float* in, const int id[3], float* out, const int od[3],
    const ResampleAxis* ax, const ResampleAxis* ay, const ResampleAxis* az, int nearest,
    hipStream_t st);

constexpr int MINMAX_BLOCKS = 1024;

// Per-workgroup minimum and maximum of img[0..n): part[2 b] = min,
// part[2 b + 1] = max, b < the returned block count.
hipError_t launch_minmax (const float* img, int64_t n, float* part, int* blocks, hipStream_t st);

// Keys of the V-optimal histogram (the equal-spacing grid of VOPT_KEYS keys).
constexpr int VOPT_KEYS = 1000;

// Voxel count and sum of (v - offset) per key, key = floor((v - offset) /
// delta) clamped to the grid, added into count[VOPT_KEYS] and
// sum[VOPT_KEYS] (zeroed by the caller).  Every partial sum of one key is
// exact (see stage_abi.cpp), so the order of the additions does not change
// a bit.
hipError_t launch_vopt_keys (const float* img, int64_t n, float offset, float delta,
    unsigned long long* count, double* sum, hipStream_t st);

// The V-optimal grouping of the keys into bins: the dynamic programme of
// bspline_initialize_mi_hist_vopt over the prefix tables s, ssq, cnt
// (VOPT_KEYS each); split[j * VOPT_KEYS + i] receives the last key of bin
// j - 1 in the best grouping of keys 0..i into j + 1 bins (row 0 unused).
// One launch a row; err (2 VOPT_KEYS doubles) holds the last two rows.
hipError_t launch_vopt_dp (const double* s, const double* ssq, const double* cnt, int bins,
    double* err, int* split, hipStream_t st);

} // namespace plmhip

#endif
