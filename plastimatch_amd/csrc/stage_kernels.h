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

} // namespace plmhip

#endif
