// Helpers shared by the C-ABI source files of libplmregisterhip.so (internal).
#ifndef PLM_BSPLINE_INTERNAL_H
#define PLM_BSPLINE_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "plm_hip_bspline.h"

namespace plmhip {

// Record msg as plm_hip_last_error () and return code.
int abi_fail (int code, const std::string& msg);
// The same for a failed HIP call.
int abi_hip_fail (hipError_t e, const char* what);
// Validate an image descriptor: non-null data, known location, dims >= 2,
// non-zero spacing, fewer than 2^31 / 3 voxels.
int abi_check_img (const plm_img_desc* d, const char* name);
// A context's device, stream and knot count.
int ctx_info (const plm_hip_ctx* c, int* device, hipStream_t* stream, int64_t* num_knots);

struct KGeom;
// The tile-kernel geometry of a B-spline grid over a fixed/moving pair
// (validated; the moving-index map also in double), its tile and knot
// counts.
int make_geometry (const plm_bsp_geom* geom, const plm_img_desc* fixed, const plm_img_desc* moving,
    KGeom& g, double moff_d[3], double mstep_d[9], int64_t& tiles, int64_t& knots);
// Cubic B-spline basis B_t(i / vpr) for i < vpr, t < 4, as
// bspline_basis_eval rounds it (src/base/bspline_xform.cxx:73-99).
void abi_basis_lut (int vpr, float* lut);

} // namespace plmhip

#endif
