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

} // namespace plmhip

#endif
