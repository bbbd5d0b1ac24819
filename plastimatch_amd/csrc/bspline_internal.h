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
// plm_hip_bspline_score_device and plm_hip_bspline_regularize (accumulate
// 1) on stream st, without stage timing, every kernel returning at once
// when *skip (a device int, may be null) is nonzero at its launch: the
// evaluation steps of an optimizer chain that may be replayed after its
// run ended.  With hook (MSE with a gradient), the hook launches in place
// of the gradient reduce and d_grad is not written.  ctx_reg_prepare
// allocates the regularizer's scratch ahead of a graph capture.
struct ReduceHook;
int ctx_eval (plm_hip_ctx* c, int metric, int grad_mode, const float* d_coeff, double* d_score,
    double* d_grad, int64_t* d_num_vox, hipStream_t st, const int* skip, const ReduceHook* hook = nullptr);
int ctx_reg_prepare (plm_hip_ctx* c);

// The optimizer's device workspace, kept by a context between runs (freed
// with it): the slot, and how the context frees what it holds.
struct LbWorkspace;
LbWorkspace** ctx_lb_workspace (plm_hip_ctx* c);
void lb_workspace_free (LbWorkspace* w);
int ctx_regularize (plm_hip_ctx* c, const plm_reg_parms* parms, const float* d_coeff, double* d_score,
    double* d_grad, hipStream_t st, const int* skip);

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
