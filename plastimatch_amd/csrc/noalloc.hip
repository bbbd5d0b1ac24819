// The allocation-free MSE evaluation, plm_hip_bspline_score_noalloc (no
// counterpart in the reference): for callers that cannot allocate device
// memory -- code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc.  The caller passes device images, coefficients and
// outputs; every intermediate of the evaluation (the pair table, the basis
// LUTs, the per-tile moving-index frames, the per-workgroup score and count
// partials and the slot-major condense buffer) lives in one file-scope
// device array of PLM_HIP_NOALLOC_SCRATCH_BYTES, which each GPU holds until
// the process ends.  A setup kernel fills it from kernel arguments (no host
// memory is referenced after the call returns) and a fill kernel clears the
// condense buffer, so a call leaves nothing behind that the next one reads.
// Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU (two streams, two threads) race.

#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include <string>

#include "plm_hip_bspline.h"
#include "bspline_internal.h"
#include "bspline_kernels.h"

namespace plmhip {

// The scratch array.  In host code its name is only a stand-in for the
// symbol: the device address comes from hipGetSymbolAddress, per device.
static __device__ __align__ (256) unsigned char scratch[PLM_HIP_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;

// Byte offsets of the evaluation's buffers in the scratch array (each
// 256-aligned).
struct Layout {
    size_t pair, lut, part_score, part_count, ctr, tbase, cond, total;
};

size_t
up256 (size_t n)
{
    return (n + 255) & ~(size_t) 255;
}

Layout
layout (int64_t tiles, int64_t knots)
{
    Layout L;
    size_t o = 0;
    L.pair = o;       o += up256 (sizeof (PairDev));
    L.lut = o;        o += up256 (sizeof (float) * 12 * MAX_VPR);
    // the score per tile; one overlap count per workgroup of the
    // persistent grid (at most one per tile, rounded up to a multiple of 8);
    // the tile kernel's counters
    L.part_score = o; o += up256 (sizeof (double) * (size_t) (tiles + 8));
    L.part_count = o; o += up256 (sizeof (int) * (size_t) (tiles + 8));
    L.ctr = o;        o += up256 (sizeof (unsigned long long) * TILE_CTR_WORDS);
    L.tbase = o;      o += up256 (sizeof (TileBase) * (size_t) tiles);
    L.cond = o;       o += up256 (sizeof (float) * 192 * (size_t) knots);
    L.total = o;
    return L;
}

// What the setup kernel writes, passed by value.
struct SetupArgs {
    KGeom g;
    double moff[3];
    double mstep[9];
    float lut[12 * MAX_VPR];
    int nlut;
    const void* fimg;
    const void* mimg;
    const void* mgrad;
};

__global__ __launch_bounds__ (256)
void k_noalloc_setup (SetupArgs a, unsigned char* base, Layout L)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    TileBase* tb = reinterpret_cast<TileBase*> (base + L.tbase);
    if (i < a.g.num_tiles) tb[i] = tile_base_of (a.g, a.moff, a.mstep, i);
    if (i < a.nlut) reinterpret_cast<float*> (base + L.lut)[i] = a.lut[i];
    // (each tile launch leaves its counters zero; cleared here all the
    // same, should an earlier call have stopped short)
    if (i < TILE_CTR_WORDS) reinterpret_cast<unsigned long long*> (base + L.ctr)[i] = 0ull;
    if (i == 0) {
        PairDev pd {};
        pd.g = a.g;
        pd.tbase = tb;
        pd.fimg = a.fimg;
        pd.mimg = a.mimg;
        pd.mgrad = a.mgrad;
        pd.mpair = nullptr;
        pd.cond = reinterpret_cast<float*> (base + L.cond);
        *reinterpret_cast<PairDev*> (base + L.pair) = pd;
    }
}

// Per-device state found on the first call on that device.
struct DevState {
    bool ready = false;
    unsigned char* ptr = nullptr;   // the device address of scratch
    int num_cus = 0;
    size_t lds_limit = 0;
};

std::mutex g_mu;
DevState g_dev[MAX_DEVICES];

int
device_state (int gpuid, DevState* out)
{
    std::lock_guard<std::mutex> lock (g_mu);
    DevState& d = g_dev[gpuid];
    if (!d.ready) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress (&p, HIP_SYMBOL (scratch));
        if (e != hipSuccess) return abi_hip_fail (e, "hipGetSymbolAddress(scratch)");
        int cus = 0, lds = 0;
        e = hipDeviceGetAttribute (&cus, hipDeviceAttributeMultiprocessorCount, gpuid);
        if (e != hipSuccess) return abi_hip_fail (e, "hipDeviceGetAttribute(CUs)");
        e = hipDeviceGetAttribute (&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, gpuid);
        if (e != hipSuccess) return abi_hip_fail (e, "hipDeviceGetAttribute(LDS)");
        d.ptr = static_cast<unsigned char*> (p);
        d.num_cus = cus;
        d.lds_limit = (size_t) lds;
        d.ready = true;
    }
    *out = d;
    return PLM_HIP_OK;
}

int
check_device_img (const plm_img_desc* d, const char* name)
{
    int rc = abi_check_img (d, name);
    if (rc) return rc;
    if (d->location != PLM_MEM_DEVICE)
        return abi_fail (PLM_HIP_ERR_ARG, std::string (name) + " must be in device memory");
    if (d->storage != PLM_STORE_F32)
        return abi_fail (PLM_HIP_ERR_UNSUPPORTED, std::string (name) + " must be float32");
    return PLM_HIP_OK;
}

} // namespace
} // namespace plmhip

extern "C" int
plm_hip_bspline_score_noalloc (const plm_bsp_geom* geom, const plm_img_desc* fixed,
    const plm_img_desc* moving, const plm_img_desc* moving_grad, int grad_mode, int gpuid,
    const float* d_coeff, double* d_score, double* d_grad, int64_t* d_num_vox, void* stream)
{
    using namespace plmhip;
    if (!geom) return abi_fail (PLM_HIP_ERR_ARG, "geom is NULL");
    if (grad_mode != PLM_GRAD_REF_NN && grad_mode != PLM_GRAD_EXACT && grad_mode != PLM_GRAD_NONE)
        return abi_fail (PLM_HIP_ERR_ARG, "bad grad_mode");
    if (!d_coeff || !d_score || !d_num_vox || (!d_grad && grad_mode != PLM_GRAD_NONE))
        return abi_fail (PLM_HIP_ERR_ARG, "NULL output or coefficient pointer");
    if (gpuid < 0 || gpuid >= MAX_DEVICES) return abi_fail (PLM_HIP_ERR_ARG, "gpuid out of range");
    int rc = check_device_img (fixed, "fixed");
    if (rc) return rc;
    rc = check_device_img (moving, "moving");
    if (rc) return rc;
    if (grad_mode == PLM_GRAD_REF_NN) {
        if (!moving_grad) return abi_fail (PLM_HIP_ERR_ARG, "PLM_GRAD_REF_NN needs moving_grad");
        if (!moving_grad->data) return abi_fail (PLM_HIP_ERR_ARG, "moving_grad data is NULL");
        if (moving_grad->location != PLM_MEM_DEVICE || moving_grad->storage != PLM_STORE_F32)
            return abi_fail (PLM_HIP_ERR_ARG, "moving_grad must be float32 in device memory");
        for (int a = 0; a < 3; a++)
            if (moving_grad->dim[a] != moving->dim[a])
                return abi_fail (PLM_HIP_ERR_GEOMETRY, "moving_grad dims differ from moving");
    }
    KGeom g;
    double moff[3], mstep[9];
    int64_t tiles = 0, knots = 0;
    rc = make_geometry (geom, fixed, moving, g, moff, mstep, tiles, knots);
    if (rc) return rc;
    const Layout L = layout (tiles, knots);
    if (L.total > (size_t) PLM_HIP_NOALLOC_SCRATCH_BYTES)
        return abi_fail (PLM_HIP_ERR_UNSUPPORTED,
            "B-spline grid too large for the allocation-free entry point: it needs "
            + std::to_string (L.total) + " bytes of scratch (192 floats a knot plus 76 bytes a tile), "
            + std::to_string ((size_t) PLM_HIP_NOALLOC_SCRATCH_BYTES) + " are reserved");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount (&ndev);
    if (e != hipSuccess || ndev < 1) return abi_fail (PLM_HIP_ERR_NO_DEVICE, "no HIP device");
    if (gpuid >= ndev) return abi_fail (PLM_HIP_ERR_ARG, "gpuid out of range");
    e = hipSetDevice (gpuid);
    if (e != hipSuccess) return abi_hip_fail (e, "hipSetDevice");
    DevState ds;
    rc = device_state (gpuid, &ds);
    if (rc) return rc;
    if (tile_lds_bytes (g, 2) > ds.lds_limit)
        return abi_fail (PLM_HIP_ERR_UNSUPPORTED, "tile does not fit in LDS; use a smaller vox_per_rgn");

    hipStream_t st = (hipStream_t) stream;
    SetupArgs a {};
    a.g = g;
    for (int k = 0; k < 3; k++) a.moff[k] = moff[k];
    for (int k = 0; k < 9; k++) a.mstep[k] = mstep[k];
    a.nlut = 4 * (g.vpr[0] + g.vpr[1] + g.vpr[2]);
    abi_basis_lut (g.vpr[0], &a.lut[0]);
    abi_basis_lut (g.vpr[1], &a.lut[4 * g.vpr[0]]);
    abi_basis_lut (g.vpr[2], &a.lut[4 * (g.vpr[0] + g.vpr[1])]);
    a.fimg = fixed->data;
    a.mimg = moving->data;
    a.mgrad = grad_mode == PLM_GRAD_REF_NN ? moving_grad->data : nullptr;
    const int n = (int) std::max<int64_t> (std::max<int64_t> (tiles, a.nlut), TILE_CTR_WORDS);
    hipLaunchKernelGGL (k_noalloc_setup, dim3 ((n + 255) / 256), dim3 (256), 0, st, a, ds.ptr, L);
    e = hipGetLastError ();
    if (e != hipSuccess) return abi_hip_fail (e, "noalloc setup");
    // slots no tile owns must read zero in the reduce: clear what the last
    // call left
    e = launch_zero (reinterpret_cast<float*> (ds.ptr + L.cond), 192 * knots, st);
    if (e != hipSuccess) return abi_hip_fail (e, "noalloc clear");

    Plan plan {};
    plan.d_pairs = reinterpret_cast<const PairDev*> (ds.ptr + L.pair);
    plan.npairs = 1;
    plan.g0 = g;
    plan.total_tiles = (int) tiles;
    plan.reduce_blocks = (int) ((3 * knots + REDUCE_BLOCK - 1) / REDUCE_BLOCK);
    plan.blut = reinterpret_cast<const float*> (ds.ptr + L.lut);
    plan.part_score = reinterpret_cast<double*> (ds.ptr + L.part_score);
    plan.part_count = reinterpret_cast<int*> (ds.ptr + L.part_count);
    plan.tile_ctr = reinterpret_cast<unsigned long long*> (ds.ptr + L.ctr);
    Tune tune {2, ds.lds_limit, ds.num_cus, 0};
    const EvalOut out {d_grad, d_score, d_num_vox};
    const int mode = grad_mode == PLM_GRAD_REF_NN ? MSE_REF_NN
        : grad_mode == PLM_GRAD_EXACT ? MSE_EXACT : MSE_SCORE_ONLY;
    e = launch_mse_plain (plan, tune, mode, d_coeff, out, st);
    if (e != hipSuccess) return abi_hip_fail (e, "noalloc evaluation");
    return PLM_HIP_OK;
}
