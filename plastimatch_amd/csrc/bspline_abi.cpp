// C ABI of libplmregisterhip.so: context lifetime, uploads, dispatch.
// Declared in include/plm_hip_bspline.h.

#include <hip/hip_runtime.h>
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "plm_hip_bspline.h"
#include "bspline_internal.h"
#include "bspline_kernels.h"

using plmhip::HGeom;
using plmhip::KGeom;

namespace {

thread_local std::string g_last_error;

int
fail (int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int
hip_fail (hipError_t e, const char* what)
{
    return fail (PLM_HIP_ERR_HIP, std::string (what) + ": " + hipGetErrorString (e));
}

#define HIP_TRY(expr)                                                   \
    do {                                                                \
        hipError_t e_ = (expr);                                         \
        if (e_ != hipSuccess) return hip_fail (e_, #expr);              \
    } while (0)

// 3x3 inverse by cofactors, float, as volume_matrix3x3inverse
// (src/base/direction_cosines.cxx:134-159).  Returns false if singular.
bool
mat3_inverse (float* out, const float* m)
{
    float det = m[0]*(m[4]*m[8]-m[5]*m[7])
        - m[1]*(m[3]*m[8]-m[5]*m[6])
        + m[2]*(m[3]*m[7]-m[4]*m[6]);
    if (std::fabs (det) < 1e-8) return false;
    out[0] =  (m[4]*m[8]-m[5]*m[7]) / det;
    out[1] = -(m[1]*m[8]-m[2]*m[7]) / det;
    out[2] =  (m[1]*m[5]-m[2]*m[4]) / det;
    out[3] = -(m[3]*m[8]-m[5]*m[6]) / det;
    out[4] =  (m[0]*m[8]-m[2]*m[6]) / det;
    out[5] = -(m[0]*m[5]-m[2]*m[3]) / det;
    out[6] =  (m[3]*m[7]-m[4]*m[6]) / det;
    out[7] = -(m[0]*m[7]-m[1]*m[6]) / det;
    out[8] =  (m[0]*m[4]-m[1]*m[3]) / det;
    return true;
}

// The same cofactor inverse in double.
bool
mat3_inverse_d (double* out, const float* mf)
{
    double m[9];
    for (int i = 0; i < 9; i++) m[i] = mf[i];
    double det = m[0]*(m[4]*m[8]-m[5]*m[7])
        - m[1]*(m[3]*m[8]-m[5]*m[6])
        + m[2]*(m[3]*m[7]-m[4]*m[6]);
    if (std::fabs (det) < 1e-8) return false;
    out[0] =  (m[4]*m[8]-m[5]*m[7]) / det;
    out[1] = -(m[1]*m[8]-m[2]*m[7]) / det;
    out[2] =  (m[1]*m[5]-m[2]*m[4]) / det;
    out[3] = -(m[3]*m[8]-m[5]*m[6]) / det;
    out[4] =  (m[0]*m[8]-m[2]*m[6]) / det;
    out[5] = -(m[0]*m[5]-m[2]*m[3]) / det;
    out[6] =  (m[3]*m[7]-m[4]*m[6]) / det;
    out[7] = -(m[0]*m[7]-m[1]*m[6]) / det;
    out[8] =  (m[0]*m[4]-m[1]*m[3]) / det;
    return true;
}

// Per-axis cubic B-spline basis B_t(i/vpr): float argument, double
// polynomial, float result, as bspline_basis_eval
// (src/base/bspline_xform.cxx:73-99).
void
basis_lut (int vpr, float* lut)
{
    for (int i = 0; i < vpr; i++) {
        float u = (float) i / (float) vpr;
        double du = u;
        lut[i*4+0] = (float) ((1.0/6.0)
            * (((-1.0 * du) * du) * du + (3.0 * du) * du - 3.0 * du + 1.0));
        lut[i*4+1] = (float) ((1.0/6.0)
            * (((3.0 * du) * du) * du - (6.0 * du) * du + 4.0));
        lut[i*4+2] = (float) ((1.0/6.0)
            * (((-3.0 * du) * du) * du + (3.0 * du) * du + 3.0 * du + 1.0));
        lut[i*4+3] = (float) ((1.0/6.0) * (((1.0 * du) * du) * du));
    }
}

// Per-stage kernel timing with HIP events recorded on the launch stream:
// one event set per evaluation while enabled, summed and recycled by
// collect().
struct StageTimer {
    bool on = false;
    int last_stages = 0;
    std::vector<std::vector<hipEvent_t>> used;
    std::vector<std::vector<hipEvent_t>> free_sets;

    // The event set for the next evaluation, or nullptr when timing is off.
    hipError_t next (hipEvent_t** ev)
    {
        *ev = nullptr;
        if (!on) return hipSuccess;
        if (free_sets.empty ()) {
            std::vector<hipEvent_t> s (plmhip::MAX_STAGE_EVENTS);
            for (auto& x : s) {
                hipError_t e = hipEventCreate (&x);
                if (e != hipSuccess) return e;
            }
            free_sets.push_back (s);
        }
        used.push_back (free_sets.back ());
        free_sets.pop_back ();
        *ev = used.back ().data ();
        return hipSuccess;
    }

    // Waits for the recorded evaluations and sums their stage times.
    hipError_t collect (double* stage_ms, int max_stages, int64_t* evals)
    {
        for (int i = 0; i < max_stages; i++) stage_ms[i] = 0.0;
        int64_t n = 0;
        for (auto& s : used) {
            hipError_t e = hipEventSynchronize (s[last_stages]);
            if (e != hipSuccess) return e;
            for (int i = 0; i < last_stages && i < max_stages; i++) {
                float ms = 0.f;
                e = hipEventElapsedTime (&ms, s[i], s[i + 1]);
                if (e != hipSuccess) return e;
                stage_ms[i] += ms;
            }
            n++;
            free_sets.push_back (s);
        }
        used.clear ();
        *evals = n;
        return hipSuccess;
    }

    void destroy ()
    {
        for (auto* sets : {&used, &free_sets})
            for (auto& s : *sets)
                for (hipEvent_t e : s) (void) hipEventDestroy (e);
        used.clear ();
        free_sets.clear ();
    }
};

} // namespace

struct plm_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    KGeom g {};
    HGeom h {};
    bool have_hist = false;
    bool have_mgrad = false;
    bool mi_ok = false;
    int storage = PLM_STORE_F32;
    int hist_blocks = 512;
    plmhip::Tune tune {2, 0, 0};
    int64_t num_knots = 0;
    int64_t num_tiles = 0;
    size_t lds_limit = 0;
    float grid_spac[3] = {0.f, 0.f, 0.f};  // vox_per_rgn * |spacing| (Bspline_header::set)

    void* d_fixed = nullptr;
    void* d_moving = nullptr;
    void* d_mpair = nullptr;        // the moving image as z-pairs (storage type)
    void* d_mgrad = nullptr;
    float* d_blut = nullptr;
    plmhip::TileBase* d_tbase = nullptr;    // per-tile moving-index frames
    float* d_cond = nullptr;
    double* d_tile_score = nullptr;
    double* d_reg_part = nullptr;           // regularization scratch (first use)
    double* d_reg_cond = nullptr;
    int* d_tile_count = nullptr;
    unsigned long long* d_tile_ctr = nullptr;   // the tile kernel's per-XCD counters
    plmhip::PairDev* d_pair = nullptr;      // this context as a one-pair table
    plmhip::Plan plan {};
    unsigned long long* d_hist = nullptr;   // joint bins then fixed bins
    float* d_table = nullptr;
    float* d_nvox_f = nullptr;
    uint8_t* d_fbin = nullptr;              // MI bin images (once per stage)
    uint8_t* d_mbin = nullptr;
    uint32_t* d_mquad = nullptr;    // moving bins of each voxel's (x, y) quad
    float* d_ptab = nullptr;        // MI pass-1 walk positions (only where they differ from direct)
    // regions of interest (plm_hip_bspline_set_roi)
    float m_dc[9] = {};             // the moving image's direction cosines and spacing
    float m_sp[3] = {};
    uint8_t* d_fmask = nullptr;     // the fixed ROI at every fixed voxel's position
    uint8_t* d_fmask1 = nullptr;    // the same at MI pass 1's walk positions (with d_ptab)
    uint8_t* d_mroi = nullptr;      // the moving ROI volume
    plmhip::RoiVol mroi {};
    // staging for the host-pointer entry point
    float* d_coeff = nullptr;
    double* d_grad = nullptr;
    double* d_score = nullptr;
    int64_t* d_nvox = nullptr;
    int64_t* d_hist_out = nullptr;

    StageTimer timer;       // plm_hip_bspline_set_timing
    plmhip::LbWorkspace* lb_ws = nullptr;   // the optimizer's buffers (first run)
};

// A batch of contexts evaluated together (plm_hip_batch_*).
struct plm_hip_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<plm_hip_ctx*> ctxs;
    int storage = PLM_STORE_F32;
    bool all_mgrad = true;
    std::vector<plmhip::PairDev> h_pairs;   // host copy of the device table
    plmhip::PairDev* d_pairs = nullptr;
    plmhip::Plan plan {};
    plmhip::Tune tune {2, 0, 0};
    double* d_part_score = nullptr;
    int* d_part_count = nullptr;
    unsigned long long* d_tile_ctr = nullptr;
    bool buffers_set = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    StageTimer timer;       // plm_hip_batch_set_timing
};

namespace {

void
free_ctx (plm_hip_ctx* c)
{
    c->timer.destroy ();
    plmhip::lb_workspace_free (c->lb_ws);
    void* bufs[] = {c->d_fixed, c->d_moving, c->d_mpair, c->d_mgrad, c->d_blut, c->d_tbase, c->d_cond,
        c->d_tile_score, c->d_tile_count, c->d_tile_ctr, c->d_pair, c->d_hist,
        c->d_table, c->d_nvox_f, c->d_fbin, c->d_mbin, c->d_mquad, c->d_ptab, c->d_coeff, c->d_grad,
        c->d_score, c->d_nvox,
        c->d_hist_out, c->d_reg_part, c->d_reg_cond, c->d_fmask, c->d_fmask1, c->d_mroi};
    for (void* b : bufs) if (b) (void) hipFree (b);
    if (c->stream) (void) hipStreamDestroy (c->stream);
    delete c;
}

// The context as the tile kernels see it; the evaluation's buffers come
// with each call, or from a batch's table.
plmhip::PairDev
pair_of (const plm_hip_ctx* c)
{
    plmhip::PairDev pd {};
    pd.g = c->g;
    pd.tbase = c->d_tbase;
    pd.fimg = c->d_fixed;
    pd.mimg = c->d_moving;
    pd.mgrad = c->d_mgrad;
    pd.mpair = c->d_mpair;
    pd.cond = c->d_cond;
    pd.fmask = c->d_fmask;
    pd.mroi = c->mroi;
    return pd;
}

template <typename T>
int
alloc (T** p, size_t count, const char* what)
{
    hipError_t e = hipMalloc ((void**) p, sizeof (T) * (count ? count : 1));
    if (e != hipSuccess) return hip_fail (e, what);
    return PLM_HIP_OK;
}

// Upload one image into a library-owned buffer, converting to BF16 if asked.
int
upload_image (plm_hip_ctx* c, const plm_img_desc* d, int64_t nvals, void** out)
{
    const size_t fbytes = sizeof (float) * nvals;
    float* tmp = nullptr;
    int rc = alloc (&tmp, nvals, "hipMalloc(image)");
    if (rc) return rc;
    hipMemcpyKind kind = d->location == PLM_MEM_DEVICE ? hipMemcpyDeviceToDevice
                                                       : hipMemcpyHostToDevice;
    hipError_t e = hipMemcpy (tmp, d->data, fbytes, kind);
    if (e != hipSuccess) { (void) hipFree (tmp); return hip_fail (e, "hipMemcpy(image)"); }
    if (c->storage == PLM_STORE_F32) {
        *out = tmp;
        return PLM_HIP_OK;
    }
    // 16 bytes of tail padding: BF16 pairs and gradient samples are read as
    // 64-bit words from the dword at or below them
    uint16_t* bf = nullptr;
    rc = alloc (&bf, nvals + 8, "hipMalloc(bf16 image)");
    if (rc) { (void) hipFree (tmp); return rc; }
    e = plmhip::launch_to_bf16 (tmp, bf, nvals, c->stream);
    if (e == hipSuccess) e = hipMemsetAsync (bf + nvals, 0, 16, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize (c->stream);
    (void) hipFree (tmp);
    if (e != hipSuccess) { (void) hipFree (bf); return hip_fail (e, "bf16 conversion"); }
    *out = bf;
    return PLM_HIP_OK;
}

// The moving-index frame of every tile (plmhip::TileBase).
std::vector<plmhip::TileBase>
tile_bases (const KGeom& g, const double moff[3], const double mstep[9])
{
    std::vector<plmhip::TileBase> out ((size_t) g.num_tiles);
    for (int t = 0; t < g.num_tiles; t++) out[(size_t) t] = plmhip::tile_base_of (g, moff, mstep, t);
    return out;
}

int
check_img (const plm_img_desc* d, const char* name)
{
    if (!d || !d->data) return fail (PLM_HIP_ERR_ARG, std::string (name) + " is NULL");
    if (d->location != PLM_MEM_HOST && d->location != PLM_MEM_DEVICE)
        return fail (PLM_HIP_ERR_ARG, std::string (name) + ": bad location");
    int64_t n = 1;
    for (int a = 0; a < 3; a++) {
        if (d->dim[a] < 2) return fail (PLM_HIP_ERR_GEOMETRY, std::string (name) + ": dim < 2");
        if (!(d->spacing[a] != 0.f)) return fail (PLM_HIP_ERR_GEOMETRY, std::string (name) + ": zero spacing");
        n *= d->dim[a];
    }
    if (3 * n >= ((int64_t) 1 << 31))
        return fail (PLM_HIP_ERR_GEOMETRY, std::string (name) + ": too many voxels for 32-bit indexing");
    return PLM_HIP_OK;
}

} // namespace

namespace plmhip {

// The tile-kernel geometry of a B-spline grid over a fixed/moving pair:
// validates the grid (Bspline_header::set, src/base/bspline_header.cxx:28-62)
// and the direction cosines, forms step/proj (src/base/direction_matrices.cxx:
// 10-20) and the moving-index map M = P step, moff = P (origin - moving
// origin) in double (moff_d, mstep_d) and float (g), and checks the
// 32-bit offset limits of the tile kernel.
int
make_geometry (const plm_bsp_geom* geom, const plm_img_desc* fixed, const plm_img_desc* moving,
    KGeom& g, double moff_d[3], double mstep_d[9], int64_t& tiles, int64_t& knots)
{
    g = KGeom {};
    tiles = 1;
    knots = 1;
    for (int a = 0; a < 3; a++) {
        int64_t vpr = geom->vox_per_rgn[a];
        if (vpr < 1 || vpr > plmhip::MAX_VPR)
            return fail (PLM_HIP_ERR_GEOMETRY, "vox_per_rgn out of range [1, 32]");
        if (geom->roi_dim[a] < 1 || geom->roi_offset[a] < 0
            || geom->roi_offset[a] + geom->roi_dim[a] > fixed->dim[a])
            return fail (PLM_HIP_ERR_GEOMETRY, "ROI outside the fixed image");
        if (geom->img_dim[a] != fixed->dim[a])
            return fail (PLM_HIP_ERR_GEOMETRY, "img_dim differs from the fixed image dims");
        // Bspline_header::set (src/base/bspline_header.cxx:28-62)
        int64_t rd = 1 + (geom->roi_dim[a] - 1) / vpr;
        g.vpr[a] = (int) vpr;
        g.rdims[a] = (int) rd;
        g.cdims[a] = (int) (rd + 3);
        g.roi_off[a] = (int) geom->roi_offset[a];
        g.roi_end[a] = (int) (geom->roi_offset[a] + geom->roi_dim[a]);
        g.fdim[a] = (int) fixed->dim[a];
        g.mdim[a] = (int) moving->dim[a];
        g.img_origin[a] = geom->img_origin[a];
        g.morigin[a] = moving->origin[a];
        g.inv_msp[a] = (float) (1.0 / moving->spacing[a]);
        g.mhi[a] = (float) moving->dim[a] - 0.5f;
        g.mmax[a] = (float) (moving->dim[a] - 2);
        tiles *= rd;
        knots *= rd + 3;
    }
    // step = dc*spacing, proj = inv(dc)/spacing
    // (src/base/direction_matrices.cxx:10-20)
    float finv[9], minv[9];
    if (!mat3_inverse (finv, fixed->direction_cosines) || !mat3_inverse (minv, moving->direction_cosines))
        return fail (PLM_HIP_ERR_GEOMETRY, "singular direction cosines");
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            g.fstep[3*i+j] = fixed->direction_cosines[3*i+j] * fixed->spacing[j];
            g.mproj[3*i+j] = minv[3*i+j] / moving->spacing[i];
        }
    // (the slot-major condense buffer, 192 floats per knot, is addressed
    // with 32-bit byte offsets)
    if (tiles >= (1 << 30) / 192 || knots >= ((int64_t) 1 << 31) / 768)
        return fail (PLM_HIP_ERR_GEOMETRY, "B-spline grid too large");
    g.num_tiles = (int) tiles;
    g.num_knots = (int) knots;
    g.tiles_per_xcd = (int) ((tiles + 7) / 8);
    // The tile kernel maps a fixed voxel to moving index coordinates as
    // mijk = P d + moff + M ijk with M = P step and moff = P (origin - morigin),
    // P the moving proj; M and moff are formed in double from the float
    // inputs, so an identity map is exactly the identity.
    {
        double pd[9];
        if (!mat3_inverse_d (pd, moving->direction_cosines))
            return fail (PLM_HIP_ERR_GEOMETRY, "singular direction cosines");
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) pd[3*i+j] /= (double) moving->spacing[i];
        for (int a = 0; a < 3; a++) {
            double o = 0.0;
            for (int c = 0; c < 3; c++)
                o += pd[3*a+c] * ((double) geom->img_origin[c] - (double) moving->origin[c]);
            moff_d[a] = o;
            g.moff[a] = (float) o;
            for (int b = 0; b < 3; b++) {
                double s = 0.0;
                for (int c = 0; c < 3; c++)
                    s += pd[3*a+c] * (double) fixed->direction_cosines[3*c+b] * (double) fixed->spacing[b];
                mstep_d[3*a+b] = s;
                g.mstep[3*a+b] = (float) s;
            }
        }
    }
    // 32-bit byte offsets and 24-bit products of the tile kernel: every
    // image (and the moving gradient, 12 bytes a voxel), plus the fixed
    // voxels a border tile addresses past the image, stays below 2 GiB; a
    // fixed slice holds < 2^24 voxels and a moving slice's gradient bytes
    // stay below 2^24 (1398101 voxels, e.g. 1024 x 1024).
    {
        const int64_t nf = fixed->dim[0] * fixed->dim[1] * fixed->dim[2];
        const int64_t nm = moving->dim[0] * moving->dim[1] * moving->dim[2];
        const int64_t f01 = fixed->dim[0] * fixed->dim[1], m01 = moving->dim[0] * moving->dim[1];
        if (f01 >= ((int64_t) 1 << 24) || 12 * m01 >= ((int64_t) 1 << 24)
            || moving->dim[0] >= ((int64_t) 1 << 23) || moving->dim[1] >= ((int64_t) 1 << 23)
            || moving->dim[2] >= ((int64_t) 1 << 23)
            || 4 * (nf + plmhip::MAX_VPR * f01) >= ((int64_t) 1 << 31)
            || 12 * nm >= ((int64_t) 1 << 31))
            return fail (PLM_HIP_ERR_UNSUPPORTED,
                "image too large: fixed slices must hold < 2^24 voxels, moving slices "
                "< 1398101 voxels, and images < 2 GiB (moving gradient at 12 bytes a voxel)");
    }
    // exact n/d via __umulhi(n, ceil(2^32/d)) for n*d < 2^32
    g.mag_vx = (uint32_t) ((((uint64_t) 1 << 32) + g.vpr[0] - 1) / g.vpr[0]);
    g.mag_vy = (uint32_t) ((((uint64_t) 1 << 32) + g.vpr[1] - 1) / g.vpr[1]);
    g.mag_vxy = (uint32_t) ((((uint64_t) 1 << 32) + g.vpr[0] * g.vpr[1] - 1) / (g.vpr[0] * g.vpr[1]));
    g.aligned = 1;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            if (i != j && (g.fstep[3*i+j] != 0.f || g.mproj[3*i+j] != 0.f)) g.aligned = 0;

    return PLM_HIP_OK;
}

void
abi_basis_lut (int vpr, float* lut)
{
    basis_lut (vpr, lut);
}

} // namespace plmhip

extern "C" {

int
plm_hip_abi_version (void)
{
    return PLM_HIP_ABI_VERSION;
}

const char*
plm_hip_last_error (void)
{
    return g_last_error.c_str ();
}

int
plm_hip_bspline_create (const plm_bsp_geom* geom, const plm_img_desc* fixed,
    const plm_img_desc* moving, const plm_img_desc* moving_grad,
    const plm_hist_desc* hist, int gpuid, plm_hip_ctx** out)
{
    if (!out) return fail (PLM_HIP_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!geom) return fail (PLM_HIP_ERR_ARG, "geom is NULL");
    int rc = check_img (fixed, "fixed");
    if (rc) return rc;
    rc = check_img (moving, "moving");
    if (rc) return rc;
    if (moving_grad) {
        if (!moving_grad->data) return fail (PLM_HIP_ERR_ARG, "moving_grad data is NULL");
        for (int a = 0; a < 3; a++)
            if (moving_grad->dim[a] != moving->dim[a])
                return fail (PLM_HIP_ERR_GEOMETRY, "moving_grad dims differ from moving");
    }
    if (fixed->storage != moving->storage
        || (fixed->storage != PLM_STORE_F32 && fixed->storage != PLM_STORE_BF16))
        return fail (PLM_HIP_ERR_ARG, "fixed and moving must use the same storage");

    KGeom g {};
    int64_t tiles = 1, knots = 1;
    double moff_d[3], mstep_d[9];
    rc = plmhip::make_geometry (geom, fixed, moving, g, moff_d, mstep_d, tiles, knots);
    if (rc) return rc;

    HGeom h {};
    bool mi_ok = false;
    if (hist) {
        if (hist->fixed_bins < 2 || hist->moving_bins < 2
            || hist->fixed_bins > plmhip::MAX_BINS_1D || hist->moving_bins > plmhip::MAX_BINS_1D
            || hist->fixed_bins * hist->moving_bins > plmhip::MAX_BINS_2D)
            return fail (PLM_HIP_ERR_UNSUPPORTED, "histogram bins out of range (<=256 each, <=4096 joint)");
        if (!(hist->fixed_delta > 0.f) || !(hist->moving_delta > 0.f))
            return fail (PLM_HIP_ERR_ARG, "histogram delta must be positive");
        h.fbins = (int) hist->fixed_bins;
        h.mbins = (int) hist->moving_bins;
        h.f_offset = hist->fixed_offset;
        h.f_delta = hist->fixed_delta;
        h.m_offset = hist->moving_offset;
        h.m_delta = hist->moving_delta;
        // flavor g MI walks the whole fixed image in pass 1
        // (src/register/bspline_mi.cxx:1567-1615); a ROI selects flavor h.
        mi_ok = true;
        for (int a = 0; a < 3; a++)
            if (geom->roi_offset[a] != 0 || geom->roi_dim[a] != fixed->dim[a]) mi_ok = false;
    }

    int ndev = 0;
    hipError_t e = hipGetDeviceCount (&ndev);
    if (e != hipSuccess || ndev < 1) return fail (PLM_HIP_ERR_NO_DEVICE, "no HIP device");
    if (gpuid < 0 || gpuid >= ndev) return fail (PLM_HIP_ERR_ARG, "gpuid out of range");
    HIP_TRY (hipSetDevice (gpuid));
    hipDeviceProp_t prop;
    HIP_TRY (hipGetDeviceProperties (&prop, gpuid));

    plm_hip_ctx* c = new plm_hip_ctx;
    c->device = gpuid;
    c->g = g;
    c->h = h;
    c->have_hist = hist != nullptr;
    c->mi_ok = mi_ok;
    c->have_mgrad = moving_grad != nullptr;
    c->storage = fixed->storage;
    for (int i = 0; i < 9; i++) c->m_dc[i] = moving->direction_cosines[i];
    for (int a = 0; a < 3; a++) c->m_sp[a] = moving->spacing[a];
    c->num_knots = knots;
    c->num_tiles = tiles;
    for (int a = 0; a < 3; a++)
        c->grid_spac[a] = (float) geom->vox_per_rgn[a] * std::fabs (geom->img_spacing[a]);
    c->lds_limit = prop.sharedMemPerBlock;
    c->tune.lds_limit = c->lds_limit;
    c->tune.num_cus = prop.multiProcessorCount;
    // MI pass 1: enough workgroups to keep several waves on every SIMD (the
    // pass is bound by its exact-order FP32 arithmetic), each flushing its
    // private 8.4 KB histogram once
    c->hist_blocks = (int) std::min<int64_t> (tiles, (int64_t) prop.multiProcessorCount * 8);
    if (plmhip::tile_lds_bytes (g, 1) + sizeof (float) * h.fbins * h.mbins > c->lds_limit) {
        free_ctx (c);
        return fail (PLM_HIP_ERR_UNSUPPORTED, "tile does not fit in LDS; use a smaller vox_per_rgn");
    }
    e = hipStreamCreate (&c->stream);
    if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "hipStreamCreate"); }

#define CTX_TRY(expr) do { int r_ = (expr); if (r_) { free_ctx (c); return r_; } } while (0)
    const int64_t nf = fixed->dim[0] * fixed->dim[1] * fixed->dim[2];
    const int64_t nm = moving->dim[0] * moving->dim[1] * moving->dim[2];
    CTX_TRY (upload_image (c, fixed, nf, &c->d_fixed));
    CTX_TRY (upload_image (c, moving, nm, &c->d_moving));
    if (moving_grad) CTX_TRY (upload_image (c, moving_grad, 3 * nm, &c->d_mgrad));
    {
        // 2 values a voxel, in the storage type
        const size_t bytes = 2 * (size_t) nm * (c->storage == PLM_STORE_F32 ? 4 : 2);
        CTX_TRY (alloc ((uint8_t**) &c->d_mpair, bytes, "hipMalloc(moving pairs)"));
        e = plmhip::launch_zpair (c->d_moving, c->storage == PLM_STORE_F32 ? 0 : 1, c->d_mpair, nm,
            moving->dim[0] * moving->dim[1], c->stream);
        if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "moving pairs"); }
    }

    std::vector<float> blut (4 * (g.vpr[0] + g.vpr[1] + g.vpr[2]));
    basis_lut (g.vpr[0], &blut[0]);
    basis_lut (g.vpr[1], &blut[4 * g.vpr[0]]);
    basis_lut (g.vpr[2], &blut[4 * (g.vpr[0] + g.vpr[1])]);
    CTX_TRY (alloc (&c->d_blut, blut.size (), "hipMalloc(lut)"));
    e = hipMemcpy (c->d_blut, blut.data (), sizeof (float) * blut.size (), hipMemcpyHostToDevice);
    if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "hipMemcpy(lut)"); }

    {
        const std::vector<plmhip::TileBase> tb = tile_bases (g, moff_d, mstep_d);
        CTX_TRY (alloc (&c->d_tbase, tb.size (), "hipMalloc(tile bases)"));
        e = hipMemcpy (c->d_tbase, tb.data (), sizeof (plmhip::TileBase) * tb.size (),
            hipMemcpyHostToDevice);
        if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "hipMemcpy(tile bases)"); }
    }

    // slot-major condense buffer; slots no tile owns stay zero for good
    CTX_TRY (alloc (&c->d_cond, (size_t) knots * 192, "hipMalloc(cond)"));
    // partials of the persistent tile kernel: the score per tile, the
    // overlap count per workgroup (at most one workgroup per tile, rounded
    // up to a multiple of 8); its tile counters, zero from here on
    CTX_TRY (alloc (&c->d_tile_score, tiles + 8, "hipMalloc(tile_score)"));
    CTX_TRY (alloc (&c->d_tile_count, tiles + 8, "hipMalloc(tile_count)"));
    CTX_TRY (alloc (&c->d_tile_ctr, plmhip::TILE_CTR_WORDS, "hipMalloc(tile counters)"));
    e = hipMemset (c->d_tile_ctr, 0, sizeof (unsigned long long) * plmhip::TILE_CTR_WORDS);
    if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "hipMemset(tile counters)"); }
    const size_t nbins = (size_t) h.fbins * h.mbins + h.fbins;
    CTX_TRY (alloc (&c->d_hist, nbins, "hipMalloc(hist)"));
    CTX_TRY (alloc (&c->d_table, (size_t) h.fbins * h.mbins, "hipMalloc(table)"));
    CTX_TRY (alloc (&c->d_nvox_f, 1, "hipMalloc(nvox_f)"));
    if (c->have_hist) {
        // bin images: the bin of every voxel is fixed for the stage
        CTX_TRY (alloc (&c->d_fbin, nf, "hipMalloc(fixed bins)"));
        CTX_TRY (alloc (&c->d_mbin, nm, "hipMalloc(moving bins)"));
        e = plmhip::launch_bin_image (c->d_fixed, c->storage == PLM_STORE_BF16, c->d_fbin, nf,
            h.f_offset, h.f_delta, h.fbins, c->stream);
        if (e == hipSuccess)
            e = plmhip::launch_bin_image (c->d_moving, c->storage == PLM_STORE_BF16, c->d_mbin, nm,
                h.m_offset, h.m_delta, h.mbins, c->stream);
        if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "bin images"); }
        CTX_TRY (alloc (&c->d_mquad, nm, "hipMalloc(moving bin quads)"));
        const int mdim[3] = {(int) moving->dim[0], (int) moving->dim[1], (int) moving->dim[2]};
        e = plmhip::launch_bin_quads (c->d_mbin, c->d_mquad, mdim, c->stream);
        if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "bin quads"); }
        // pass 1 walks the fixed image as the reference's LOOP_Z/Y/X do;
        // keep those positions where they differ from the direct formula
        // (not on dyadic axis-aligned grids)
        unsigned int* d_bad = nullptr;
        CTX_TRY (alloc (&d_bad, 1, "hipMalloc(mismatch)"));
        unsigned int bad = 0;
        e = hipMemcpy (d_bad, &bad, sizeof bad, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = plmhip::launch_pass1_positions (g, nullptr, d_bad, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize (c->stream);
        if (e == hipSuccess) e = hipMemcpy (&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost);
        (void) hipFree (d_bad);
        if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "pass-1 positions"); }
        if (bad > 0) {
            CTX_TRY (alloc (&c->d_ptab, 3 * (size_t) nf, "hipMalloc(pass-1 positions)"));
            unsigned int* d_b2 = nullptr;
            CTX_TRY (alloc (&d_b2, 1, "hipMalloc(mismatch)"));
            e = plmhip::launch_pass1_positions (g, c->d_ptab, d_b2, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize (c->stream);
            (void) hipFree (d_b2);
            if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "pass-1 positions"); }
        }
    }
    CTX_TRY (alloc (&c->d_coeff, 3 * knots, "hipMalloc(coeff)"));
    e = hipMemset (c->d_coeff, 0, sizeof (float) * 3 * knots);
    if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "hipMemset(coeff)"); }
    CTX_TRY (alloc (&c->d_grad, 3 * knots, "hipMalloc(grad)"));
    CTX_TRY (alloc (&c->d_score, 1, "hipMalloc(score)"));
    CTX_TRY (alloc (&c->d_nvox, 1, "hipMalloc(nvox)"));
    CTX_TRY (alloc (&c->d_hist_out, nbins, "hipMalloc(hist_out)"));
    {
        // the context as a table of one pair; evaluation buffers come with
        // every call
        const plmhip::PairDev pd = pair_of (c);
        CTX_TRY (alloc (&c->d_pair, 1, "hipMalloc(pair)"));
        e = hipMemcpy (c->d_pair, &pd, sizeof (pd), hipMemcpyHostToDevice);
        if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "hipMemcpy(pair)"); }
        plmhip::Plan& pl = c->plan;
        pl.d_pairs = c->d_pair;
        pl.npairs = 1;
        pl.g0 = g;
        pl.total_tiles = (int) tiles;
        pl.reduce_blocks = (int) ((3 * knots + plmhip::REDUCE_BLOCK - 1) / plmhip::REDUCE_BLOCK);
        pl.blut = c->d_blut;
        pl.part_score = c->d_tile_score;
        pl.part_count = c->d_tile_count;
        pl.tile_ctr = c->d_tile_ctr;
    }
#undef CTX_TRY
    e = plmhip::launch_zero (c->d_cond, (int64_t) knots * 192, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize (c->stream);
    if (e != hipSuccess) { free_ctx (c); return hip_fail (e, "init"); }
    *out = c;
    return PLM_HIP_OK;
}

} // extern "C"

namespace {

// The checks and launches of one evaluation (ev: stage timing events or
// null; skip: see ctx_eval).
int
eval_launch (plm_hip_ctx* c, int metric, int grad_mode, const float* d_coeff, double* d_score,
    double* d_grad, int64_t* d_num_vox, int64_t* d_hist, hipStream_t st, hipEvent_t* ev,
    const int* skip, const plmhip::ReduceHook* hook = nullptr)
{
    if (!d_coeff || !d_score || !d_num_vox || (!d_grad && grad_mode != PLM_GRAD_NONE))
        return fail (PLM_HIP_ERR_ARG, "NULL output or coefficient pointer");
    if (metric != PLM_METRIC_MSE && metric != PLM_METRIC_MI)
        return fail (PLM_HIP_ERR_ARG, "unknown metric");
    if (metric == PLM_METRIC_MSE) {
        if (grad_mode != PLM_GRAD_REF_NN && grad_mode != PLM_GRAD_EXACT && grad_mode != PLM_GRAD_NONE)
            return fail (PLM_HIP_ERR_ARG, "bad grad_mode");
        if (grad_mode == PLM_GRAD_REF_NN && !c->have_mgrad)
            return fail (PLM_HIP_ERR_ARG, "PLM_GRAD_REF_NN needs moving_grad");
    } else {
        if (grad_mode == PLM_GRAD_NONE) return fail (PLM_HIP_ERR_ARG, "PLM_GRAD_NONE is MSE only");
        if (!c->have_hist) return fail (PLM_HIP_ERR_ARG, "MI needs a histogram description");
        if (!c->mi_ok) return fail (PLM_HIP_ERR_UNSUPPORTED, "MI flavor g does not take a ROI");
    }
    hipError_t e;
    plmhip::EvalOut out {d_grad, d_score, d_num_vox};
    out.skip = skip;
    if (metric == PLM_METRIC_MSE) {
        const int mode = grad_mode == PLM_GRAD_REF_NN ? plmhip::MSE_REF_NN
            : grad_mode == PLM_GRAD_EXACT ? plmhip::MSE_EXACT : plmhip::MSE_SCORE_ONLY;
        e = plmhip::launch_mse (c->plan, c->tune, mode, c->storage == PLM_STORE_BF16,
            d_coeff, out, st, ev, hook);
        if (e != hipSuccess) return hip_fail (e, "MSE launch");
    } else {
        e = plmhip::launch_mi (c->plan, c->tune, c->h, c->hist_blocks, c->d_fbin, c->d_mquad,
            d_coeff, c->d_hist, c->d_hist + (size_t) c->h.fbins * c->h.mbins, c->d_table,
            c->d_nvox_f, out, d_hist, c->d_ptab, c->d_ptab ? c->d_fmask1 : c->d_fmask, c->mroi, st, ev);
        if (e != hipSuccess) return hip_fail (e, "MI launch");
    }
    return PLM_HIP_OK;
}

} // namespace

namespace plmhip {

int
ctx_eval (plm_hip_ctx* c, int metric, int grad_mode, const float* d_coeff, double* d_score,
    double* d_grad, int64_t* d_num_vox, hipStream_t st, const int* skip, const ReduceHook* hook)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    HIP_TRY (hipSetDevice (c->device));
    return eval_launch (c, metric, grad_mode, d_coeff, d_score, d_grad, d_num_vox, nullptr, st, nullptr,
        skip, hook);
}

} // namespace plmhip

extern "C" {

int
plm_hip_bspline_score_device (plm_hip_ctx* c, int metric, int grad_mode,
    const float* d_coeff, double* d_score, double* d_grad, int64_t* d_num_vox,
    int64_t* d_hist, void* stream)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    hipStream_t st = stream ? (hipStream_t) stream : c->stream;
    HIP_TRY (hipSetDevice (c->device));
    hipEvent_t* ev = nullptr;
    HIP_TRY (c->timer.next (&ev));
    const int rc = eval_launch (c, metric, grad_mode, d_coeff, d_score, d_grad, d_num_vox, d_hist, st, ev,
        nullptr);
    if (rc) return rc;
    if (ev) c->timer.last_stages = metric == PLM_METRIC_MSE ? plmhip::MSE_STAGES : plmhip::MI_STAGES;
    return PLM_HIP_OK;
}

int
plm_hip_bspline_set_timing (plm_hip_ctx* c, int enable)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    c->timer.on = enable != 0;
    return PLM_HIP_OK;
}

int
plm_hip_bspline_get_timing (plm_hip_ctx* c, double* stage_ms, int max_stages,
    int64_t* evals)
{
    if (!c || !stage_ms || !evals) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    HIP_TRY (hipSetDevice (c->device));
    HIP_TRY (c->timer.collect (stage_ms, max_stages, evals));
    return PLM_HIP_OK;
}

int
plm_hip_bspline_score (plm_hip_ctx* c, int metric, int grad_mode, const float* coeff,
    double* score, double* grad, int64_t* num_vox, int64_t* hist_out)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (!coeff || !score || !grad || !num_vox) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    if (hist_out && metric != PLM_METRIC_MI) return fail (PLM_HIP_ERR_ARG, "hist_out is MI only");
    HIP_TRY (hipSetDevice (c->device));
    const size_t nc = 3 * (size_t) c->num_knots;
    HIP_TRY (hipMemcpyAsync (c->d_coeff, coeff, sizeof (float) * nc, hipMemcpyHostToDevice, c->stream));
    int rc = plm_hip_bspline_score_device (c, metric, grad_mode, c->d_coeff, c->d_score,
        c->d_grad, c->d_nvox, hist_out ? c->d_hist_out : nullptr, c->stream);
    if (rc) return rc;
    HIP_TRY (hipMemcpyAsync (grad, c->d_grad, sizeof (double) * nc, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY (hipMemcpyAsync (score, c->d_score, sizeof (double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY (hipMemcpyAsync (num_vox, c->d_nvox, sizeof (int64_t), hipMemcpyDeviceToHost, c->stream));
    if (hist_out) {
        const size_t nbins = (size_t) c->h.fbins * c->h.mbins + c->h.fbins;
        HIP_TRY (hipMemcpyAsync (hist_out, c->d_hist_out, sizeof (int64_t) * nbins,
            hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY (hipStreamSynchronize (c->stream));
    if (*num_vox < 1) {
        // bspline_score_normalize (src/register/bspline_mse.cxx:53-65); the
        // MI flavor has no guard in the reference and gives the same here.
        *score = (double) FLT_MAX;
        for (size_t i = 0; i < nc; i++) grad[i] = 0.0;
        return PLM_HIP_WARN_NO_OVERLAP;
    }
    return PLM_HIP_OK;
}

int
plm_hip_bspline_set_tuning (plm_hip_ctx* c, int vox_per_thread)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (vox_per_thread != 1 && vox_per_thread != 2 && vox_per_thread != 4)
        return fail (PLM_HIP_ERR_ARG, "vox_per_thread must be 1, 2 or 4");
    c->tune.vox_per_thread = vox_per_thread;
    return PLM_HIP_OK;
}

int
plm_hip_bspline_set_occupancy (plm_hip_ctx* c, int wg_per_cu)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (wg_per_cu < 0) return fail (PLM_HIP_ERR_ARG, "workgroups per CU must be >= 0");
    c->tune.wg_per_cu = wg_per_cu;
    return PLM_HIP_OK;
}

int
plm_hip_bspline_get_tuning (plm_hip_ctx* c, int* vox_per_thread, int* wg_per_cu)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (vox_per_thread) *vox_per_thread = c->tune.vox_per_thread;
    if (wg_per_cu) *wg_per_cu = c->tune.wg_per_cu;
    return PLM_HIP_OK;
}

int
plm_hip_bspline_autotune (plm_hip_ctx* c, int metric, int grad_mode, const float* d_coeff,
    int reps, int* chosen, double* eval_ms)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (reps < 1) reps = 1;
    const float* coeff = d_coeff ? d_coeff : c->d_coeff;
    const bool was_timing = c->timer.on;
    double stage[plmhip::MAX_STAGE_EVENTS];
    int64_t n = 0;
    int rc = plm_hip_bspline_get_timing (c, stage, plmhip::MAX_STAGE_EVENTS, &n);
    if (rc) return rc;
    c->timer.on = true;
    const int cands[3] = {1, 2, 4};
    int best = c->tune.vox_per_thread, best_wg = c->tune.wg_per_cu;
    double best_ms = 1e300;
    for (int k = 0; k < 12; k++) {
        c->tune.vox_per_thread = cands[k % 3];
        c->tune.wg_per_cu = plmhip::TUNE_WG_CANDS[k / 3];
        for (int r = 0; r < reps + 1; r++) {
            rc = plm_hip_bspline_score_device (c, metric, grad_mode, coeff, c->d_score, c->d_grad,
                c->d_nvox, nullptr, c->stream);
            if (rc) { c->timer.on = was_timing; return rc; }
            if (r == 0) {
                // the first call of each variant is a warm-up
                rc = plm_hip_bspline_get_timing (c, stage, plmhip::MAX_STAGE_EVENTS, &n);
                if (rc) { c->timer.on = was_timing; return rc; }
            }
        }
        rc = plm_hip_bspline_get_timing (c, stage, plmhip::MAX_STAGE_EVENTS, &n);
        if (rc) { c->timer.on = was_timing; return rc; }
        double total = 0.0;
        for (int i = 0; i < plmhip::MAX_STAGE_EVENTS; i++) total += stage[i];
        total /= (double) (n ? n : 1);
        if (total < best_ms) { best_ms = total; best = cands[k % 3]; best_wg = plmhip::TUNE_WG_CANDS[k / 3]; }
    }
    c->tune.vox_per_thread = best;
    c->tune.wg_per_cu = best_wg;
    c->timer.on = was_timing;
    if (chosen) *chosen = best;
    if (eval_ms) *eval_ms = best_ms;
    return PLM_HIP_OK;
}

} // extern "C"

namespace {

// A ROI descriptor: one byte a voxel, dims >= 1, non-zero spacing, fewer
// than 2^31 voxels.
int
check_roi (const plm_img_desc* d, const char* name)
{
    if (!d->data) return fail (PLM_HIP_ERR_ARG, std::string (name) + " data is NULL");
    if (d->location != PLM_MEM_HOST && d->location != PLM_MEM_DEVICE)
        return fail (PLM_HIP_ERR_ARG, std::string (name) + ": bad location");
    if (d->storage != PLM_STORE_U8)
        return fail (PLM_HIP_ERR_ARG, std::string (name) + " must be PLM_STORE_U8");
    int64_t n = 1;
    for (int a = 0; a < 3; a++) {
        if (d->dim[a] < 1) return fail (PLM_HIP_ERR_GEOMETRY, std::string (name) + ": dim < 1");
        if (!(d->spacing[a] != 0.f)) return fail (PLM_HIP_ERR_GEOMETRY, std::string (name) + ": zero spacing");
        n *= d->dim[a];
    }
    if (n >= ((int64_t) 1 << 31))
        return fail (PLM_HIP_ERR_GEOMETRY, std::string (name) + ": too many voxels for 32-bit indexing");
    return PLM_HIP_OK;
}

// The ROI's bytes in a new device buffer.
int
upload_roi (const plm_img_desc* d, uint8_t** out)
{
    const size_t n = (size_t) (d->dim[0] * d->dim[1] * d->dim[2]);
    int rc = alloc (out, n, "hipMalloc(roi)");
    if (rc) return rc;
    hipError_t e = hipMemcpy (*out, d->data, n, d->location == PLM_MEM_DEVICE ? hipMemcpyDeviceToDevice
                                                                              : hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void) hipFree (*out);
        *out = nullptr;
        return hip_fail (e, "hipMemcpy(roi)");
    }
    return PLM_HIP_OK;
}

// The kernels' view of ROI d (its bytes at dimg): the float proj the
// reference forms for a volume (inv(dc) / spacing, src/base/
// direction_matrices.cxx:10-20), and for a moving ROI the map from the
// moving image's index coordinates to the ROI's, R = P_roi S_m and
// r0 = P_roi (moving origin - ROI origin), S_m = dc * spacing of the moving
// image, in double.
int
roi_vol (const plm_img_desc* d, const uint8_t* dimg, const plm_hip_ctx* c, bool moving, plmhip::RoiVol& r)
{
    r = plmhip::RoiVol {};
    r.img = dimg;
    float inv[9];
    double pd[9];
    if (!mat3_inverse (inv, d->direction_cosines) || !mat3_inverse_d (pd, d->direction_cosines))
        return fail (PLM_HIP_ERR_GEOMETRY, "singular ROI direction cosines");
    for (int a = 0; a < 3; a++) {
        r.dim[a] = (int) d->dim[a];
        r.origin[a] = d->origin[a];
        for (int b = 0; b < 3; b++) {
            r.proj[3*a+b] = inv[3*a+b] / d->spacing[a];
            pd[3*a+b] /= (double) d->spacing[a];
        }
    }
    if (!moving) return PLM_HIP_OK;
    for (int a = 0; a < 3; a++) {
        double o = 0.0;
        for (int k = 0; k < 3; k++) o += pd[3*a+k] * ((double) c->g.morigin[k] - (double) d->origin[k]);
        r.r0[a] = o;
        for (int b = 0; b < 3; b++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += pd[3*a+k] * (double) c->m_dc[3*k+b] * (double) c->m_sp[b];
            r.Rd[3*a+b] = v;
            r.R[3*a+b] = (float) v;
        }
    }
    return PLM_HIP_OK;
}

void
clear_roi (plm_hip_ctx* c)
{
    for (uint8_t** p : {&c->d_fmask, &c->d_fmask1, &c->d_mroi}) {
        if (*p) (void) hipFree (*p);
        *p = nullptr;
    }
    c->mroi = plmhip::RoiVol {};
}

int
build_roi (plm_hip_ctx* c, const plm_img_desc* fixed_roi, const plm_img_desc* moving_roi)
{
    if (fixed_roi) {
        uint8_t* tmp = nullptr;
        int rc = upload_roi (fixed_roi, &tmp);
        if (rc) return rc;
        plmhip::RoiVol fr;
        rc = roi_vol (fixed_roi, tmp, c, false, fr);
        const size_t nf = (size_t) c->g.fdim[0] * c->g.fdim[1] * c->g.fdim[2];
        if (!rc) rc = alloc (&c->d_fmask, nf, "hipMalloc(fixed ROI mask)");
        if (!rc && c->d_ptab) rc = alloc (&c->d_fmask1, nf, "hipMalloc(fixed ROI mask)");
        hipError_t e = hipSuccess;
        if (!rc) e = plmhip::launch_roi_fixed (c->g, fr, nullptr, c->d_fmask, c->stream);
        if (!rc && e == hipSuccess && c->d_ptab)
            e = plmhip::launch_roi_fixed (c->g, fr, c->d_ptab, c->d_fmask1, c->stream);
        if (!rc && e == hipSuccess) e = hipStreamSynchronize (c->stream);
        (void) hipFree (tmp);
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail (e, "fixed ROI mask");
    }
    if (moving_roi) {
        int rc = upload_roi (moving_roi, &c->d_mroi);
        if (!rc) rc = roi_vol (moving_roi, c->d_mroi, c, true, c->mroi);
        if (rc) return rc;
    }
    return PLM_HIP_OK;
}

} // namespace

extern "C" {

int
plm_hip_bspline_set_roi (plm_hip_ctx* c, const plm_img_desc* fixed_roi, const plm_img_desc* moving_roi)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    int rc = fixed_roi ? check_roi (fixed_roi, "fixed_roi") : PLM_HIP_OK;
    if (!rc && moving_roi) rc = check_roi (moving_roi, "moving_roi");
    if (rc) return rc;
    HIP_TRY (hipSetDevice (c->device));
    HIP_TRY (hipStreamSynchronize (c->stream));
    clear_roi (c);
    rc = build_roi (c, fixed_roi, moving_roi);
    if (rc) clear_roi (c);
    // the one-pair table the single-context launches read
    const plmhip::PairDev pd = pair_of (c);
    hipError_t e = hipMemcpy (c->d_pair, &pd, sizeof (pd), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail (e, "hipMemcpy(pair)");
    return rc;
}

int
plm_hip_bspline_set_hist_keys (plm_hip_ctx* c, const int32_t* fixed_key_lut, const int32_t* moving_key_lut)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (!c->have_hist) return fail (PLM_HIP_ERR_ARG, "the context has no histogram");
    constexpr int K = PLM_HIP_VOPT_KEYS;
    const int32_t* luts[2] = {fixed_key_lut, moving_key_lut};
    const int nbins[2] = {c->h.fbins, c->h.mbins};
    for (int s = 0; s < 2; s++) {
        if (!luts[s]) continue;
        for (int k = 0; k < K; k++)
            if (luts[s][k] < 0 || luts[s][k] >= nbins[s])
                return fail (PLM_HIP_ERR_ARG, s ? "moving_key_lut entry out of range" : "fixed_key_lut entry out of range");
    }
    if (!fixed_key_lut && !moving_key_lut) return PLM_HIP_OK;
    HIP_TRY (hipSetDevice (c->device));
    HIP_TRY (hipStreamSynchronize (c->stream));
    int* d_lut = nullptr;
    int rc = alloc (&d_lut, 2 * (size_t) K, "hipMalloc(key tables)");
    if (rc) return rc;
    const int64_t nf = (int64_t) c->g.fdim[0] * c->g.fdim[1] * c->g.fdim[2];
    const int64_t nm = (int64_t) c->g.mdim[0] * c->g.mdim[1] * c->g.mdim[2];
    const int bf16 = c->storage == PLM_STORE_BF16;
    hipError_t e = hipSuccess;
    if (fixed_key_lut) {
        e = hipMemcpy (d_lut, fixed_key_lut, sizeof (int) * K, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = plmhip::launch_bin_image (c->d_fixed, bf16, c->d_fbin, nf, c->h.f_offset, c->h.f_delta,
                c->h.fbins, c->stream, d_lut, K);
    }
    if (e == hipSuccess && moving_key_lut) {
        e = hipMemcpy (d_lut + K, moving_key_lut, sizeof (int) * K, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = plmhip::launch_bin_image (c->d_moving, bf16, c->d_mbin, nm, c->h.m_offset, c->h.m_delta,
                c->h.mbins, c->stream, d_lut + K, K);
        if (e == hipSuccess) e = plmhip::launch_bin_quads (c->d_mbin, c->d_mquad, c->g.mdim, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize (c->stream);
    (void) hipFree (d_lut);
    if (e != hipSuccess) return hip_fail (e, "V-optimal bin images");
    return PLM_HIP_OK;
}

int
plm_hip_bspline_destroy (plm_hip_ctx* c)
{
    if (!c) return PLM_HIP_OK;
    (void) hipSetDevice (c->device);
    (void) hipStreamSynchronize (c->stream);
    free_ctx (c);
    return PLM_HIP_OK;
}

int
plm_hip_volume_gradient (const plm_img_desc* vol, float* out, int out_location, int gpuid)
{
    int rc = check_img (vol, "vol");
    if (rc) return rc;
    if (!out) return fail (PLM_HIP_ERR_ARG, "out is NULL");
    if (vol->storage != PLM_STORE_F32) return fail (PLM_HIP_ERR_ARG, "vol must be float32");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount (&ndev);
    if (e != hipSuccess || ndev < 1) return fail (PLM_HIP_ERR_NO_DEVICE, "no HIP device");
    if (gpuid < 0 || gpuid >= ndev) return fail (PLM_HIP_ERR_ARG, "gpuid out of range");
    HIP_TRY (hipSetDevice (gpuid));
    plmhip::Mat3 inv;
    if (!mat3_inverse (inv.m, vol->direction_cosines))
        return fail (PLM_HIP_ERR_GEOMETRY, "singular direction cosines");
    const int64_t n = vol->dim[0] * vol->dim[1] * vol->dim[2];
    const int dim[3] = {(int) vol->dim[0], (int) vol->dim[1], (int) vol->dim[2]};
    float* d_in = nullptr;
    float* d_out = nullptr;
    if (vol->location == PLM_MEM_HOST) {
        rc = alloc (&d_in, n, "hipMalloc(gradient input)");
        if (rc) return rc;
        e = hipMemcpy (d_in, vol->data, sizeof (float) * n, hipMemcpyHostToDevice);
        if (e != hipSuccess) { (void) hipFree (d_in); return hip_fail (e, "hipMemcpy(gradient input)"); }
    }
    if (out_location == PLM_MEM_HOST) {
        rc = alloc (&d_out, 3 * n, "hipMalloc(gradient output)");
        if (rc) { if (d_in) (void) hipFree (d_in); return rc; }
    }
    const float* src = d_in ? d_in : (const float*) vol->data;
    float* dst = d_out ? d_out : out;
    e = plmhip::launch_volume_gradient (src, dst, dim, vol->spacing, inv, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize ();
    if (e == hipSuccess && d_out)
        e = hipMemcpy (out, d_out, sizeof (float) * 3 * n, hipMemcpyDeviceToHost);
    if (d_in) (void) hipFree (d_in);
    if (d_out) (void) hipFree (d_out);
    if (e != hipSuccess) return hip_fail (e, "volume gradient");
    return PLM_HIP_OK;
}

int
plm_hip_bspline_interpolate_vf (plm_hip_ctx* c, const float* d_coeff, float* d_vf, void* stream)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (!d_coeff || !d_vf) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    HIP_TRY (hipSetDevice (c->device));
    hipStream_t st = stream ? (hipStream_t) stream : c->stream;
    hipError_t e = plmhip::launch_vf (c->plan, c->tune, d_coeff, d_vf, st);
    if (e != hipSuccess) return hip_fail (e, "displacement field launch");
    return PLM_HIP_OK;
}

namespace {

// V[4j+i] = sum_{a,b} I[a+b] Qn1[4j+a] Qn2[4i+b], I[n] = gs^(n+1)/(n+1):
// one region's integral of a product of two basis polynomials (eval_integral,
// src/register/bspline_regularize_analytic.cxx:167-199).
void
reg_integral (double* V, const double* Qn1, const double* Qn2, double gs)
{
    double I[7];
    double pw = gs;
    for (int n = 0; n < 7; n++) {
        I[n] = pw / (double) (n + 1);
        pw *= gs;
    }
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 4; i++) {
            double acc = 0.0;
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++) acc += I[a + b] * (Qn1[4 * j + a] * Qn2[4 * i + b]);
            V[4 * j + i] = acc;
        }
}

// The per-axis factors and weights of the analytic regularizer
// (init_analytic :201-281 and analytic_init :638-940): Q_n = B R d^n, the
// cubic basis coefficients scaled by the reciprocal grid spacing and
// differentiated n times; factor k < 4 pairs Q_k with Q_k, 4 pairs (Q_1,
// Q_0), 5 pairs (Q_0, Q_1).  The weights are formed in float as the
// reference forms them.
plmhip::RegOp
reg_operator (const float grid_spac[3], const plm_reg_parms& rp)
{
    static const double B[16] = {
        1.0 / 6.0, -1.0 / 2.0, 1.0 / 2.0, -1.0 / 6.0,
        2.0 / 3.0, 0.0, -1.0, 1.0 / 2.0,
        1.0 / 6.0, 1.0 / 2.0, 1.0 / 2.0, -1.0 / 2.0,
        0.0, 0.0, 0.0, 1.0 / 6.0};
    // d^n on polynomial coefficients (rows: power of u)
    static const double D[3][16] = {
        {0, 0, 0, 0, 1, 0, 0, 0, 0, 2, 0, 0, 0, 0, 3, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 2, 0, 0, 0, 0, 6, 0, 0},
        {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 6, 0, 0, 0}};
    auto mul = [] (double* m1, const double* m2, const double* m3) {
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 4; j++) {
                double acc = 0.0;
                for (int k = 0; k < 4; k++) acc += m2[4 * i + k] * m3[4 * k + j];
                m1[4 * i + j] = acc;
            }
    };
    plmhip::RegOp op {};
    double (*F[3])[16] = {op.X, op.Y, op.Z};
    for (int a = 0; a < 3; a++) {
        const double gs = (double) grid_spac[a];
        const double r = 1.0 / gs;
        const double R[16] = {1.0, 0, 0, 0, 0, r, 0, 0, 0, 0, r * r, 0, 0, 0, 0, r * r * r};
        double Q[4][16];
        mul (Q[0], B, R);
        for (int n = 1; n < 4; n++) mul (Q[n], Q[0], D[n - 1]);
        for (int n = 0; n < 4; n++) reg_integral (F[a][n], Q[n], Q[n], gs);
        reg_integral (F[a][4], Q[1], Q[0], gs);
        reg_integral (F[a][5], Q[0], Q[1], gs);
    }
    const float cp = rp.curvature_penalty, mix = rp.curvature_mixed_weight;
    const float m = rp.linear_elastic_multiplier, l1 = rp.lame_coefficient_1, l2 = rp.lame_coefficient_2;
    op.cp = cp;
    op.cpm = (float) (cp * mix);
    op.dp = rp.diffusion_penalty;
    op.tdp = rp.total_displacement_penalty;
    op.top = rp.third_order_penalty;
    op.e1 = (float) (l1 * m);
    op.e2 = (float) (l2 * m);
    op.ed = (float) ((m) * (2 * l1 + l2) / 2);
    op.eh = (float) ((m) * (l1) / 2);
    return op;
}

} // namespace

int
plm_hip_bspline_regularize (plm_hip_ctx* c, const plm_reg_parms* parms, const float* d_coeff,
    double* d_score, double* d_grad, int accumulate, void* stream)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (!parms || !d_coeff || !d_score || !d_grad) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    if (accumulate != 0 && accumulate != 1) return fail (PLM_HIP_ERR_ARG, "accumulate must be 0 or 1");
    HIP_TRY (hipSetDevice (c->device));
    hipStream_t st = stream ? (hipStream_t) stream : c->stream;
    const int rc = plmhip::ctx_reg_prepare (c);
    if (rc) return rc;
    const plmhip::RegOp op = reg_operator (c->grid_spac, *parms);
    hipError_t e = plmhip::launch_regularize (c->g, op, d_coeff, c->d_reg_part, c->d_reg_cond, d_score,
        d_grad, accumulate, st);
    if (e != hipSuccess) return hip_fail (e, "regularization launch");
    return PLM_HIP_OK;
}

} // extern "C"

namespace plmhip {

int
ctx_reg_prepare (plm_hip_ctx* c)
{
    if (c->d_reg_part) return PLM_HIP_OK;
    HIP_TRY (hipSetDevice (c->device));
    int rc = alloc (&c->d_reg_part, (size_t) c->num_tiles, "hipMalloc(reg partials)");
    if (rc) return rc;
    rc = alloc (&c->d_reg_cond, (size_t) 192 * c->num_knots, "hipMalloc(reg slots)");
    if (rc) return rc;
    // slots no region owns (knots near the grid's edge) stay zero for good
    HIP_TRY (hipMemset (c->d_reg_cond, 0, sizeof (double) * 192 * (size_t) c->num_knots));
    return PLM_HIP_OK;
}

int
ctx_regularize (plm_hip_ctx* c, const plm_reg_parms* parms, const float* d_coeff, double* d_score,
    double* d_grad, hipStream_t st, const int* skip)
{
    HIP_TRY (hipSetDevice (c->device));
    const int rc = ctx_reg_prepare (c);
    if (rc) return rc;
    const RegOp op = reg_operator (c->grid_spac, *parms);
    hipError_t e = launch_regularize (c->g, op, d_coeff, c->d_reg_part, c->d_reg_cond, d_score, d_grad, 1,
        st, skip);
    if (e != hipSuccess) return hip_fail (e, "regularization launch");
    return PLM_HIP_OK;
}

} // namespace plmhip

extern "C" {

int
plm_hip_bspline_warp (plm_hip_ctx* c, const float* d_coeff, float* d_out, float default_val,
    int linear_interp, void* stream)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (!d_coeff || !d_out) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    if (linear_interp != 0 && linear_interp != 1)
        return fail (PLM_HIP_ERR_ARG, "linear_interp must be 0 or 1");
    HIP_TRY (hipSetDevice (c->device));
    hipStream_t st = stream ? (hipStream_t) stream : c->stream;
    hipError_t e = plmhip::launch_warp (c->plan, c->tune, c->storage == PLM_STORE_BF16, d_coeff,
        d_out, default_val, linear_interp, st);
    if (e != hipSuccess) return hip_fail (e, "warp launch");
    return PLM_HIP_OK;
}

int
plm_hip_bspline_score_vf_device (plm_hip_ctx* c, const float* d_vf, double* d_score,
    double* d_grad, int64_t* d_num_vox, void* stream)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    if (!d_vf || !d_score || !d_grad || !d_num_vox) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    if (c->d_fmask || c->d_mroi)
        return fail (PLM_HIP_ERR_UNSUPPORTED, "the field-driven MSE takes no region of interest");
    HIP_TRY (hipSetDevice (c->device));
    hipStream_t st = stream ? (hipStream_t) stream : c->stream;
    hipEvent_t* ev = nullptr;
    HIP_TRY (c->timer.next (&ev));
    const plmhip::EvalOut out {d_grad, d_score, d_num_vox};
    hipError_t e = plmhip::launch_mse_from_vf (c->plan, c->tune, c->storage == PLM_STORE_BF16,
        d_vf, out, st, ev);
    if (e != hipSuccess) return hip_fail (e, "MSE from field launch");
    if (ev) c->timer.last_stages = plmhip::MSE_STAGES;
    return PLM_HIP_OK;
}

int
plm_hip_bspline_num_knots (const plm_hip_ctx* c, int64_t* n)
{
    if (!c || !n) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    *n = c->num_knots;
    return PLM_HIP_OK;
}

int
plm_hip_bspline_num_tiles (const plm_hip_ctx* c, int64_t* n)
{
    if (!c || !n) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    *n = c->num_tiles;
    return PLM_HIP_OK;
}

} // extern "C"

namespace plmhip {

int
abi_fail (int code, const std::string& msg)
{
    return fail (code, msg);
}

int
abi_hip_fail (hipError_t e, const char* what)
{
    return hip_fail (e, what);
}

int
abi_check_img (const plm_img_desc* d, const char* name)
{
    return check_img (d, name);
}

LbWorkspace**
ctx_lb_workspace (plm_hip_ctx* c)
{
    return &c->lb_ws;
}

int
ctx_info (const plm_hip_ctx* c, int* device, hipStream_t* stream, int64_t* num_knots)
{
    if (!c) return fail (PLM_HIP_ERR_ARG, "ctx is NULL");
    *device = c->device;
    *stream = c->stream;
    *num_knots = c->num_knots;
    return PLM_HIP_OK;
}

} // namespace plmhip

namespace {

void
drop_graph (plm_hip_batch* b)
{
    if (b->exec) (void) hipGraphExecDestroy (b->exec);
    if (b->graph) (void) hipGraphDestroy (b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void
free_batch (plm_hip_batch* b)
{
    drop_graph (b);
    b->timer.destroy ();
    void* bufs[] = {b->d_pairs, b->d_part_score, b->d_part_count, b->d_tile_ctr};
    for (void* p : bufs) if (p) (void) hipFree (p);
    if (b->stream) (void) hipStreamDestroy (b->stream);
    delete b;
}

int
check_batch_mode (const plm_hip_batch* b, int grad_mode)
{
    if (grad_mode != PLM_GRAD_REF_NN && grad_mode != PLM_GRAD_EXACT)
        return fail (PLM_HIP_ERR_ARG, "bad grad_mode");
    if (grad_mode == PLM_GRAD_REF_NN && !b->all_mgrad)
        return fail (PLM_HIP_ERR_ARG, "PLM_GRAD_REF_NN needs moving_grad in every context");
    return PLM_HIP_OK;
}

} // namespace

extern "C" {

int
plm_hip_batch_create (plm_hip_ctx* const* ctxs, int n, plm_hip_batch** out)
{
    if (!out) return fail (PLM_HIP_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (!ctxs || n < 1 || n > plmhip::MAX_BATCH)
        return fail (PLM_HIP_ERR_ARG, "a batch holds 1 to 64 contexts");
    for (int i = 0; i < n; i++) {
        const plm_hip_ctx* c = ctxs[i];
        if (!c) return fail (PLM_HIP_ERR_ARG, "NULL context in the batch");
        if (c->device != ctxs[0]->device)
            return fail (PLM_HIP_ERR_ARG, "batched contexts must share a device");
        if (c->storage != ctxs[0]->storage)
            return fail (PLM_HIP_ERR_ARG, "batched contexts must share the storage type");
        for (int a = 0; a < 3; a++)
            if (c->g.vpr[a] != ctxs[0]->g.vpr[a])
                return fail (PLM_HIP_ERR_ARG, "batched contexts must share vox_per_rgn");
        // each pair condenses into its context's own slot buffer, so a
        // context listed twice would have two pairs overwrite one buffer
        for (int j = 0; j < i; j++)
            if (ctxs[j] == c) return fail (PLM_HIP_ERR_ARG, "a context appears twice in the batch");
    }
    HIP_TRY (hipSetDevice (ctxs[0]->device));
    plm_hip_batch* b = new plm_hip_batch;
    b->device = ctxs[0]->device;
    b->storage = ctxs[0]->storage;
    b->ctxs.assign (ctxs, ctxs + n);
    b->tune = ctxs[0]->tune;
    hipError_t e = hipStreamCreate (&b->stream);
    if (e != hipSuccess) { free_batch (b); return hip_fail (e, "hipStreamCreate"); }
    // the flattened tile queue and the reduce workgroups, pair after pair;
    // until plm_hip_batch_set_buffers, each pair reads and writes its
    // context's own staging buffers
    int64_t tiles = 0, blocks = 0;
    b->h_pairs.resize (n);
    for (int i = 0; i < n; i++) {
        plm_hip_ctx* c = ctxs[i];
        plmhip::PairDev& pd = b->h_pairs[i];
        pd = pair_of (c);
        pd.coeff = c->d_coeff;
        pd.grad = c->d_grad;
        pd.score = c->d_score;
        pd.nvox = c->d_nvox;
        pd.tile0 = (int) tiles;
        pd.blk0 = (int) blocks;
        tiles += c->num_tiles;
        blocks += (3 * c->num_knots + plmhip::REDUCE_BLOCK - 1) / plmhip::REDUCE_BLOCK;
        b->all_mgrad = b->all_mgrad && c->have_mgrad;
    }
    if (tiles >= (1 << 30) / plmhip::MAX_BATCH || blocks >= (1 << 30)) {
        free_batch (b);
        return fail (PLM_HIP_ERR_GEOMETRY, "batch too large");
    }
#define BATCH_TRY(expr) do { int r_ = (expr); if (r_) { free_batch (b); return r_; } } while (0)
    BATCH_TRY (alloc (&b->d_pairs, (size_t) n, "hipMalloc(pairs)"));
    BATCH_TRY (alloc (&b->d_part_score, (size_t) n * (tiles + 8), "hipMalloc(partials)"));
    BATCH_TRY (alloc (&b->d_part_count, (size_t) n * (tiles + 8), "hipMalloc(partials)"));
    BATCH_TRY (alloc (&b->d_tile_ctr, plmhip::TILE_CTR_WORDS, "hipMalloc(tile counters)"));
#undef BATCH_TRY
    e = hipMemset (b->d_tile_ctr, 0, sizeof (unsigned long long) * plmhip::TILE_CTR_WORDS);
    if (e != hipSuccess) { free_batch (b); return hip_fail (e, "hipMemset(tile counters)"); }
    e = hipMemcpy (b->d_pairs, b->h_pairs.data (), sizeof (plmhip::PairDev) * n,
        hipMemcpyHostToDevice);
    if (e != hipSuccess) { free_batch (b); return hip_fail (e, "hipMemcpy(pairs)"); }
    plmhip::Plan& pl = b->plan;
    pl.d_pairs = b->d_pairs;
    pl.npairs = n;
    pl.g0 = ctxs[0]->g;
    pl.total_tiles = (int) tiles;
    pl.reduce_blocks = (int) blocks;
    pl.blut = ctxs[0]->d_blut;      // equal tile edges, equal LUTs
    pl.part_score = b->d_part_score;
    pl.part_count = b->d_part_count;
    pl.tile_ctr = b->d_tile_ctr;
    *out = b;
    return PLM_HIP_OK;
}

int
plm_hip_batch_set_buffers (plm_hip_batch* b, const float* const* d_coeff,
    double* const* d_grad, double* d_score, int64_t* d_num_vox)
{
    if (!b) return fail (PLM_HIP_ERR_ARG, "batch is NULL");
    if (!d_coeff || !d_grad || !d_score || !d_num_vox) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    const int n = (int) b->ctxs.size ();
    for (int i = 0; i < n; i++)
        if (!d_coeff[i] || !d_grad[i]) return fail (PLM_HIP_ERR_ARG, "NULL buffer pointer");
    HIP_TRY (hipSetDevice (b->device));
    HIP_TRY (hipStreamSynchronize (b->stream));
    for (int i = 0; i < n; i++) {
        plmhip::PairDev& pd = b->h_pairs[i];
        pd.coeff = d_coeff[i];
        pd.grad = d_grad[i];
        pd.score = d_score + i;
        pd.nvox = d_num_vox + i;
    }
    HIP_TRY (hipMemcpy (b->d_pairs, b->h_pairs.data (), sizeof (plmhip::PairDev) * n,
        hipMemcpyHostToDevice));
    drop_graph (b);
    b->buffers_set = true;
    return PLM_HIP_OK;
}

int
plm_hip_batch_score_device (plm_hip_batch* b, int grad_mode, void* stream)
{
    if (!b) return fail (PLM_HIP_ERR_ARG, "batch is NULL");
    int rc = check_batch_mode (b, grad_mode);
    if (rc) return rc;
    HIP_TRY (hipSetDevice (b->device));
    hipStream_t st = stream ? (hipStream_t) stream : b->stream;
    hipEvent_t* ev = nullptr;
    HIP_TRY (b->timer.next (&ev));
    const plmhip::EvalOut none {nullptr, nullptr, nullptr};
    hipError_t e = plmhip::launch_mse (b->plan, b->tune, grad_mode, b->storage == PLM_STORE_BF16,
        nullptr, none, st, ev);
    if (e != hipSuccess) return hip_fail (e, "batched MSE launch");
    if (ev) b->timer.last_stages = plmhip::MSE_STAGES;
    return PLM_HIP_OK;
}

int
plm_hip_batch_capture (plm_hip_batch* b, int grad_mode)
{
    if (!b) return fail (PLM_HIP_ERR_ARG, "batch is NULL");
    int rc = check_batch_mode (b, grad_mode);
    if (rc) return rc;
    HIP_TRY (hipSetDevice (b->device));
    HIP_TRY (hipStreamSynchronize (b->stream));
    drop_graph (b);
    // the evaluation has no memset and no host work: two kernel nodes
    HIP_TRY (hipStreamBeginCapture (b->stream, hipStreamCaptureModeThreadLocal));
    const plmhip::EvalOut none {nullptr, nullptr, nullptr};
    hipError_t e = plmhip::launch_mse (b->plan, b->tune, grad_mode, b->storage == PLM_STORE_BF16,
        nullptr, none, b->stream, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture (b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void) hipGraphDestroy (g);
        return hip_fail (e != hipSuccess ? e : e2, "graph capture");
    }
    b->graph = g;
    e = hipGraphInstantiate (&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) { drop_graph (b); return hip_fail (e, "hipGraphInstantiate"); }
    return PLM_HIP_OK;
}

int
plm_hip_batch_launch (plm_hip_batch* b, void* stream)
{
    if (!b) return fail (PLM_HIP_ERR_ARG, "batch is NULL");
    if (!b->exec) return fail (PLM_HIP_ERR_ARG, "no captured evaluation (plm_hip_batch_capture)");
    HIP_TRY (hipSetDevice (b->device));
    HIP_TRY (hipGraphLaunch (b->exec, stream ? (hipStream_t) stream : b->stream));
    return PLM_HIP_OK;
}

int
plm_hip_batch_set_tuning (plm_hip_batch* b, int vox_per_thread)
{
    if (!b) return fail (PLM_HIP_ERR_ARG, "batch is NULL");
    if (vox_per_thread != 1 && vox_per_thread != 2 && vox_per_thread != 4)
        return fail (PLM_HIP_ERR_ARG, "vox_per_thread must be 1, 2 or 4");
    if (b->tune.vox_per_thread != vox_per_thread) drop_graph (b);
    b->tune.vox_per_thread = vox_per_thread;
    return PLM_HIP_OK;
}

int
plm_hip_batch_set_occupancy (plm_hip_batch* b, int wg_per_cu)
{
    if (!b) return fail (PLM_HIP_ERR_ARG, "batch is NULL");
    if (wg_per_cu < 0) return fail (PLM_HIP_ERR_ARG, "workgroups per CU must be >= 0");
    if (b->tune.wg_per_cu != wg_per_cu) drop_graph (b);
    b->tune.wg_per_cu = wg_per_cu;
    return PLM_HIP_OK;
}

int
plm_hip_batch_get_tuning (plm_hip_batch* b, int* vox_per_thread, int* wg_per_cu)
{
    if (!b) return fail (PLM_HIP_ERR_ARG, "batch is NULL");
    if (vox_per_thread) *vox_per_thread = b->tune.vox_per_thread;
    if (wg_per_cu) *wg_per_cu = b->tune.wg_per_cu;
    return PLM_HIP_OK;
}

int
plm_hip_batch_set_timing (plm_hip_batch* b, int enable)
{
    if (!b) return fail (PLM_HIP_ERR_ARG, "batch is NULL");
    b->timer.on = enable != 0;
    return PLM_HIP_OK;
}

int
plm_hip_batch_get_timing (plm_hip_batch* b, double* stage_ms, int max_stages, int64_t* evals)
{
    if (!b || !stage_ms || !evals) return fail (PLM_HIP_ERR_ARG, "NULL argument");
    HIP_TRY (hipSetDevice (b->device));
    HIP_TRY (b->timer.collect (stage_ms, max_stages, evals));
    return PLM_HIP_OK;
}

int
plm_hip_batch_autotune (plm_hip_batch* b, int grad_mode, int reps, int* chosen, double* eval_ms)
{
    if (!b) return fail (PLM_HIP_ERR_ARG, "batch is NULL");
    int rc = check_batch_mode (b, grad_mode);
    if (rc) return rc;
    if (reps < 1) reps = 1;
    const bool was_timing = b->timer.on;
    double stage[plmhip::MAX_STAGE_EVENTS];
    int64_t n = 0;
    rc = plm_hip_batch_get_timing (b, stage, plmhip::MAX_STAGE_EVENTS, &n);
    if (rc) return rc;
    b->timer.on = true;
    const int cands[3] = {1, 2, 4};
    const int before = b->tune.vox_per_thread, before_wg = b->tune.wg_per_cu;
    int best = before, best_wg = before_wg;
    double best_ms = 1e300;
    for (int k = 0; k < 12; k++) {
        b->tune.vox_per_thread = cands[k % 3];
        b->tune.wg_per_cu = plmhip::TUNE_WG_CANDS[k / 3];
        for (int r = 0; r < reps + 1; r++) {
            rc = plm_hip_batch_score_device (b, grad_mode, b->stream);
            if (rc == PLM_HIP_OK && r == 0)     // the first call of each variant is a warm-up
                rc = plm_hip_batch_get_timing (b, stage, plmhip::MAX_STAGE_EVENTS, &n);
            if (rc) break;
        }
        if (rc == PLM_HIP_OK) rc = plm_hip_batch_get_timing (b, stage, plmhip::MAX_STAGE_EVENTS, &n);
        if (rc) {
            b->timer.on = was_timing;
            b->tune.vox_per_thread = before;
            b->tune.wg_per_cu = before_wg;
            return rc;
        }
        double total = 0.0;
        for (int i = 0; i < plmhip::MAX_STAGE_EVENTS; i++) total += stage[i];
        total /= (double) (n ? n : 1);
        if (total < best_ms) { best_ms = total; best = cands[k % 3]; best_wg = plmhip::TUNE_WG_CANDS[k / 3]; }
    }
    b->tune.vox_per_thread = best;
    b->tune.wg_per_cu = best_wg;
    if (best != before || best_wg != before_wg) drop_graph (b);
    b->timer.on = was_timing;
    if (chosen) *chosen = best;
    if (eval_ms) *eval_ms = best_ms;
    return PLM_HIP_OK;
}

int
plm_hip_batch_destroy (plm_hip_batch* b)
{
    if (!b) return PLM_HIP_OK;
    (void) hipSetDevice (b->device);
    (void) hipStreamSynchronize (b->stream);
    free_batch (b);
    return PLM_HIP_OK;
}

} // extern "C"
