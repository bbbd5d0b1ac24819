// Kernel-side geometry and launchers of libplmregisterhip.so (internal).
#ifndef PLM_BSPLINE_KERNELS_H
#define PLM_BSPLINE_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plmhip {

constexpr int TILE_THREADS = 256;
constexpr int MAX_BINS_1D = 256;     // per-axis histogram bins
constexpr int MAX_BINS_2D = 4096;    // joint bins held in LDS (32 KB int64)
constexpr int MAX_VPR = 32;          // tile edge bound
constexpr int MAX_BATCH = 64;        // pairs in one batched evaluation

// Everything a tile kernel needs, passed by value.
struct KGeom {
    int vpr[3];
    int rdims[3];
    int cdims[3];
    int roi_off[3];
    int roi_end[3];          // roi_offset + roi_dim
    int fdim[3];
    int mdim[3];
    float img_origin[3];     // grid origin (Bspline_header::img_origin)
    float fstep[9];          // fixed step = dc * spacing
    float morigin[3];
    float mproj[9];          // moving proj = inv(dc) / spacing
    float inv_msp[3];        // (float)(1.0 / moving spacing)
    float mstep[9];          // M = moving proj * fixed step (moving index per fixed index)
    float moff[3];           // moving proj * (img_origin - moving origin)
    float mhi[3];            // mdim - 0.5: moving index coordinates inside the image are below it
    float mmax[3];           // mdim - 2: the largest floor corner of the trilinear clamp
    int num_tiles;
    int num_knots;
    int tiles_per_xcd;
    uint32_t mag_vx;         // ceil(2^32 / vpr0), for n / vpr0 with __umulhi
    uint32_t mag_vy;         // ceil(2^32 / vpr1)
    uint32_t mag_vxy;        // ceil(2^32 / (vpr0*vpr1))
    int aligned;             // fixed step and moving proj are diagonal
};

// dc/dv slots of a tile: its voxel count rounded up to whole batches of
// TILE_THREADS * vox_per_thread.
__host__ __device__ inline int
tile_slots (int nvox, int vox_per_thread)
{
    const int s = TILE_THREADS * vox_per_thread;
    return (nvox + s - 1) / s * s;
}

// Floats of the tile kernel's column-term planes (three float4 planes of
// Horner coefficients per column), which the condense reuses for the
// tile's dc/dv contracted over x and y.
__host__ __device__ inline int
tile_col_floats (int vx, int vy, int vz)
{
    const int a = 12 * vx * vy;
    const int b = 48 * vz;
    return a > b ? a : b;
}

// Floats of the tile kernel's dc/dv planes (3 x slots), which the first
// displacement stage (64 floats per x) uses before any dc/dv is written.
__host__ __device__ inline int
tile_dv_floats (int vx, int slots)
{
    const int a = 3 * slots;
    const int b = 64 * vx;
    return a > b ? a : b;
}

struct Mat3 {
    float m[9];
};

struct HGeom {
    int fbins;
    int mbins;
    float f_offset, f_delta;
    float m_offset, m_delta;
};

// Launch-shape choices made by the autotuner.
struct Tune {
    int vox_per_thread;      // voxels each tile-kernel thread keeps in flight: 1, 2 or 4
    size_t lds_limit;        // LDS bytes a workgroup may use
    int num_cus;             // compute units of the device (persistent grid size)
    int wg_per_cu = 0;       // tile-kernel workgroups per CU (0: as many as fit)
};

// Workgroups-per-CU candidates of the autotuners (0: as many as fit).  Fewer
// than fit can be faster: the tiles in flight on a CU share its L1.
constexpr int TUNE_WG_CANDS[4] = {0, 6, 5, 4};

// LDS bytes of the tile kernel at vox_per_thread voxels per batch, static
// arrays included (the MI table comes on top).
size_t tile_lds_bytes (const KGeom& g, int vox_per_thread);

// Moving-index frame of one tile.  The tile's first fixed voxel maps to
// moving index coordinates B = moff + M (bx0, by0, bz0) (formed in double on
// the host); with the integer part Bi = clamp(floor(B), 0, mdim - 1) taken
// out, the tile kernel works in the local coordinate ml = m - Bi, which stays
// a few voxels in size wherever the tile lands inside the moving image, so
// its float rounding is that of a small number instead of one as large as
// the image (at 512 voxels a float position is good to 3e-5 voxel, the local
// one to ~1e-6).  Every threshold of the inside test and the clamp moves into
// the same frame exactly (integers and half-integers are exact floats).
struct TileBase {
    float bf[3];             // B - Bi, rounded to float
    float thr_lo[3];         // -0.5 - Bi: inside iff thr_lo < ml < thr_hi (Volume::is_inside)
    float thr_hi[3];         // mdim - 0.5 - Bi
    float cl_lo[3];          // -Bi: the clamped floor corner lies in [cl_lo, cl_hi] (li_clamp)
    float cl_hi[3];          // mdim - 2 - Bi
    int boff;                // element offset of voxel Bi in the moving image
};

// The frame of tile t: the base B of the tile's first voxel in double, its
// integer part clamped to the moving image (so every local floor stays
// within +-mdim of it, and B - Bi is small wherever the tile lands inside
// the image), and the inside-test and clamp bounds moved into that frame.
// Host and device (the allocation-free entry point builds the table on the
// device) run the same double arithmetic.
__host__ __device__ inline TileBase
tile_base_of (const KGeom& g, const double moff[3], const double mstep[9], int t)
{
    const int p0 = t % g.rdims[0], p1 = (t / g.rdims[0]) % g.rdims[1];
    const int p2 = t / (g.rdims[0] * g.rdims[1]);
    const double f0 = (double) (g.roi_off[0] + g.vpr[0] * p0);
    const double f1 = (double) (g.roi_off[1] + g.vpr[1] * p1);
    const double f2 = (double) (g.roi_off[2] + g.vpr[2] * p2);
    TileBase tb;
    int64_t bi[3];
    for (int a = 0; a < 3; a++) {
        const double B = moff[a] + mstep[3*a] * f0 + mstep[3*a+1] * f1 + mstep[3*a+2] * f2;
        const bool finite = (B - B) == 0.0;
        double fl = finite ? floor (B) : 0.0;
        const double top = (double) (g.mdim[a] - 1);
        fl = fl < 0.0 ? 0.0 : (fl > top ? top : fl);
        bi[a] = (int64_t) fl;
        tb.bf[a] = (float) (finite ? B - fl : 0.0);
        tb.thr_lo[a] = (float) (-0.5 - fl);
        tb.thr_hi[a] = (float) ((double) g.mdim[a] - 0.5 - fl);
        tb.cl_lo[a] = (float) -fl;
        tb.cl_hi[a] = (float) ((double) (g.mdim[a] - 2) - fl);
    }
    tb.boff = (int) (bi[0] + (int64_t) g.mdim[0] * (bi[1] + (int64_t) g.mdim[1] * bi[2]));
    return tb;
}

// A region-of-interest volume (one byte a voxel, a geometry of its own) as
// inside_roi tests it (src/util/bspline_correspond.cxx:52-74): the byte at
// the floor of proj (xyz - origin), each index clamped into the volume.
struct RoiVol {
    const uint8_t* img;      // null: no ROI
    int dim[3];
    float origin[3];
    float proj[9];           // inv(dc) / spacing, formed as for the images
    // The tile kernels' map from the moving image's index coordinates to
    // the ROI's: roi = R mijk + r0 (a moving ROI only; double on the host).
    float R[9];
    double Rd[9];
    double r0[3];
};

// One fixed/moving pair as the tile and reduce kernels see it (a device
// table of these drives a launch; a single context is a table of one).
struct PairDev {
    KGeom g;
    const TileBase* tbase;   // per tile of the pair
    const uint8_t* fmask;    // the fixed ROI's byte at every fixed voxel (null: none)
    RoiVol mroi;             // the moving ROI (img null: none)
    const void* fimg;        // fixed, moving, moving gradient (the batch's storage type)
    const void* mimg;
    const void* mgrad;
    const void* mpair;       // the moving image as z-pairs (the batch's storage type)
    const float* coeff;      // coefficients of the evaluation (batched launches)
    float* cond;             // slot-major condense buffer, 64 * 3 * num_knots
    double* grad;            // outputs (batched launches): 3 * num_knots
    double* score;
    int64_t* nvox;
    int tile0;               // first tile of the pair in the flattened tile queue
    int blk0;                // first reduce workgroup of the pair
};

// The tile kernel's per-XCD tile counters: XCD x's 64-bit word at
// ctr[TILE_CTR_STRIDE * x] (each on a 128-byte line of its own), zero
// between launches (every launch leaves them zero).  One set per plan;
// launches that share a set must not overlap.
constexpr int TILE_CTR_STRIDE = 16;
constexpr int TILE_CTR_WORDS = 8 * TILE_CTR_STRIDE;

struct TileQueue {
    int total_tiles;
    int tiles_per_xcd;       // ceil(total_tiles / 8)
    int npairs;
    int pad;
    unsigned long long* ctr; // TILE_CTR_WORDS counters (launches with counted tiles)
};

// A batched launch or MI pass 2 takes its tiles from the counters when its
// workgroups would otherwise walk at least this many tiles each at a
// fixed stride.
constexpr int TILE_DYNAMIC_ROUNDS = 12;

// Per-evaluation outputs of pair 0 that override its table entry (the
// single-pair entry points pass them with every call); null fields fall
// back to the table.
struct EvalOut {
    double* grad;
    double* score;
    int64_t* nvox;
    // when not null and nonzero at launch time, every kernel of the
    // evaluation returns at once (an optimizer chain enqueued after its run
    // ended)
    const int* skip = nullptr;
};

// Everything a launch needs about its set of pairs.
struct Plan {
    const PairDev* d_pairs;  // device table
    int npairs;
    KGeom g0;                // pair 0's geometry (every pair shares the tile edges)
    int total_tiles;
    int reduce_blocks;       // sum over pairs of ceil(3 * num_knots / REDUCE_BLOCK)
    const float* blut;       // basis LUTs of the shared tile edges
    double* part_score;      // per tile of the flattened queue (score kinds)
    int* part_count;         // per (pair, workgroup) overlap counts, npairs * (total_tiles + 8)
    unsigned long long* tile_ctr;   // TILE_CTR_WORDS, zeroed once at allocation
};

constexpr int REDUCE_BLOCK = 64;     // gradient components per reduce workgroup

// Stage timing: when ev is not null, ev[0..n] are recorded on st between
// the launches (MSE: tiles | reduce, 3 events; MI: zero | hist | score |
// tiles | reduce, 6 events).
constexpr int MSE_STAGES = 2;
constexpr int MI_STAGES = 5;
constexpr int MAX_STAGE_EVENTS = 6;

// What an MSE evaluation computes.
constexpr int MSE_REF_NN = 0;        // score + gradient, nearest-voxel moving gradient
constexpr int MSE_EXACT = 1;         // score + gradient, analytic trilinear derivative
constexpr int MSE_SCORE_ONLY = 2;    // score and overlap count (single-pair plans)

// A launch that takes the place of a single-pair evaluation's gradient
// reduce (the optimizer's chain folds its post pass into it): called with
// the tile kernel's partial count and whether it wrote per-tile scores.
struct ReduceHook {
    hipError_t (*launch) (void* user, const Plan& plan, int nparts, int tile_scores, const EvalOut& out0,
        hipStream_t st);
    void* user;
};

// MSE over every pair of the plan (coeff0/out0 override pair 0's table
// entry when not null); with hook (gradient modes only) the hook launches
// in place of k_reduce.
hipError_t launch_mse (const Plan& plan, const Tune& tune, int mode, int storage,
    const float* coeff0, const EvalOut& out0, hipStream_t st, hipEvent_t* ev,
    const ReduceHook* hook = nullptr);

// The same for a single-pair float plan with no z-pair image: the tile
// kernel reads the trilinear corners from the moving image itself (the
// allocation-free entry point, which has no z-pair copy), at
// vox_per_thread 2.  No stage timing.
hipError_t launch_mse_plain (const Plan& plan, const Tune& tune, int mode, const float* coeff0,
    const EvalOut& out0, hipStream_t st);

// Dense displacement field of a single-pair plan: float3 (mm) per fixed
// voxel, the ROI filled and the rest zero.
hipError_t launch_vf (const Plan& plan, const Tune& tune, const float* coeff0, float* vf,
    hipStream_t st);

// MSE score and exact-derivative gradient of a single-pair plan from a
// stored displacement field instead of the coefficients.
hipError_t launch_mse_from_vf (const Plan& plan, const Tune& tune, int storage, const float* vf,
    const EvalOut& out0, hipStream_t st, hipEvent_t* ev);

// Warped moving image of a single-pair plan (trilinear if linear, else the
// nearest voxel): one float per fixed voxel, fill where the moving position
// leaves the moving image and outside the ROI.
hipError_t launch_warp (const Plan& plan, const Tune& tune, int storage, const float* coeff0,
    float* out, float fill, int linear, hipStream_t st);
// Analytic regularization operator: the 4x4 region-integral matrices of
// each axis (index 0..3: both factors of derivative order n; 4: orders
// (1, 0); 5: orders (0, 1)) and the term weights, formed on the host as
// Bspline_regularize::analytic_init does.
struct RegOp {
    double X[6][16];         // grid spacing 0: acts on the knot index's z digit
    double Y[6][16];
    double Z[6][16];         // grid spacing 2: acts on the x digit
    double cp, cpm, dp, tdp, top;   // curvature, mixed curvature, diffusion, displacement, 3rd order
    double e1, e2;           // elastic cross terms: lame_1 * m, lame_2 * m
    double ed, eh;           // elastic diagonal: m (2 lame_1 + lame_2) / 2, m lame_1 / 2
};
// Regularization score and gradient of coefficients coeff over g's grid:
// part (num_tiles doubles) and cond (64 * 3 * num_knots doubles) are
// scratch; score = the score, grad = (accumulate ? grad : 0) + gradient.
hipError_t launch_regularize (const KGeom& g, const RegOp& op, const float* coeff, double* part,
    double* cond, double* score, double* grad, int accumulate, hipStream_t st,
    const int* skip = nullptr);
// Mattes MI of a single-pair plan.  ptab: the fixed positions of pass 1's
// walk (launch_pass1_positions), or null where they equal the direct ones.
// fmask1: the fixed ROI at pass 1's positions (null: none); mroi: the
// moving ROI, tested in pass 1 at the reference's float moving position.
hipError_t launch_mi (const Plan& plan, const Tune& tune, const HGeom& h, int hist_blocks,
    const uint8_t* fbin, const uint32_t* mquad, const float* coeff, unsigned long long* jhist,
    unsigned long long* fhist, float* table, float* nvox_f, const EvalOut& out0,
    int64_t* hist_out, const float* ptab, const uint8_t* fmask1, const RoiVol& mroi,
    hipStream_t st, hipEvent_t* ev);
// The fixed ROI r at every voxel of g's fixed image: out[v] = 1 where
// inside_roi holds at the voxel's position (ptab's where given, else
// origin + step ijk), else 0.
hipError_t launch_roi_fixed (const KGeom& g, const RoiVol& r, const float* ptab, uint8_t* out,
    hipStream_t st);
// The reference's pass-1 walk of the fixed image (LOOP_Z/Y/X): adds to
// *mismatch the voxels whose walk position differs from origin + step ijk,
// and with out (float3 per fixed voxel) not null writes the walk positions.
hipError_t launch_pass1_positions (const KGeom& g, float* out, unsigned int* mismatch,
    hipStream_t st);

// Bin image of an image (storage 0 float, 1 BF16): the equal-spacing bin
// floor((v - offset) / delta) clamped to bins, or with lut (device, keys
// entries) the bin lut[key] of that key clamped to keys.
hipError_t launch_bin_image (const void* img, int storage, uint8_t* out, int64_t n,
    float offset, float delta, int bins, hipStream_t st, const int* lut = nullptr, int keys = 0);

hipError_t launch_zero (float* p, int64_t n, hipStream_t st);
// The moving bin-quad image of a bin image of dim voxels: record v holds the
// bins of v, v + 1, v + dim0, v + dim0 + 1 (low byte first).
hipError_t launch_bin_quads (const uint8_t* bins, uint32_t* out, const int dim[3], hipStream_t st);
// The moving image as z-pairs: out[2v] = m[v], out[2v + 1] = m[v + d01]
// (0 on the last slice), so the two records at voxel v hold the trilinear
// corners (x, z), (x, z + 1), (x + 1, z), (x + 1, z + 1) of one row: 16
// bytes for float storage (storage 0), 8 for BF16 (storage 1).
hipError_t launch_zpair (const void* m, int storage, void* out, int64_t n, int64_t d01,
    hipStream_t st);
hipError_t launch_volume_gradient (const float* img, float* out, const int dim[3],
    const float sp[3], const Mat3& inv, hipStream_t st);
hipError_t launch_to_bf16 (const float* src, uint16_t* dst, int64_t n, hipStream_t st);

} // namespace plmhip

#endif
