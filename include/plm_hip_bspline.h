/* plm_hip_bspline.h -- C ABI of libplmregisterhip.so, the MI355X (gfx950)
 * B-spline score/gradient flavor ("threading=hip").
 *
 * The library replaces the compiled-out CUDA plugin ABI
 * (src/register/cuda/bspline_cuda.h:123-336, loaded through
 * LOAD_LIBRARY_SAFE / LOAD_SYMBOL, src/sys/cuda_delayload.h:26-59) with
 * plain-old-data arguments and integer status codes.  Nothing in these
 * signatures is a C++ class or a HIP type; streams travel as void*.
 *
 * Conventions, as in the reference:
 *   - volumes are x-fastest (index i + dim0*(j + dim1*k),
 *     src/base/volume_macros.h:13-23);
 *   - coefficients and gradients are interleaved, coeff[3*knot+d], with
 *     knot = (k*cdims1 + j)*cdims0 + i (src/base/bspline_interpolate.cxx:37-43);
 *   - the moving-image gradient is float3 interleaved per voxel
 *     (src/base/volume_grad.cxx:136-145).
 *
 * Every function returns PLM_HIP_OK (0) on success, a positive warning, or
 * a negative error code; plm_hip_last_error() describes the last failure of
 * the calling thread.  No function exits the process.
 */
#ifndef PLM_HIP_BSPLINE_H
#define PLM_HIP_BSPLINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PLM_HIP_ABI_VERSION 1

/* status codes */
#define PLM_HIP_OK                 0
#define PLM_HIP_WARN_NO_OVERLAP    1   /* fewer than 1 voxel overlaps: score
                                          FLT_MAX, gradient 0, as
                                          bspline_score_normalize does
                                          (src/register/bspline_mse.cxx:53-65) */
#define PLM_HIP_ERR_ARG           -1
#define PLM_HIP_ERR_HIP           -2
#define PLM_HIP_ERR_GEOMETRY      -3
#define PLM_HIP_ERR_UNSUPPORTED   -4
#define PLM_HIP_ERR_NO_DEVICE     -5

/* metric (src/register/similarity_metric_type.h) */
#define PLM_METRIC_MSE 0
#define PLM_METRIC_MI  1

/* gradient semantics of the MSE dc/dv:
 *   REF_NN -- diff * m_grad[round(mijk)], flavor g
 *             (src/register/bspline_mse.cxx:679-681); needs moving_grad;
 *   EXACT  -- diff * analytic trilinear derivative, flavor p
 *             (src/register/bspline_mse.cxx:1316-1351); the exact adjoint. */
#define PLM_GRAD_REF_NN 0
#define PLM_GRAD_EXACT  1
/* MSE score and overlap count only, no gradient (single contexts): the
 * forward pass of the differentiable binding. */
#define PLM_GRAD_NONE   2

/* where a pointer lives */
#define PLM_MEM_HOST   0
#define PLM_MEM_DEVICE 1

/* how the library keeps an image in HBM (arithmetic is always FP32) */
#define PLM_STORE_F32  0
#define PLM_STORE_BF16 1
#define PLM_STORE_U8   2   /* region-of-interest volumes: one byte a voxel, non-zero inside */

/* B-spline grid: the arguments of Bspline_header::set
 * (src/base/bspline_header.cxx:28-62). rdims/cdims are derived. */
typedef struct plm_bsp_geom {
    float img_origin[3];
    float img_spacing[3];
    int64_t img_dim[3];
    int64_t roi_offset[3];
    int64_t roi_dim[3];
    int64_t vox_per_rgn[3];
    float direction_cosines[9];
} plm_bsp_geom;

/* One image: the Volume fields the path reads (src/base/volume.h:39-57).
 * data is float32 (float3 for moving_grad).  step/proj are derived from
 * direction_cosines and spacing as compute_direction_matrices does
 * (src/base/direction_matrices.cxx:10-20). */
typedef struct plm_img_desc {
    const void* data;
    int32_t location;          /* PLM_MEM_HOST or PLM_MEM_DEVICE */
    int32_t storage;           /* PLM_STORE_F32 or PLM_STORE_BF16 */
    int64_t dim[3];
    float origin[3];
    float spacing[3];
    float direction_cosines[9];
} plm_img_desc;

/* Equal-spacing joint histogram (src/register/histogram.h:19-35,
 * src/register/joint_histogram.cxx:115-139). */
typedef struct plm_hist_desc {
    int64_t fixed_bins;
    int64_t moving_bins;
    float fixed_offset;
    float fixed_delta;
    float moving_offset;
    float moving_delta;
} plm_hist_desc;

/* Regularization weights: the fields of Regularization_parms
 * (src/register/regularization_parms.h:20-28) the analytic flavors read. */
typedef struct plm_reg_parms {
    float total_displacement_penalty;
    float diffusion_penalty;
    float curvature_penalty;
    float lame_coefficient_1;
    float lame_coefficient_2;
    float linear_elastic_multiplier;
    float third_order_penalty;
    float curvature_mixed_weight;
} plm_reg_parms;

typedef struct plm_hip_ctx plm_hip_ctx;

int plm_hip_abi_version (void);
const char* plm_hip_last_error (void);

/* Replaces CUDA_bspline_mse_init_j / CUDA_bspline_mi_init_a
 * (src/register/cuda/bspline_cuda.h:123-336, bspline_cuda.cu:388-682) and
 * the per-eval bspline_cuda_state_create
 * (src/register/bspline_state.cxx:165-237): uploads the images once per
 * stage, builds the basis LUTs and the per-tile work buffers.
 * moving_grad may be NULL (then only PLM_GRAD_EXACT and MI are
 * available); hist may be NULL (then MI is unavailable). */
int plm_hip_bspline_create (
    const plm_bsp_geom* geom,
    const plm_img_desc* fixed,
    const plm_img_desc* moving,
    const plm_img_desc* moving_grad,
    const plm_hist_desc* hist,
    int gpuid,
    plm_hip_ctx** out);

/* Replaces CUDA_bspline_mse_j / CUDA_bspline_mi_a
 * (src/register/cuda/bspline_cuda.cxx:607-801), i.e. one evaluation of
 * bspline_score_g_mse (src/register/bspline_mse.cxx:519-720),
 * bspline_score_p_mse (:1149-1384) or bspline_score_g_mi
 * (src/register/bspline_mi.cxx:1505-1737).  Host pointers; blocks until
 * done.  grad has 3*num_knots doubles (already normalized: 2/N for MSE).
 * hist_out, if not NULL, receives Bf*Bm joint bins (units of 2^-32 voxel)
 * followed by Bf fixed-bin counts (MI only). */
int plm_hip_bspline_score (
    plm_hip_ctx* ctx,
    int metric,
    int grad_mode,
    const float* coeff,
    double* score,
    double* grad,
    int64_t* num_vox,
    int64_t* hist_out);

/* The same evaluation with every pointer in device memory, enqueued on
 * stream (a hipStream_t, or NULL for the context's own stream); returns
 * without waiting.  No counterpart in the reference: it is what an
 * optimizer loop that keeps coefficients and gradient resident calls.
 * The evaluations of one context share its intermediate buffers: they must
 * not overlap (enqueue them on one stream, or order the streams); those of
 * different contexts may. */
int plm_hip_bspline_score_device (
    plm_hip_ctx* ctx,
    int metric,
    int grad_mode,
    const float* d_coeff,
    double* d_score,
    double* d_grad,
    int64_t* d_num_vox,
    int64_t* d_hist,
    void* stream);

/* Scratch that the allocation-free entry point below reserves on each GPU
 * it runs on, for the whole life of the process. */
#define PLM_HIP_NOALLOC_SCRATCH_BYTES (32u << 20)

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  One MSE evaluation -- what plm_hip_bspline_score_device
 * computes for PLM_METRIC_MSE, the same gradient bits and overlap count --
 * without a context: fixed, moving and (for PLM_GRAD_REF_NN) moving_grad
 * are float32 images in device memory (PLM_MEM_DEVICE), read in place;
 * d_coeff, d_score, d_grad (may be NULL for PLM_GRAD_NONE) and d_num_vox are
 * device buffers.  Enqueued on stream (a hipStream_t; NULL for the null
 * stream) on device gpuid, which it makes current; returns without waiting.
 * Calls no hipMalloc, hipFree or allocating routine: every intermediate
 * lives in a static device array of PLM_HIP_NOALLOC_SCRATCH_BYTES per GPU
 * (192 floats a knot, 76 bytes a tile and 26 KB besides; a grid that needs
 * more returns PLM_HIP_ERR_UNSUPPORTED, e.g. above ~40000 knots).  The first
 * call on a device looks up that array's address (make it outside a graph
 * capture).  Every call on a GPU shares the array and the next call
 * overwrites it: calls that overlap on one GPU (two streams or threads)
 * race. */
int plm_hip_bspline_score_noalloc (
    const plm_bsp_geom* geom,
    const plm_img_desc* fixed,
    const plm_img_desc* moving,
    const plm_img_desc* moving_grad,
    int grad_mode,
    int gpuid,
    const float* d_coeff,
    double* d_score,
    double* d_grad,
    int64_t* d_num_vox,
    void* stream);

/* Regions of interest of a context: the fixed and moving ROI volumes of
 * Bspline_state (fixed_roi / moving_roi, src/register/bspline_state.h:44-45,
 * set by Bspline_stage, src/register/bspline_stage.cxx:129-205), uchar
 * images (storage PLM_STORE_U8) with geometries of their own, on the host
 * or the device; either may be NULL, and NULL for both clears them.  With a
 * ROI set, MSE follows flavor i (src/register/bspline_mse.cxx:82-290) for
 * PLM_GRAD_REF_NN and PLM_GRAD_NONE and flavor r (:1628-1862) for
 * PLM_GRAD_EXACT: a fixed voxel counts only where the fixed ROI holds at
 * its position and the moving ROI at its moving position (inside_roi,
 * src/util/bspline_correspond.cxx:52-74, 113-145).  MI follows flavor h
 * (src/register/bspline_mi.cxx:1232-1496): pass 1 tests both ROIs at its
 * walk positions, pass 2 at the tiles' positions.  The histogram ranges
 * stay what the context was created with (Joint_histogram::initialize
 * spans the whole images whatever the ROIs).  No evaluation of the context
 * may be in flight; a batch takes its contexts' ROIs when it is created.
 * plm_hip_bspline_score_vf_device refuses a context with a ROI. */
int plm_hip_bspline_set_roi (plm_hip_ctx* ctx, const plm_img_desc* fixed_roi,
    const plm_img_desc* moving_roi);

/* V-optimal bins for a context's MI (Histogram::type HIST_VOPT and its
 * key_lut, src/register/histogram.h:14-35): fixed_key_lut / moving_key_lut
 * (PLM_HIP_VOPT_KEYS ints each, host memory, as plm_hip_hist_vopt writes
 * them; NULL keeps that side's equal-spacing bins) map the key
 * floor((v - offset) / delta) of a voxel to its bin, as add_pvi_8 and the
 * pass-2 derivative do (src/register/joint_histogram.cxx:458-481,
 * bspline_mi.txx:82-94); offset and delta are the hist the context was
 * created with, which must then be the key grid's (plm_hip_hist_vopt's).
 * Every entry must lie in [0, bins - 1].  The bin images are rebuilt; no
 * evaluation of the context may be in flight. */
int plm_hip_bspline_set_hist_keys (plm_hip_ctx* ctx, const int32_t* fixed_key_lut,
    const int32_t* moving_key_lut);

/* Replaces CUDA_bspline_mse_cleanup_j / CUDA_bspline_mi_cleanup_a and
 * bspline_cuda_state_destroy (src/register/bspline_state.cxx:239-271). */
int plm_hip_bspline_destroy (plm_hip_ctx* ctx);

/* Moving-image gradient for PLM_GRAD_REF_NN, computed on the GPU:
 * central differences with edge replicate, rotated by inv(dc), float3
 * interleaved -- the volume_gradient the stage builds once per stage
 * (src/base/volume_grad.cxx:9-76,136-145, called at
 * src/register/registration_data.cxx:383).  vol->data and out may be host
 * or device pointers (out_location says which); out holds 3*npix floats. */
int plm_hip_volume_gradient (
    const plm_img_desc* vol,
    float* out,
    int out_location,
    int gpuid);

/* Per-stage kernel timing with HIP events recorded on the launch stream,
 * the counterpart of the CUDA flavor's PROFILE_MSE stage timers
 * (src/register/cuda/bspline_cuda.h:28, bspline_cuda.cu:1371-1436).
 * While enabled, every evaluation records its stage boundaries;
 * get_timing waits for them and returns the summed milliseconds per stage
 * (MSE: tiles, reduce; MI: zero, histogram, score, tiles, reduce) and the
 * number of evaluations, then starts a new window.  One
 * window should hold one metric. */
int plm_hip_bspline_set_timing (plm_hip_ctx* ctx, int enable);
int plm_hip_bspline_get_timing (plm_hip_ctx* ctx, double* stage_ms, int max_stages,
    int64_t* evals);

/* Launch-shape tuning of the tile kernel: how many voxels each thread keeps
 * in flight (1, 2 or 4) and how many of its persistent workgroups share a
 * compute unit (set_occupancy; 0 = as many as fit, the default).  autotune
 * times every pair of the two (vox_per_thread 1, 2, 4 by workgroups per CU
 * as many as fit, 6, 5, 4) on this context's own images (reps evaluations
 * each, after one warm-up, HIP events on the context stream), keeps the
 * fastest whole evaluation and reports it; get_tuning reads the current
 * shape.  d_coeff is a device coefficient vector or NULL (zeros).  The
 * counterpart of the reference's launch-configuration search
 * (CUDA_exec_conf_1tpe / _1bpe, src/cuda/cuda_kernel_util.cu:48-188;
 * divideWork, src/opencl/autotune_opencl.cxx:7-53). */
int plm_hip_bspline_set_tuning (plm_hip_ctx* ctx, int vox_per_thread);
int plm_hip_bspline_set_occupancy (plm_hip_ctx* ctx, int workgroups_per_cu);
int plm_hip_bspline_get_tuning (plm_hip_ctx* ctx, int* vox_per_thread, int* workgroups_per_cu);
int plm_hip_bspline_autotune (plm_hip_ctx* ctx, int metric, int grad_mode,
    const float* d_coeff, int reps, int* chosen_vox_per_thread, double* eval_ms);

/* Dense displacement field of coefficients d_coeff, float3 (mm)
 * interleaved per fixed voxel: the ROI voxels are filled, the rest set to
 * zero -- bspline_interpolate_vf (src/base/bspline_interpolate.cxx:112-142),
 * with device pointers, enqueued on stream (NULL: the context's). */
int plm_hip_bspline_interpolate_vf (plm_hip_ctx* ctx, const float* d_coeff, float* d_vf,
    void* stream);

/* Analytic smoothness regularization of coefficients d_coeff over the
 * context's B-spline grid: *d_score = the regularization score (double)
 * and d_grad (3 * num_knots doubles) = its gradient, added to what d_grad
 * holds when accumulate is 1 (the reference adds it to the similarity
 * gradient).  Replaces Bspline_regularize::compute_score_analytic_omp,
 * flavor c (src/register/bspline_regularize_analytic.cxx:943-1007, with the
 * matrices of analytic_init :638-940), called from bspline_score
 * (src/register/bspline.cxx:619-622).  Device pointers, enqueued on stream
 * (NULL: the context's); the first call allocates the context's
 * regularization scratch. */
int plm_hip_bspline_regularize (plm_hip_ctx* ctx, const plm_reg_parms* parms,
    const float* d_coeff, double* d_score, double* d_grad, int accumulate, void* stream);

/* Warped moving image of coefficients d_coeff: one float per fixed voxel
 * (fixed-image dims, x fastest) into d_out, the moving image sampled where
 * the fixed voxel lands -- trilinear (linear_interp = 1) or the nearest
 * voxel (0) -- and default_val where that position leaves the moving image
 * and on voxels outside the ROI.  Replaces bspline_warp
 * (src/base/bspline_warp.cxx:171-214; bspline_warp_dcos :27-167) for float
 * images; the moving image is read at the context's storage precision.
 * Device pointers, enqueued on stream (NULL: the context's). */
int plm_hip_bspline_warp (plm_hip_ctx* ctx, const float* d_coeff, float* d_out,
    float default_val, int linear_interp, void* stream);

/* MSE score and exact-derivative gradient (as PLM_GRAD_EXACT) taking the
 * displacement from a stored field d_vf (as plm_hip_bspline_interpolate_vf
 * writes it) instead of interpolating the coefficients: the stored-field
 * form of the backward pass.  No counterpart in the reference, whose
 * flavors always interpolate (bspline_interp_pix_c,
 * src/base/bspline_interpolate.cxx:78-110). */
int plm_hip_bspline_score_vf_device (plm_hip_ctx* ctx, const float* d_vf, double* d_score,
    double* d_grad, int64_t* d_num_vox, void* stream);

/* Geometry queries, for the host adapter. */
int plm_hip_bspline_num_knots (const plm_hip_ctx* ctx, int64_t* num_knots);
int plm_hip_bspline_num_tiles (const plm_hip_ctx* ctx, int64_t* num_tiles);

/* ------------------------------------------------------------------------
 * Batched MSE evaluation of many stages at once (the multi-atlas case).
 *
 * A batch strings together up to 64 contexts (same device, same storage,
 * equal vox_per_rgn) and evaluates all of them with one persistent
 * tile-kernel launch over the flattened queue of their tiles, plus one
 * reduce launch; the evaluation can be captured once in a hipGraph and
 * replayed every optimizer iteration.  The reference has no batched form:
 * its flavors score one pair per call (bspline_score,
 * src/register/bspline.cxx:583-601), so these entry points have no
 * counterpart to replace; per pair the results are those of
 * plm_hip_bspline_score_device (the same gradient bits, the score to
 * summation order).
 *
 * The batch borrows each context's images and condense buffer: the
 * contexts must outlive the batch, and a context must not be evaluated on
 * its own while a batched evaluation that includes it is in flight. */
typedef struct plm_hip_batch plm_hip_batch;

int plm_hip_batch_create (plm_hip_ctx* const* ctxs, int n, plm_hip_batch** out);

/* Device buffers of the evaluation: d_coeff[i] and d_grad[i] are host
 * arrays of n device pointers (pair i's 3*num_knots coefficients and FP64
 * gradient), d_score and d_num_vox device arrays of n.  Kept in the batch's
 * device table (this call synchronizes); a captured graph refers to them,
 * so setting new buffers drops the graph. */
int plm_hip_batch_set_buffers (plm_hip_batch* b, const float* const* d_coeff,
    double* const* d_grad, double* d_score, int64_t* d_num_vox);

/* One batched evaluation (grad_mode as plm_hip_bspline_score), enqueued
 * on stream (NULL: the batch's own stream). */
int plm_hip_batch_score_device (plm_hip_batch* b, int grad_mode, void* stream);

/* Capture one evaluation into a hipGraph (on the batch's stream) and
 * instantiate it; plm_hip_batch_launch replays it on stream.  Changing the
 * buffers or the tuning (vox_per_thread or workgroups per CU) drops the
 * graph (launch then returns
 * PLM_HIP_ERR_ARG until the next capture). */
int plm_hip_batch_capture (plm_hip_batch* b, int grad_mode);
int plm_hip_batch_launch (plm_hip_batch* b, void* stream);

/* Tuning and stage timing of the batched tile kernel, as
 * plm_hip_bspline_set_tuning / _autotune / _set_timing / _get_timing. */
int plm_hip_batch_set_tuning (plm_hip_batch* b, int vox_per_thread);
int plm_hip_batch_set_occupancy (plm_hip_batch* b, int workgroups_per_cu);
int plm_hip_batch_get_tuning (plm_hip_batch* b, int* vox_per_thread, int* workgroups_per_cu);
int plm_hip_batch_autotune (plm_hip_batch* b, int grad_mode, int reps,
    int* chosen_vox_per_thread, double* eval_ms);
int plm_hip_batch_set_timing (plm_hip_batch* b, int enable);
int plm_hip_batch_get_timing (plm_hip_batch* b, double* stage_ms, int max_stages,
    int64_t* evals);

int plm_hip_batch_destroy (plm_hip_batch* b);

/* ------------------------------------------------------------------------
 * Stage preprocessing on the device. */

/* Resample kinds of a stage (Stage_parms::resample_type with
 * Shared_parms::legacy_subsampling, src/register/registration_resample.cxx:
 * 29-50). */
#define PLM_RESAMPLE_VOXEL_RATE    0   /* RESAMPLE_AUTO / _VOXEL_RATE: volume_subsample_vox */
#define PLM_RESAMPLE_VOXEL_LEGACY  1   /* the same with legacy_subsampling: volume_subsample_vox_legacy */
#define PLM_RESAMPLE_MM            2   /* RESAMPLE_MM: volume_resample_spacing (rate = spacing in mm) */
#define PLM_RESAMPLE_PCT           3   /* RESAMPLE_PCT: volume_resample_percent (rate = fraction) */

/* The output geometry registration_resample_volume gives a stage
 * (src/register/registration_resample.cxx:13-58; the geometry functions of
 * src/base/volume_resample.cxx:298-413).  *identity = 1 when the reference
 * clones the volume instead (voxel rates of 1, 1, 1). */
int plm_hip_resample_geometry (int kind, const plm_img_desc* in, const float rate[3],
    int64_t dim[3], float origin[3], float spacing[3], int* identity);

/* Replaces volume_resample / volume_resample_nn for float volumes
 * (volume_resample_float_li / _nn, src/base/volume_resample.cxx:26-135,
 * 227-295): in resampled onto dim / origin / spacing (same direction
 * cosines), trilinear (nearest = 0) or nearest voxel (1), 0 outside; the
 * reference's result bit for bit.  in may be on the host or the device; out
 * holds dim[0]*dim[1]*dim[2] floats at out_location. */
int plm_hip_volume_resample (const plm_img_desc* in, const int64_t dim[3],
    const float origin[3], const float spacing[3], int nearest, float* out, int out_location,
    int gpuid);

/* Replaces registration_resample_volume
 * (src/register/registration_resample.cxx:13-58): the geometry above, then
 * the clone or the resample.  Size out with plm_hip_resample_geometry. */
int plm_hip_stage_resample (int kind, const float rate[3], const plm_img_desc* in, int nearest,
    float* out, int out_location, int gpuid);

/* Replaces bspline_initialize_mi_hist_eqsp
 * (src/register/joint_histogram.cxx:115-139): the image's minimum and
 * maximum by a device scan (exact, order independent), then
 * delta = (max - min) / (bins - 1), offset = min - delta / 2. */
int plm_hip_hist_eqsp (const plm_img_desc* vol, int64_t bins, float* offset, float* delta,
    int gpuid);

/* Keys of the V-optimal histogram (VOPT_RES, src/register/histogram.h:12). */
#define PLM_HIP_VOPT_KEYS 1000

/* Replaces bspline_initialize_mi_hist_vopt
 * (src/register/joint_histogram.cxx:195-324), the histogram_type=vopt
 * initialization (registration_parms.cxx:478-479): the image's range split
 * into PLM_HIP_VOPT_KEYS equal keys (offset and delta of that grid, the
 * equal-spacing formulas with the keys for bins), then the grouping of
 * consecutive keys into bins that the reference's dynamic programme picks;
 * key_lut (PLM_HIP_VOPT_KEYS ints, host memory) receives the bin of every
 * key.  The same offset, delta and table as the reference, bit for bit: the
 * range scan, the key counts and sums (exact in any order) and the
 * programme run on the device, the running sums over the keys on the host.
 * bins must be in [2, PLM_HIP_VOPT_KEYS] (the reference falls back to equal
 * spacing above, which its caller does here); a constant image is refused.
 * Use the result with plm_hip_bspline_set_hist_keys. */
int plm_hip_hist_vopt (const plm_img_desc* vol, int64_t bins, float* offset, float* delta,
    int32_t* key_lut, int gpuid);

/* ------------------------------------------------------------------------
 * The optimizer loop on the device: L-BFGS with the More-Thuente line
 * search, the algorithm of the liblbfgs 1.9 the reference vendors
 * (libs/liblbfgs-1.9/lib/lbfgs.c:256-653, 823-1330) as
 * bspline_optimize_liblbfgs runs it
 * (src/register/bspline_optimize_liblbfgs.cxx:20-109): default parameters,
 * orthantwise_c = 0, every line-search evaluation reported to the progress
 * rule.  Coefficients, gradients, the search direction and the m history
 * pairs stay in device memory (double); per evaluation the host reads back
 * only the score and a few dot products.  The direction is the two-loop
 * recursion's d = -H g evaluated in the compact form of the same matrix, so
 * it agrees with liblbfgs's to rounding, not bit for bit. */

/* Return codes of the optimizer (plm_lbfgs_result::status): liblbfgs's
 * values (libs/liblbfgs-1.9/include/lbfgs.h:72-145). */
#define PLM_LBFGS_SUCCESS                   0
#define PLM_LBFGS_STOP                      1
#define PLM_LBFGS_ALREADY_MINIMIZED         2
#define PLM_LBFGSERR_ROUNDING_ERROR     (-1001)
#define PLM_LBFGSERR_MINIMUMSTEP        (-1000)
#define PLM_LBFGSERR_MAXIMUMSTEP         (-999)
#define PLM_LBFGSERR_MAXIMUMLINESEARCH   (-998)
#define PLM_LBFGSERR_MAXIMUMITERATION    (-997)
#define PLM_LBFGSERR_WIDTHTOOSMALL       (-996)
#define PLM_LBFGSERR_INVALIDPARAMETERS   (-995)
#define PLM_LBFGSERR_INCREASEGRADIENT    (-994)
#define PLM_LBFGSERR_OUTOFINTERVAL      (-1003)
#define PLM_LBFGSERR_INCORRECT_TMINMAX  (-1002)

/* The lbfgs_parameter_t fields the More-Thuente path reads
 * (libs/liblbfgs-1.9/include/lbfgs.h:200-330), and the reference's
 * progress limits Bspline_parms::max_its / max_feval
 * (src/register/bspline_parms.cxx:16-17): the run stops once the iteration
 * number exceeds max_its or the evaluations exceed max_feval
 * (progress, src/register/bspline_optimize_liblbfgs.cxx:55-80); a negative
 * limit is no limit. */
typedef struct plm_lbfgs_parms {
    int m;                 /* history pairs, 1..16 */
    double epsilon;
    int past;              /* 0, or the period of the delta test, <= 64 */
    double delta;
    int max_iterations;
    int max_linesearch;
    double min_step;
    double max_step;
    double ftol;
    double gtol;
    double xtol;
    int max_its;
    int max_feval;
} plm_lbfgs_parms;

/* liblbfgs's defaults (lbfgs_parameter_init, lbfgs.c:113-118, 240-243)
 * and the reference's max_its = max_feval = 10. */
void plm_hip_lbfgs_parms_init (plm_lbfgs_parms* p);

/* Called after every evaluation inside the line search (ls = 0, fx the
 * line search's best value so far) and after every accepted step (ls =
 * the evaluations it took), as liblbfgs's progress callback; d_x and d_g
 * are device pointers valid during the call.  Nonzero stops the run. */
typedef int (*plm_lbfgs_progress_fn) (void* instance, const double* d_x, const double* d_g,
    double fx, double xnorm, double gnorm, double step, int n, int k, int ls);

typedef struct plm_lbfgs_result {
    int status;            /* PLM_LBFGS_* / PLM_LBFGSERR_* */
    int iterations;        /* k of the last progress report (Bspline_state::it) */
    int fevals;            /* evaluations (Bspline_state::feval) */
    double fx;             /* objective at exit, as lbfgs()'s *ptr_fx */
} plm_lbfgs_result;

/* Minimize score + regularization over the context's coefficients.  The
 * objective is plm_hip_bspline_score_device's score (metric, grad_mode)
 * plus, when reg is not NULL, plm_hip_bspline_regularize's score, with
 * the sum of their FP64 gradients.  d_x: 3 * num_knots doubles, the start
 * point in, lbfgs()'s x at exit out; d_coeff (may be NULL): the
 * coefficients of the last evaluation, as Bspline_xform::coeff holds them
 * when the reference's run ends.  parms NULL: the defaults.  Runs on the
 * context's stream and returns when the run has ended. */
int plm_hip_bspline_optimize_lbfgs (plm_hip_ctx* ctx, int metric, int grad_mode,
    const plm_reg_parms* reg, const plm_lbfgs_parms* parms, double* d_x, float* d_coeff,
    plm_lbfgs_progress_fn progress, void* instance, plm_lbfgs_result* result);

/* One progress report of a run: the arguments liblbfgs's progress
 * callback receives (fx, xnorm, gnorm, step, k, ls). */
typedef struct plm_lbfgs_report {
    double fx;
    double xnorm;
    double gnorm;
    double step;
    int k;
    int ls;
} plm_lbfgs_report;

/* The same run with the whole loop on the device: after every evaluation
 * one GPU thread advances the line search and the direction solve (the
 * same double arithmetic as the host loop, so the same run bit for bit),
 * and the evaluation chain -- trial point, score, regularization, post
 * pass, that step -- reads everything it needs from device memory, so it
 * is captured once in a hipGraph and replayed back to back with no host
 * work between evaluations; the reference's max_its / max_feval rule
 * stops it.  For callers that cannot keep a host thread in the loop: on
 * one MI355X the host-driven loop above is as fast or faster (the scalar
 * logic is latency bound on one device thread).  No counterpart in the
 * reference, whose liblbfgs calls back into the host every evaluation
 * (bspline_optimize_liblbfgs.cxx:83-109).
 * reports (may be NULL) receives the first max_reports progress reports in
 * order; n_reports (may be NULL) the number made. */
int plm_hip_bspline_optimize_lbfgs_resident (plm_hip_ctx* ctx, int metric, int grad_mode,
    const plm_reg_parms* reg, const plm_lbfgs_parms* parms, double* d_x, float* d_coeff,
    plm_lbfgs_report* reports, int max_reports, int* n_reports, plm_lbfgs_result* result);

#ifdef __cplusplus
}
#endif
#endif
