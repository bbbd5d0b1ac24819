/* CPU oracle for the B-spline score/gradient path (TEST INFRASTRUCTURE).
 *
 * This header and plm_oracle.c restate, in plain C99 + OpenMP, the CPU
 * flavors of Plastimatch's B-spline MSE and Mattes-MI score and gradient.
 * They exist only to check the HIP product path: the tests, the smoke
 * check and the CPU-baseline leg of bench.py may load liboracle.so; the
 * product library (libplmregisterhip.so) never does.
 *
 * Float semantics follow the reference's float operations one by one
 * (compiled with -ffp-contract=off so no FMA is formed); the *_f64 entry
 * points redo the same algorithm in double precision and are the FP64
 * yardstick the relative errors are reported against.
 */
#ifndef PLM_ORACLE_H
#define PLM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Volume header: mirrors the Volume fields the path reads
   (src/base/volume.h:39-57). */
typedef struct orc_vol_hdr {
    int64_t dim[3];
    float origin[3];
    float spacing[3];
    float dc[9];            /* direction cosines, row-major */
} orc_vol_hdr;

/* Optional regions of interest: uchar volumes with their own geometry
   (Bspline_state::fixed_roi / moving_roi, src/register/bspline_state.h).
   A NULL header or image means no ROI of that kind. */
typedef struct orc_rois {
    const orc_vol_hdr* fixed_hdr;
    const uint8_t* fixed;
    const orc_vol_hdr* moving_hdr;
    const uint8_t* moving;
} orc_rois;

/* B-spline grid: the arguments of Bspline_header::set
   (src/base/bspline_header.cxx:28-62). */
typedef struct orc_geom {
    float img_origin[3];
    float img_spacing[3];
    int64_t img_dim[3];
    int64_t roi_offset[3];
    int64_t roi_dim[3];
    int64_t vox_per_rgn[3];
    float dc[9];
} orc_geom;

/* Histogram parameters (src/register/histogram.h:19-35): equal spacing,
   or V-optimal where a side has a key table (f_keys / m_keys non-NULL:
   ORC_VOPT_KEYS entries, the bin of each equal-spacing key, with offset
   and delta those of the key grid). */
#define ORC_VOPT_KEYS 1000
typedef struct orc_hist {
    int64_t fixed_bins;
    int64_t moving_bins;
    float f_offset, f_delta;
    float m_offset, m_delta;
    const int32_t* f_keys;
    const int32_t* m_keys;
} orc_hist;

int64_t orc_header (const orc_geom* g, int64_t rdims[3], int64_t cdims[3]);
void orc_basis_lut (int64_t vpr, float* lut);
void orc_direction_matrices (const float dc[9], const float spacing[3],
    float step[9], float proj[9], float inv_dc[9]);
void orc_interp_vf (const orc_geom* g, const float* coeff, float* vf);
/* Warped moving image over the fixed grid (bspline_warp_dcos, float):
   linear_interp 1 = trilinear, 0 = nearest voxel; default_val outside. */
int orc_warp (const orc_geom* g, const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, int linear_interp, float default_val, int nthreads, float* out);
void orc_volume_gradient (const orc_vol_hdr* h, const float* img, float* out);
void orc_hist_eqsp (const float* img, int64_t npix, int64_t bins,
    float* offset, float* delta);
/* V-optimal histogram (bspline_initialize_mi_hist_vopt): offset and delta
   of the ORC_VOPT_KEYS-key grid and the key -> bin table; key_count and
   key_sum (optional, ORC_VOPT_KEYS each) receive the per-key voxel counts
   and sums of (v - offset).  Returns -1 for bins outside [2, keys] or a
   constant image. */
int orc_hist_vopt (const float* img, int64_t npix, int64_t bins, float* offset,
    float* delta, int32_t* key_lut, double* key_count, double* key_sum);
/* Its partition step alone: bins contiguous groups of keys from the
   prefix tables s (running sum of key averages), ssq (of their squares)
   and cnt (of key counts). */
void orc_vopt_partition (const double* s, const double* ssq, const double* cnt,
    int64_t keys, int64_t bins, int32_t* key_lut);
int64_t orc_pass1_position_mismatch (const orc_vol_hdr* fh, const orc_geom* g);

/* flavor: 'g' (tile condense, nearest-voxel m_grad), 'p' (tile condense,
   analytic trilinear derivative), 'c' (voxel-order scatter, m_grad). */
int orc_mse_score (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img, const float* m_grad,
    const float* coeff, int flavor, int nthreads, const orc_rois* rois,
    double* raw_score, float* score, float* grad, int64_t* num_vox);

/* grad_mode 0 = nearest-voxel m_grad (flavor g), 1 = analytic (flavor p). */
int orc_mse_score_f64 (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img, const float* m_grad,
    const float* coeff, int grad_mode, int nthreads, const orc_rois* rois,
    double* score, double* grad, int64_t* num_vox);

void orc_pv_quantize (const float w[8], int64_t q[8]);

/* Regularization weights (src/register/regularization_parms.h:20-28). */
typedef struct orc_reg_parms {
    float total_displacement_penalty;
    float diffusion_penalty;
    float curvature_penalty;
    float lame_coefficient_1;
    float lame_coefficient_2;
    float linear_elastic_multiplier;
    float third_order_penalty;
    float curvature_mixed_weight;
} orc_reg_parms;

/* Analytic regularization (flavor c): score and the gradient it adds,
   both double, for the coefficients of g's grid. */
int orc_reg_analytic (const orc_geom* g, const orc_reg_parms* rp, const float* coeff,
    int nthreads, double* score, double* grad);
/* Its 32 64x64 V matrices (32 * 4096 doubles). */
int orc_reg_matrices (const orc_geom* g, const orc_reg_parms* rp, double* V);

/* FP64 sum of squared differences over the support of coefficient ci
   with that coefficient moved by delta (central differences of the FP64
   MSE without a whole-image evaluation); n and nclamp count the support's
   inside and clamped voxels. */
int orc_mse_f64_support (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, int64_t ci, double delta, const orc_rois* rois,
    double* raw, int64_t* n, int64_t* nclamp);

int orc_mi_score (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, const orc_hist* hp, int nthreads, const orc_rois* rois,
    float* score, float* grad, int64_t* num_vox,
    double* f_hist, double* m_hist, double* j_hist,
    int64_t* fx_f_hist, int64_t* fx_j_hist);

int orc_mi_score_f64 (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, const orc_hist* hp, int nthreads, const orc_rois* rois,
    double* score, double* grad, int64_t* num_vox);
/* The same on given histograms (f, m, joint, concatenated; count nv_in):
   FP64 score and pass 2 only. */
int orc_mi_score_f64_hist (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, const orc_hist* hp, int nthreads, const orc_rois* rois,
    const double* hist_in, int64_t nv_in,
    double* score, double* grad, int64_t* num_vox);

/* Stage subsampling (resample_oracle.c).  Kinds follow the stage's
   resample type and legacy_subsampling flag
   (src/register/registration_resample.cxx:29-50). */
enum {
    ORC_RESAMPLE_VOXEL_RATE = 0,    /* RESAMPLE_AUTO / _VOXEL_RATE: volume_subsample_vox */
    ORC_RESAMPLE_VOXEL_LEGACY = 1,  /* the same with legacy_subsampling */
    ORC_RESAMPLE_MM = 2,            /* volume_resample_spacing */
    ORC_RESAMPLE_PCT = 3            /* volume_resample_percent */
};
/* Output geometry; returns 1 when the volume is cloned (rate 1, 1, 1),
   0 when it is resampled, -1 for a kind or rate the reference cannot use. */
int orc_resample_geometry (int kind, const orc_vol_hdr* in, const float rate[3],
    int64_t dim[3], float origin[3], float spacing[3]);
/* volume_resample_float_li (nearest = 0) or _nn (nearest = 1) onto the
   given geometry; out holds dim[0]*dim[1]*dim[2] floats. */
void orc_volume_resample (const orc_vol_hdr* in, const float* img, const int64_t dim[3],
    const float origin[3], const float spacing[3], int nearest, float* out);

/* L-BFGS (lbfgs_oracle.c): the reference's vendored liblbfgs 1.9 on its
   More-Thuente path.  Return codes keep liblbfgs's values
   (libs/liblbfgs-1.9/include/lbfgs.h:72-145). */
enum {
    ORC_LBFGS_SUCCESS = 0,
    ORC_LBFGS_STOP = 1,
    ORC_LBFGS_ALREADY_MINIMIZED = 2,
    ORC_LBFGSERR_OUTOFMEMORY = -1022,
    ORC_LBFGSERR_INVALID_N = -1020,
    ORC_LBFGSERR_INVALID_EPSILON = -1017,
    ORC_LBFGSERR_INVALID_TESTPERIOD = -1016,
    ORC_LBFGSERR_INVALID_DELTA = -1015,
    ORC_LBFGSERR_INVALID_MINSTEP = -1013,
    ORC_LBFGSERR_INVALID_MAXSTEP = -1012,
    ORC_LBFGSERR_INVALID_FTOL = -1011,
    ORC_LBFGSERR_INVALID_GTOL = -1009,
    ORC_LBFGSERR_INVALID_XTOL = -1008,
    ORC_LBFGSERR_INVALID_MAXLINESEARCH = -1007,
    ORC_LBFGSERR_OUTOFINTERVAL = -1003,
    ORC_LBFGSERR_INCORRECT_TMINMAX = -1002,
    ORC_LBFGSERR_ROUNDING_ERROR = -1001,
    ORC_LBFGSERR_MINIMUMSTEP = -1000,
    ORC_LBFGSERR_MAXIMUMSTEP = -999,
    ORC_LBFGSERR_MAXIMUMLINESEARCH = -998,
    ORC_LBFGSERR_MAXIMUMITERATION = -997,
    ORC_LBFGSERR_WIDTHTOOSMALL = -996,
    ORC_LBFGSERR_INVALIDPARAMETERS = -995,
    ORC_LBFGSERR_INCREASEGRADIENT = -994
};

typedef double (*orc_lbfgs_eval_fn) (void* inst, const double* x, double* g, int n,
    double step);
typedef int (*orc_lbfgs_progress_fn) (void* inst, const double* x, const double* g,
    double fx, double xnorm, double gnorm, double step, int n, int k, int ls);

/* The fields of lbfgs_parameter_t that the More-Thuente path reads. */
typedef struct orc_lbfgs_parms {
    int m;
    double epsilon;
    int past;
    double delta;
    int max_iterations;
    int max_linesearch;
    double min_step;
    double max_step;
    double ftol;
    double gtol;
    double xtol;
} orc_lbfgs_parms;

void orc_lbfgs_parms_init (orc_lbfgs_parms* p);
int orc_lbfgs (int n, double* x, double* fx, orc_lbfgs_eval_fn eval,
    orc_lbfgs_progress_fn progress, void* inst, const orc_lbfgs_parms* p);

#ifdef __cplusplus
}
#endif
#endif
