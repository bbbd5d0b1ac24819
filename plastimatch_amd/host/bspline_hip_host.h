/* C++ host side of the threading=hip B-spline flavor.
 *
 * Mirrors the objects and the dispatch the reference's score path uses,
 * restricted to what the path reads and writes:
 *   Volume            src/base/volume.h:39-57
 *   Bspline_xform     src/base/bspline_xform.h:22-96, bspline_header.cxx:28-62
 *   Histogram,
 *   Joint_histogram   src/register/histogram.h:19-35, joint_histogram.h:14-46
 *   Bspline_parms     src/register/bspline_parms.h:34-92 (BTHR_HIP added)
 *   Bspline_score     src/register/bspline_score.h:33-95
 *   Bspline_state     src/register/bspline_state.h:21-67
 *   Bspline_optimize  src/register/bspline_optimize.h:15-31
 *   bspline_score()   src/register/bspline.cxx:558-632
 * The HIP library is loaded at run time (dlopen/dlsym, as LOAD_LIBRARY_SAFE /
 * LOAD_SYMBOL do, src/sys/cuda_delayload.h:26-59) and driven only through
 * its C ABI (include/plm_hip_bspline.h).  Failures are reported as return
 * codes and messages; nothing here calls exit().
 */
#ifndef BSPLINE_HIP_HOST_H
#define BSPLINE_HIP_HOST_H

#include <stdint.h>
#include <string>
#include <vector>

#include "plm_hip_bspline.h"

typedef int64_t plm_long;

enum Similarity_metric_type {
    SIMILARITY_METRIC_MSE,
    SIMILARITY_METRIC_MI_MATTES
};

enum BsplineThreading {
    BTHR_CPU,
    BTHR_CUDA,
    BTHR_HIP
};

class Volume {
public:
    plm_long dim[3];
    plm_long npix;
    float origin[3];
    float spacing[3];
    float direction_cosines[9];
    int vox_planes;               /* 1 = scalar float, 3 = float3 interleaved */
    std::vector<float> img;
public:
    Volume (const plm_long dim[3], const float origin[3], const float spacing[3],
        const float* dc, int vox_planes);
    plm_img_desc desc (int storage) const;
};

/* A uchar volume (PT_UCHAR, src/base/volume.h:39-57): the fixed and moving
   regions of interest. */
class Volume_uchar {
public:
    plm_long dim[3];
    plm_long npix;
    float origin[3];
    float spacing[3];
    float direction_cosines[9];
    std::vector<uint8_t> img;
public:
    Volume_uchar (const plm_long dim[3], const float origin[3], const float spacing[3],
        const float* dc);
    plm_img_desc desc () const;
};

/* update_roi (src/register/bspline_stage.cxx:71-93): voxels of image
   outside [min_val, max_val] leave roi; with fill_empty_roi, voxels inside
   the range join it.  roi and image share one grid, voxel for voxel
   (false, and nothing changed, if their dims differ). */
bool update_roi (Volume_uchar* roi, const Volume* image, float min_val, float max_val,
    bool fill_empty_roi);

class Bspline_xform {
public:
    float img_origin[3];
    float img_spacing[3];
    plm_long img_dim[3];
    plm_long roi_offset[3];
    plm_long roi_dim[3];
    plm_long vox_per_rgn[3];
    float grid_spac[3];
    plm_long rdims[3];
    plm_long cdims[3];
    float dc[9];
    plm_long num_knots;
    plm_long num_coeff;
    std::vector<float> coeff;
public:
    void initialize (const float img_origin[3], const float img_spacing[3],
        const plm_long img_dim[3], const plm_long roi_offset[3],
        const plm_long roi_dim[3], const plm_long vox_per_rgn[3],
        const float direction_cosines[9]);
    void fill_coefficients (float val);
    void jitter_if_zero ();
    plm_bsp_geom geom () const;
    /* MGH_GPUIT_BSP text file (Bspline_xform::save,
       src/base/bspline_xform.cxx:101-145); false if it cannot be opened. */
    bool save (const char* filename) const;
};

/* MetaImage files as the reference reads and writes them (read_mha /
   write_mha, src/base/mha_io.cxx:26-257): data after the header
   (ElementDataFile = LOCAL, or a raw file next to it on read), MET_FLOAT,
   MET_SHORT or MET_UCHAR (widened to float here: the path reads float),
   float vector fields with ElementNumberOfChannels = 3, TransformMatrix by
   columns, either byte order.  nullptr and a message in *err on failure. */
Volume* read_mha (const std::string& filename, std::string* err);
/* Float volume or float3 field with the reference's header. */
bool write_mha (const std::string& filename, const Volume& vol, std::string* err);

/* Read an MGH_GPUIT_BSP file (bspline_xform_load,
   src/base/bspline_xform.cxx:147-279); nullptr and a message in *err on a
   file the reference would also refuse. */
Bspline_xform* bspline_xform_load (const char* filename, std::string* err);

/* Mi_hist_type (src/register/histogram.h:14-17) */
enum Mi_hist_type { HIST_EQSP, HIST_VOPT };

class Histogram {
public:
    Mi_hist_type type;
    plm_long bins;
    float offset;
    float delta;
    std::vector<int32_t> key_lut;       /* HIST_VOPT: bin of each key */
    Histogram (Mi_hist_type type = HIST_EQSP, plm_long bins = 32)
        : type (type), bins (bins), offset (0.f), delta (0.f) {}
};

class Joint_histogram {
public:
    Histogram moving;
    Histogram fixed;
    std::vector<double> f_hist;
    std::vector<double> m_hist;
    std::vector<double> j_hist;
    std::string last_error;
public:
    Joint_histogram (plm_long fixed_bins, plm_long moving_bins);
    Joint_histogram (Mi_hist_type type, plm_long fixed_bins, plm_long moving_bins);
    /* Equal spacing on the host; V-optimal bins through the plugin
       (plm_hip_hist_vopt, device gpuid).  Returns false with last_error
       set when the plugin cannot build them. */
    bool initialize (const Volume* fixed, const Volume* moving, int gpuid = 0);
    plm_hist_desc desc () const;
};

class Bspline_parms {
public:
    enum BsplineThreading threading;
    char implementation;
    int gpuid;
    int storage;                  /* PLM_STORE_F32 or PLM_STORE_BF16 */
    int grad_mode;                /* PLM_GRAD_REF_NN (flavor g) or PLM_GRAD_EXACT (p) */
    plm_long mi_hist_fixed_bins;
    plm_long mi_hist_moving_bins;
    /* user intensity ranges of MI (Bspline_parms, src/register/
       bspline_parms.h:34-92); a set range becomes a ROI (bspline_stage_rois) */
    float mi_fixed_image_minVal;
    float mi_fixed_image_maxVal;
    float mi_moving_image_minVal;
    float mi_moving_image_maxVal;
    int max_its;
    int max_feval;
public:
    Bspline_parms ();
};

class Metric_score {
public:
    float score;
    double time;
    plm_long num_vox;
};

class Bspline_score {
public:
    float total_score;
    std::vector<float> total_grad;
    plm_long num_coeff;
    float curr_smetric;
    std::vector<float> curr_smetric_grad;
    plm_long curr_num_vox;
    /* FP64 side channel of the last evaluation, for error reporting */
    double curr_smetric_f64;
    std::vector<double> curr_smetric_grad_f64;
    std::vector<Metric_score> metric_record;
public:
    Bspline_score ();
    void set_num_coeff (plm_long n);
    void reset_score ();
    void accumulate (float lambda);
};

class Bspline_state {
public:
    int it;
    int feval;
    Bspline_score ssd;
    void* dev_ptrs;               /* plm_hip_ctx*, created once per stage */
    Similarity_metric_type metric_type;
    float metric_lambda;
    Volume* fixed;
    Volume* moving;
    Volume* moving_grad;
    Joint_histogram* mi_hist;
    Volume_uchar* fixed_roi;      /* not owned; either may be null */
    Volume_uchar* moving_roi;
    std::string last_error;
public:
    Bspline_state ();
    ~Bspline_state ();
};

class Bspline_optimize {
public:
    Bspline_parms* parms;
    Bspline_state* bst;
    Bspline_xform* bxf;
public:
    Bspline_optimize (Bspline_parms* p, Bspline_state* s, Bspline_xform* x)
        : parms (p), bst (s), bxf (x) {}
    Bspline_parms* get_bspline_parms () { return parms; }
    Bspline_state* get_bspline_state () { return bst; }
    Bspline_xform* get_bspline_xform () { return bxf; }
    /* Steepest descent with a trust-region step, one bspline_score per
       function evaluation.  Returns 0 or a negative status. */
    int optimize ();
};

/* Loads libplmregisterhip.so (once) and resolves every symbol; returns
   false with a message when the library or a symbol is missing. */
bool hip_plugin_load (std::string* err);

/* The stage's ROI preparation (src/register/bspline_stage.cxx:150-192):
   each image's intensity range, when set (min or max non-zero, min < max),
   narrows that image's ROI, or makes one on the image's grid, which *made
   then holds (owned by the caller).  Returns false if a given ROI's dims
   differ from its image's. */
bool bspline_stage_rois (const Bspline_parms* parms, const Volume* fixed, const Volume* moving,
    Volume_uchar** fixed_roi, Volume_uchar** moving_roi, std::vector<Volume_uchar*>* made);

/* Creates the device state for the stage (images uploaded once, and the
   state's ROIs, if any, handed to the library). */
int bspline_state_initialize_hip (Bspline_state* bst, Bspline_xform* bxf,
    Bspline_parms* parms);

/* One evaluation: fills bst->ssd like the reference's bspline_score. */
int bspline_score (Bspline_optimize* bod);

/* Moving-image gradient on the GPU (volume_gradient). */
int volume_gradient_hip (const Volume* vol, Volume* out, int gpuid);

#endif
