/* C++ host side of the threading=hip B-spline flavor; see bspline_hip_host.h. */
#include "bspline_hip_host.h"

#include <algorithm>
#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <chrono>
#include <string>

/* ---------------------------------------------------------------------
   Plugin loading (the LOAD_LIBRARY_SAFE / LOAD_SYMBOL pattern of
   src/sys/cuda_delayload.h:26-59, without the exit(0) on failure)
   --------------------------------------------------------------------- */

namespace {

struct Hip_plugin {
    void* handle = nullptr;
#define PLM_SYM(name) decltype (&::name) name = nullptr;
    PLM_SYM (plm_hip_abi_version)
    PLM_SYM (plm_hip_last_error)
    PLM_SYM (plm_hip_bspline_create)
    PLM_SYM (plm_hip_bspline_score)
    PLM_SYM (plm_hip_bspline_score_device)
    PLM_SYM (plm_hip_bspline_destroy)
    PLM_SYM (plm_hip_volume_gradient)
    PLM_SYM (plm_hip_bspline_num_knots)
    PLM_SYM (plm_hip_bspline_set_roi)
    PLM_SYM (plm_hip_hist_vopt)
    PLM_SYM (plm_hip_bspline_set_hist_keys)
#undef PLM_SYM
};

Hip_plugin g_plugin;

std::string
plugin_dir ()
{
    Dl_info info;
    if (dladdr ((void*) &plugin_dir, &info) && info.dli_fname) {
        std::string p (info.dli_fname);
        size_t s = p.rfind ('/');
        if (s != std::string::npos) return p.substr (0, s);
    }
    return ".";
}

std::string
hip_error ()
{
    const char* m = g_plugin.plm_hip_last_error ? g_plugin.plm_hip_last_error () : nullptr;
    return m ? m : "";
}

} // namespace

bool
hip_plugin_load (std::string* err)
{
    if (g_plugin.handle) return true;
    std::vector<std::string> cands;
    if (const char* env = getenv ("PLM_HIP_LIBRARY")) cands.push_back (env);
    cands.push_back (plugin_dir () + "/../libplmregisterhip.so");
    cands.push_back (plugin_dir () + "/libplmregisterhip.so");
    cands.push_back ("libplmregisterhip.so");
    void* h = nullptr;
    std::string tried;
    for (const std::string& c : cands) {
        h = dlopen (c.c_str (), RTLD_NOW | RTLD_LOCAL);
        if (h) break;
        tried += " " + c;
    }
    if (!h) {
        if (err) *err = "cannot load libplmregisterhip.so (tried" + tried + ")";
        return false;
    }
#define PLM_LOAD(name)                                                      \
    g_plugin.name = (decltype (g_plugin.name)) dlsym (h, #name);            \
    if (!g_plugin.name) {                                                   \
        if (err) *err = "missing symbol " #name;                            \
        dlclose (h);                                                        \
        return false;                                                       \
    }
    PLM_LOAD (plm_hip_abi_version)
    PLM_LOAD (plm_hip_last_error)
    PLM_LOAD (plm_hip_bspline_create)
    PLM_LOAD (plm_hip_bspline_score)
    PLM_LOAD (plm_hip_bspline_score_device)
    PLM_LOAD (plm_hip_bspline_destroy)
    PLM_LOAD (plm_hip_volume_gradient)
    PLM_LOAD (plm_hip_bspline_num_knots)
    PLM_LOAD (plm_hip_bspline_set_roi)
    PLM_LOAD (plm_hip_hist_vopt)
    PLM_LOAD (plm_hip_bspline_set_hist_keys)
#undef PLM_LOAD
    if (g_plugin.plm_hip_abi_version () != PLM_HIP_ABI_VERSION) {
        if (err) *err = "libplmregisterhip.so ABI version mismatch";
        dlclose (h);
        return false;
    }
    g_plugin.handle = h;
    return true;
}

/* ---------------------------------------------------------------------
   Volume, Bspline_xform, Joint_histogram
   --------------------------------------------------------------------- */

static const float IDENTITY[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};

Volume::Volume (const plm_long d[3], const float o[3], const float sp[3], const float* dc,
    int planes)
{
    for (int a = 0; a < 3; a++) {
        dim[a] = d[a];
        origin[a] = o[a];
        spacing[a] = sp[a];
    }
    memcpy (direction_cosines, dc ? dc : IDENTITY, sizeof direction_cosines);
    vox_planes = planes;
    npix = dim[0] * dim[1] * dim[2];
    img.assign ((size_t) npix * planes, 0.f);
}

plm_img_desc
Volume::desc (int storage) const
{
    plm_img_desc d;
    d.data = img.data ();
    d.location = PLM_MEM_HOST;
    d.storage = storage;
    for (int a = 0; a < 3; a++) {
        d.dim[a] = dim[a];
        d.origin[a] = origin[a];
        d.spacing[a] = spacing[a];
    }
    memcpy (d.direction_cosines, direction_cosines, sizeof d.direction_cosines);
    return d;
}

Volume_uchar::Volume_uchar (const plm_long d[3], const float o[3], const float sp[3], const float* dc)
{
    for (int a = 0; a < 3; a++) {
        dim[a] = d[a];
        origin[a] = o[a];
        spacing[a] = sp[a];
    }
    memcpy (direction_cosines, dc ? dc : IDENTITY, sizeof direction_cosines);
    npix = dim[0] * dim[1] * dim[2];
    img.assign ((size_t) npix, 0);
}

plm_img_desc
Volume_uchar::desc () const
{
    plm_img_desc d;
    d.data = img.data ();
    d.location = PLM_MEM_HOST;
    d.storage = PLM_STORE_U8;
    for (int a = 0; a < 3; a++) {
        d.dim[a] = dim[a];
        d.origin[a] = origin[a];
        d.spacing[a] = spacing[a];
    }
    memcpy (d.direction_cosines, direction_cosines, sizeof d.direction_cosines);
    return d;
}

bool
update_roi (Volume_uchar* roi, const Volume* image, float min_val, float max_val, bool fill_empty_roi)
{
    for (int a = 0; a < 3; a++)
        if (roi->dim[a] != image->dim[a]) return false;
    for (plm_long p = 0; p < roi->npix; p++) {
        const float v = image->img[(size_t) p];
        const bool in_range = v >= min_val && v <= max_val;
        if (fill_empty_roi && in_range) roi->img[(size_t) p] = 1;
        else if (!in_range && roi->img[(size_t) p] > 0) roi->img[(size_t) p] = 0;
    }
    return true;
}

bool
bspline_stage_rois (const Bspline_parms* parms, const Volume* fixed, const Volume* moving,
    Volume_uchar** fixed_roi, Volume_uchar** moving_roi, std::vector<Volume_uchar*>* made)
{
    struct Side { float lo, hi; const Volume* img; Volume_uchar** roi; };
    const Side sides[2] = {
        {parms->mi_moving_image_minVal, parms->mi_moving_image_maxVal, moving, moving_roi},
        {parms->mi_fixed_image_minVal, parms->mi_fixed_image_maxVal, fixed, fixed_roi}};
    for (const Side& sd : sides) {
        if (!((sd.lo != 0 || sd.hi != 0) && sd.lo < sd.hi)) continue;
        const bool fill = *sd.roi == nullptr;
        if (fill) {
            *sd.roi = new Volume_uchar (sd.img->dim, sd.img->origin, sd.img->spacing,
                sd.img->direction_cosines);
            made->push_back (*sd.roi);
        }
        if (!update_roi (*sd.roi, sd.img, sd.lo, sd.hi, fill)) return false;
    }
    return true;
}

/* rdims = 1 + (roi_dim-1)/vpr, cdims = rdims + 3, as Bspline_header::set
   (src/base/bspline_header.cxx:28-62); coefficients start at zero as in
   Bspline_xform::allocate (src/base/bspline_xform.cxx:366-368). */
void
Bspline_xform::initialize (const float o[3], const float sp[3], const plm_long d[3],
    const plm_long roff[3], const plm_long rdim[3], const plm_long vpr[3], const float* dcos)
{
    for (int a = 0; a < 3; a++) {
        img_origin[a] = o[a];
        img_spacing[a] = sp[a];
        img_dim[a] = d[a];
        roi_offset[a] = roff[a];
        roi_dim[a] = rdim[a];
        vox_per_rgn[a] = vpr[a];
        grid_spac[a] = vpr[a] * fabsf (sp[a]);
        rdims[a] = 1 + (rdim[a] - 1) / vpr[a];
        cdims[a] = 3 + rdims[a];
    }
    memcpy (dc, dcos ? dcos : IDENTITY, sizeof dc);
    num_knots = cdims[0] * cdims[1] * cdims[2];
    num_coeff = 3 * num_knots;
    coeff.assign ((size_t) num_coeff, 0.f);
}

Volume*
read_mha (const std::string& fn, std::string* err)
{
    FILE* fp = fopen (fn.c_str (), "rb");
    if (!fp) {
        if (err) *err = "cannot open " + fn;
        return nullptr;
    }
    auto fail = [&] (const std::string& m) -> Volume* {
        fclose (fp);
        if (err) *err = fn + ": " + m;
        return nullptr;
    };
    plm_long dim[3] = {0, 0, 0};
    float sp[3] = {1.f, 1.f, 1.f}, org[3] = {0.f, 0.f, 0.f};
    float dc[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    int channels = 1, esize = 0;
    char etype = 0;                     // 'f', 's', 'u'
    bool big = false;
    std::string datafile;
    char line[512];
    for (;;) {
        if (!fgets (line, sizeof line, fp)) return fail ("no ElementDataFile line");
        size_t len = strlen (line);
        while (len > 0 && (line[len - 1] == '\n' || line[len - 1] == '\r' || line[len - 1] == ' '))
            line[--len] = 0;
        if (strncmp (line, "ElementDataFile = ", 18) == 0) {
            datafile = line + 18;
            break;
        }
        long a, b, c;
        if (sscanf (line, "DimSize = %ld %ld %ld", &a, &b, &c) == 3) {
            dim[0] = a; dim[1] = b; dim[2] = c;
        } else if (sscanf (line, "Offset = %g %g %g", &org[0], &org[1], &org[2]) == 3) {
        } else if (sscanf (line, "ElementSpacing = %g %g %g", &sp[0], &sp[1], &sp[2]) == 3) {
        } else if (sscanf (line, "TransformMatrix = %g %g %g %g %g %g %g %g %g", &dc[0], &dc[3], &dc[6], &dc[1],
                       &dc[4], &dc[7], &dc[2], &dc[5], &dc[8]) == 9) {
        } else if (sscanf (line, "ElementNumberOfChannels = %d", &channels) == 1) {
        } else if (strcmp (line, "ElementType = MET_FLOAT") == 0) {
            etype = 'f'; esize = 4;
        } else if (strcmp (line, "ElementType = MET_SHORT") == 0) {
            etype = 's'; esize = 2;
        } else if (strcmp (line, "ElementType = MET_UCHAR") == 0) {
            etype = 'u'; esize = 1;
        } else if (strcmp (line, "BinaryDataByteOrderMSB = True") == 0) {
            big = true;
        }
    }
    if (!etype) return fail ("cannot interpret the element type");
    if (channels != 1 && (channels != 3 || etype != 'f')) return fail ("only float3 vector fields are read");
    if (dim[0] < 1 || dim[1] < 1 || dim[2] < 1) return fail ("bad DimSize");
    Volume* v = new Volume (dim, org, sp, dc, channels);
    const size_t n = v->img.size ();
    std::vector<unsigned char> raw (n * (size_t) esize);
    size_t got;
    if (datafile == "LOCAL") {
        got = fread (raw.data (), 1, raw.size (), fp);
    } else {
        const std::string dir = fn.substr (0, fn.rfind ('/') + 1);
        FILE* r = fopen ((dir + datafile).c_str (), "rb");
        if (!r) {
            delete v;
            return fail ("cannot open " + datafile);
        }
        got = fread (raw.data (), 1, raw.size (), r);
        fclose (r);
    }
    fclose (fp);
    if (got != raw.size ()) {
        delete v;
        if (err) *err = fn + ": short data";
        return nullptr;
    }
    for (size_t i = 0; i < n; i++) {
        unsigned char* p = raw.data () + i * esize;
        if (big) std::reverse (p, p + esize);   // data are little-endian on this host
        if (etype == 'f') {
            float f;
            memcpy (&f, p, 4);
            v->img[i] = f;
        } else if (etype == 's') {
            int16_t s16;
            memcpy (&s16, p, 2);
            v->img[i] = (float) s16;
        } else {
            v->img[i] = (float) *p;
        }
    }
    return v;
}

bool
write_mha (const std::string& fn, const Volume& vol, std::string* err)
{
    FILE* fp = fopen (fn.c_str (), "wb");
    if (!fp) {
        if (err) *err = "cannot open " + fn;
        return false;
    }
    const float* dc = vol.direction_cosines;
    fprintf (fp,
        "ObjectType = Image\n"
        "NDims = 3\n"
        "BinaryData = True\n"
        "BinaryDataByteOrderMSB = False\n"
        "TransformMatrix = %g %g %g %g %g %g %g %g %g\n"
        "Offset = %g %g %g\n"
        "CenterOfRotation = 0 0 0\n"
        "ElementSpacing = %g %g %g\n"
        "DimSize = %d %d %d\n"
        "AnatomicalOrientation = RAI\n"
        "%s"
        "ElementType = MET_FLOAT\n"
        "ElementDataFile = LOCAL\n",
        dc[0], dc[3], dc[6], dc[1], dc[4], dc[7], dc[2], dc[5], dc[8], vol.origin[0], vol.origin[1],
        vol.origin[2], vol.spacing[0], vol.spacing[1], vol.spacing[2], (int) vol.dim[0], (int) vol.dim[1],
        (int) vol.dim[2], vol.vox_planes == 3 ? "ElementNumberOfChannels = 3\n" : "");
    const size_t n = vol.img.size ();
    const bool ok = fwrite (vol.img.data (), sizeof (float), n, fp) == n;
    if (fclose (fp) != 0 || !ok) {
        if (err) *err = "cannot write " + fn;
        return false;
    }
    return true;
}

bool
Bspline_xform::save (const char* filename) const
{
    FILE* fp = fopen (filename, "wb");
    if (!fp) return false;
    fprintf (fp, "MGH_GPUIT_BSP <experimental>\n");
    fprintf (fp, "img_origin = %f %f %f\n", img_origin[0], img_origin[1], img_origin[2]);
    fprintf (fp, "img_spacing = %f %f %f\n", img_spacing[0], img_spacing[1], img_spacing[2]);
    fprintf (fp, "img_dim = %u %u %u\n", (unsigned) img_dim[0], (unsigned) img_dim[1], (unsigned) img_dim[2]);
    fprintf (fp, "roi_offset = %d %d %d\n", (int) roi_offset[0], (int) roi_offset[1], (int) roi_offset[2]);
    fprintf (fp, "roi_dim = %d %d %d\n", (int) roi_dim[0], (int) roi_dim[1], (int) roi_dim[2]);
    fprintf (fp, "vox_per_rgn = %d %d %d\n", (int) vox_per_rgn[0], (int) vox_per_rgn[1], (int) vox_per_rgn[2]);
    fprintf (fp, "direction_cosines = %f %f %f %f %f %f %f %f %f\n", dc[0], dc[1], dc[2], dc[3], dc[4],
        dc[5], dc[6], dc[7], dc[8]);
    /* planar: every knot's x, then y, then z */
    for (int d = 0; d < 3; d++)
        for (plm_long k = 0; k < num_knots; k++)
            fprintf (fp, "%.20f\n", coeff[(size_t) (3 * k + d)]);
    return fclose (fp) == 0;
}

Bspline_xform*
bspline_xform_load (const char* filename, std::string* err)
{
    float o[3] = {0.f, 0.f, 0.f}, sp[3] = {1.f, 1.f, 1.f};
    float dcos[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    plm_long dim[3] = {0, 0, 0}, roff[3] = {0, 0, 0}, rdim[3] = {0, 0, 0}, vpr[3] = {0, 0, 0};
    FILE* fp = fopen (filename, "r");
    if (!fp) {
        if (err) *err = std::string ("cannot open ") + filename;
        return nullptr;
    }
    auto fail = [&] (const std::string& m) -> Bspline_xform* {
        fclose (fp);
        if (err) *err = std::string (filename) + ": " + m;
        return nullptr;
    };
    char line[1024];
    if (!fgets (line, sizeof line, fp) || strncmp (line, "MGH_GPUIT_BSP", 13) != 0)
        return fail ("not an MGH_GPUIT_BSP file");
    /* header lines until the first line without '=' (the first coefficient) */
    for (;;) {
        if (!fgets (line, sizeof line, fp)) return fail ("header ends without coefficients");
        if (!strchr (line, '=')) break;
        unsigned a, b, c;
        if (sscanf (line, "img_origin = %f %f %f", &o[0], &o[1], &o[2]) == 3) continue;
        if (sscanf (line, "img_spacing = %f %f %f", &sp[0], &sp[1], &sp[2]) == 3) continue;
        if (sscanf (line, "img_dim = %u %u %u", &a, &b, &c) == 3) {
            dim[0] = a; dim[1] = b; dim[2] = c;
            continue;
        }
        if (sscanf (line, "roi_offset = %u %u %u", &a, &b, &c) == 3) {
            roff[0] = a; roff[1] = b; roff[2] = c;
            continue;
        }
        if (sscanf (line, "roi_dim = %u %u %u", &a, &b, &c) == 3) {
            rdim[0] = a; rdim[1] = b; rdim[2] = c;
            continue;
        }
        if (sscanf (line, "vox_per_rgn = %u %u %u", &a, &b, &c) == 3) {
            vpr[0] = a; vpr[1] = b; vpr[2] = c;
            continue;
        }
        if (sscanf (line, "direction_cosines = %f %f %f %f %f %f %f %f %f", &dcos[0], &dcos[1], &dcos[2],
                &dcos[3], &dcos[4], &dcos[5], &dcos[6], &dcos[7], &dcos[8]) == 9)
            continue;
        return fail (std::string ("bad header line: ") + line);
    }
    for (int a = 0; a < 3; a++)
        if (vpr[a] < 1 || rdim[a] < 1) return fail ("no coefficients");
    Bspline_xform* bxf = new Bspline_xform;
    bxf->initialize (o, sp, dim, roff, rdim, vpr, dcos);
    for (int d = 0; d < 3; d++) {
        for (plm_long k = 0; k < bxf->num_knots; k++) {
            if ((d != 0 || k != 0) && !fgets (line, sizeof line, fp)) {
                delete bxf;
                return fail ("too few coefficients");
            }
            if (sscanf (line, "%f", &bxf->coeff[(size_t) (3 * k + d)]) != 1) {
                delete bxf;
                return fail ("bad coefficient line");
            }
        }
    }
    fclose (fp);
    return bxf;
}

void
Bspline_xform::fill_coefficients (float val)
{
    for (auto& c : coeff) c = val;
}

/* All-zero coefficients become 0.01 (src/base/bspline_xform.cxx:630-644). */
void
Bspline_xform::jitter_if_zero ()
{
    for (float c : coeff)
        if (c != 0.0f) return;
    fill_coefficients (0.01f);
}

plm_bsp_geom
Bspline_xform::geom () const
{
    plm_bsp_geom g;
    for (int a = 0; a < 3; a++) {
        g.img_origin[a] = img_origin[a];
        g.img_spacing[a] = img_spacing[a];
        g.img_dim[a] = img_dim[a];
        g.roi_offset[a] = roi_offset[a];
        g.roi_dim[a] = roi_dim[a];
        g.vox_per_rgn[a] = vox_per_rgn[a];
    }
    memcpy (g.direction_cosines, dc, sizeof g.direction_cosines);
    return g;
}

Joint_histogram::Joint_histogram (plm_long fb, plm_long mb)
    : Joint_histogram (HIST_EQSP, fb, mb)
{
}

Joint_histogram::Joint_histogram (Mi_hist_type type, plm_long fb, plm_long mb)
    : moving (type, mb), fixed (type, fb), f_hist (fb, 0.0), m_hist (mb, 0.0), j_hist (fb * mb, 0.0)
{
}

/* delta = (max-min)/(bins-1), offset = min - 0.5*delta, float, as
   bspline_initialize_mi_hist_eqsp (src/register/joint_histogram.cxx:115-139). */
static void
hist_eqsp (Histogram* h, const Volume* v)
{
    float mn = v->img[0], mx = v->img[0];
    for (plm_long i = 0; i < v->npix; i++) {
        if (v->img[i] < mn) mn = v->img[i];
        else if (v->img[i] > mx) mx = v->img[i];
    }
    h->delta = (mx - mn) / (float) (h->bins - 1);
    h->offset = (float) (mn - 0.5 * h->delta);
}

/* bspline_initialize_mi_hist (src/register/joint_histogram.cxx:327-346):
   more V-optimal bins than keys fall back to equal spacing. */
bool
Joint_histogram::initialize (const Volume* f, const Volume* m, int gpuid)
{
    Histogram* hs[2] = {&fixed, &moving};
    const Volume* vs[2] = {f, m};
    for (int k = 0; k < 2; k++) {
        Histogram* h = hs[k];
        if (h->type == HIST_VOPT && h->bins > PLM_HIP_VOPT_KEYS) {
            printf ("WARNING: Falling back to EQSP histograms.\n"
                    "         (Reason: # bins > %i)\n", PLM_HIP_VOPT_KEYS);
            h->type = HIST_EQSP;
        }
        if (h->type == HIST_EQSP) {
            hist_eqsp (h, vs[k]);
            h->key_lut.clear ();
            continue;
        }
        if (!hip_plugin_load (&last_error)) return false;
        const plm_img_desc d = vs[k]->desc (PLM_STORE_F32);
        h->key_lut.assign (PLM_HIP_VOPT_KEYS, 0);
        if (g_plugin.plm_hip_hist_vopt (&d, h->bins, &h->offset, &h->delta, h->key_lut.data (), gpuid) < 0) {
            last_error = hip_error ();
            return false;
        }
    }
    return true;
}

plm_hist_desc
Joint_histogram::desc () const
{
    plm_hist_desc d;
    d.fixed_bins = fixed.bins;
    d.moving_bins = moving.bins;
    d.fixed_offset = fixed.offset;
    d.fixed_delta = fixed.delta;
    d.moving_offset = moving.offset;
    d.moving_delta = moving.delta;
    return d;
}

/* ---------------------------------------------------------------------
   Parms, score, state
   --------------------------------------------------------------------- */

Bspline_parms::Bspline_parms ()
    : threading (BTHR_HIP), implementation ('g'), gpuid (0), storage (PLM_STORE_F32),
      grad_mode (PLM_GRAD_REF_NN), mi_hist_fixed_bins (32), mi_hist_moving_bins (32),
      mi_fixed_image_minVal (0), mi_fixed_image_maxVal (0), mi_moving_image_minVal (0),
      mi_moving_image_maxVal (0), max_its (10), max_feval (10)
{
}

Bspline_score::Bspline_score ()
    : total_score (0), num_coeff (0), curr_smetric (0), curr_num_vox (0), curr_smetric_f64 (0)
{
}

void
Bspline_score::set_num_coeff (plm_long n)
{
    num_coeff = n;
    total_grad.assign ((size_t) n, 0.f);
    curr_smetric_grad.assign ((size_t) n, 0.f);
    curr_smetric_grad_f64.assign ((size_t) n, 0.0);
}

/* src/register/bspline_score.cxx:52-64 */
void
Bspline_score::reset_score ()
{
    total_score = 0;
    for (auto& g : total_grad) g = 0.f;
    metric_record.clear ();
    curr_num_vox = 0;
    curr_smetric = 0;
    for (auto& g : curr_smetric_grad) g = 0.f;
}

/* total += lambda * curr, then curr is cleared
   (src/register/bspline_score.cxx:66-77). */
void
Bspline_score::accumulate (float lambda)
{
    total_score += lambda * curr_smetric;
    for (plm_long i = 0; i < num_coeff; i++)
        total_grad[i] += lambda * curr_smetric_grad[i];
    curr_smetric = 0;
    curr_num_vox = 0;
    for (auto& g : curr_smetric_grad) g = 0.f;
}

Bspline_state::Bspline_state ()
    : it (0), feval (0), dev_ptrs (nullptr), metric_type (SIMILARITY_METRIC_MSE),
      metric_lambda (1.0f), fixed (nullptr), moving (nullptr), moving_grad (nullptr),
      mi_hist (nullptr), fixed_roi (nullptr), moving_roi (nullptr)
{
}

Bspline_state::~Bspline_state ()
{
    if (dev_ptrs && g_plugin.plm_hip_bspline_destroy)
        g_plugin.plm_hip_bspline_destroy ((plm_hip_ctx*) dev_ptrs);
}

int
bspline_state_initialize_hip (Bspline_state* bst, Bspline_xform* bxf, Bspline_parms* parms)
{
    std::string err;
    if (!hip_plugin_load (&err)) {
        bst->last_error = err;
        return PLM_HIP_ERR_NO_DEVICE;
    }
    bst->ssd.set_num_coeff (bxf->num_coeff);
    if (bst->dev_ptrs) {
        g_plugin.plm_hip_bspline_destroy ((plm_hip_ctx*) bst->dev_ptrs);
        bst->dev_ptrs = nullptr;
    }
    plm_bsp_geom g = bxf->geom ();
    plm_img_desc fd = bst->fixed->desc (parms->storage);
    plm_img_desc md = bst->moving->desc (parms->storage);
    plm_img_desc gd;
    if (bst->moving_grad) gd = bst->moving_grad->desc (parms->storage);
    plm_hist_desc hd;
    if (bst->mi_hist) hd = bst->mi_hist->desc ();
    plm_hip_ctx* ctx = nullptr;
    int rc = g_plugin.plm_hip_bspline_create (&g, &fd, &md, bst->moving_grad ? &gd : nullptr,
        bst->mi_hist ? &hd : nullptr, parms->gpuid, &ctx);
    if (rc < 0) {
        bst->last_error = hip_error ();
        return rc;
    }
    bst->dev_ptrs = ctx;
    if (bst->mi_hist && (!bst->mi_hist->fixed.key_lut.empty () || !bst->mi_hist->moving.key_lut.empty ())) {
        const Joint_histogram* jh = bst->mi_hist;
        rc = g_plugin.plm_hip_bspline_set_hist_keys (ctx,
            jh->fixed.key_lut.empty () ? nullptr : jh->fixed.key_lut.data (),
            jh->moving.key_lut.empty () ? nullptr : jh->moving.key_lut.data ());
        if (rc < 0) {
            bst->last_error = hip_error ();
            return rc;
        }
    }
    if (bst->fixed_roi || bst->moving_roi) {
        plm_img_desc fr, mr;
        if (bst->fixed_roi) fr = bst->fixed_roi->desc ();
        if (bst->moving_roi) mr = bst->moving_roi->desc ();
        rc = g_plugin.plm_hip_bspline_set_roi (ctx, bst->fixed_roi ? &fr : nullptr,
            bst->moving_roi ? &mr : nullptr);
        if (rc < 0) {
            bst->last_error = hip_error ();
            return rc;
        }
    }
    return PLM_HIP_OK;
}

int
volume_gradient_hip (const Volume* vol, Volume* out, int gpuid)
{
    std::string err;
    if (!hip_plugin_load (&err)) return PLM_HIP_ERR_NO_DEVICE;
    plm_img_desc d = vol->desc (PLM_STORE_F32);
    return g_plugin.plm_hip_volume_gradient (&d, out->img.data (), PLM_MEM_HOST, gpuid);
}

/* ---------------------------------------------------------------------
   Score dispatch
   --------------------------------------------------------------------- */

/* The BTHR_HIP branch the reference's bspline_score_mse / bspline_score_mi
   dispatchers gain (src/register/bspline_mse.cxx:1877-1963,
   src/register/bspline_mi.cxx:3035-3141). */
static int
bspline_score_hip (Bspline_optimize* bod, int metric)
{
    Bspline_parms* parms = bod->get_bspline_parms ();
    Bspline_state* bst = bod->get_bspline_state ();
    Bspline_xform* bxf = bod->get_bspline_xform ();
    Bspline_score* ssd = &bst->ssd;
    if (!bst->dev_ptrs) {
        int rc = bspline_state_initialize_hip (bst, bxf, parms);
        if (rc < 0) return rc;
    }
    std::vector<int64_t> hist;
    if (metric == PLM_METRIC_MI && bst->mi_hist)
        hist.resize ((size_t) (bst->mi_hist->fixed.bins * bst->mi_hist->moving.bins
                + bst->mi_hist->fixed.bins));
    double score = 0.0;
    int64_t nvox = 0;
    int rc = g_plugin.plm_hip_bspline_score ((plm_hip_ctx*) bst->dev_ptrs, metric,
        metric == PLM_METRIC_MSE ? parms->grad_mode : PLM_GRAD_REF_NN, bxf->coeff.data (),
        &score, ssd->curr_smetric_grad_f64.data (), &nvox, hist.empty () ? nullptr : hist.data ());
    if (rc < 0) {
        bst->last_error = hip_error ();
        return rc;
    }
    ssd->curr_smetric_f64 = score;
    ssd->curr_smetric = (float) score;
    ssd->curr_num_vox = nvox;
    for (plm_long i = 0; i < bxf->num_coeff; i++)
        ssd->curr_smetric_grad[i] = (float) ssd->curr_smetric_grad_f64[i];
    if (!hist.empty ()) {
        Joint_histogram* jh = bst->mi_hist;
        const plm_long fb = jh->fixed.bins, mb = jh->moving.bins;
        for (auto& v : jh->m_hist) v = 0.0;
        for (plm_long a = 0; a < fb; a++) {
            jh->f_hist[a] = (double) hist[fb * mb + a];
            for (plm_long b = 0; b < mb; b++) {
                double j = (double) hist[a * mb + b] / 4294967296.0;
                jh->j_hist[a * mb + b] = j;
                jh->m_hist[b] += j;
            }
        }
    }
    return rc;
}

/* Per evaluation: reset, one metric, record, accumulate, report
   (src/register/bspline.cxx:558-632). */
int
bspline_score (Bspline_optimize* bod)
{
    Bspline_parms* parms = bod->get_bspline_parms ();
    Bspline_state* bst = bod->get_bspline_state ();
    bst->ssd.reset_score ();
    if (parms->threading != BTHR_HIP) {
        bst->last_error = "only threading=hip is built in this library";
        return PLM_HIP_ERR_UNSUPPORTED;
    }
    auto t0 = std::chrono::steady_clock::now ();
    int rc = bspline_score_hip (bod, bst->metric_type == SIMILARITY_METRIC_MI_MATTES
        ? PLM_METRIC_MI : PLM_METRIC_MSE);
    if (rc < 0) return rc;
    double t = std::chrono::duration<double> (std::chrono::steady_clock::now () - t0).count ();
    Metric_score ms;
    ms.score = bst->ssd.curr_smetric;
    ms.time = t;
    ms.num_vox = bst->ssd.curr_num_vox;
    bst->ssd.metric_record.push_back (ms);
    bst->ssd.accumulate (bst->metric_lambda);
    return rc;
}

/* ---------------------------------------------------------------------
   Optimizer: steepest descent along -grad/|grad| with a trust region
   updated from the gain ratio against the linear model (the rules of
   src/register/bspline_optimize_steepest.cxx): shrink on increase, shrink
   and accept below 0.25, grow x3 above 0.75.
   --------------------------------------------------------------------- */

int
Bspline_optimize::optimize ()
{
    Bspline_score* ssd = &bst->ssd;
    const plm_long n = bxf->num_coeff;
    std::vector<float> x (bxf->coeff), h (n);
    float alpha = 1.0f;
    bst->it = 0;
    bst->feval = 0;
    int rc = bspline_score (this);
    if (rc < 0) return rc;
    float old_score = ssd->total_score;
    for (;;) {
        double gnorm = 0.0;
        for (plm_long i = 0; i < n; i++) gnorm += (double) ssd->total_grad[i] * ssd->total_grad[i];
        gnorm = sqrt (gnorm);
        if (gnorm == 0.0) break;
        double htg = 0.0;
        for (plm_long i = 0; i < n; i++) {
            h[i] = (float) (-ssd->total_grad[i] / gnorm);
            htg -= h[i] * ssd->total_grad[i];
        }
        bool accepted = false;
        while (!accepted && bst->feval < parms->max_feval) {
            bst->feval++;
            for (plm_long i = 0; i < n; i++) bxf->coeff[i] = x[i] + alpha * h[i];
            rc = bspline_score (this);
            if (rc < 0) return rc;
            double gr = (old_score - ssd->total_score) / htg;
            if (gr < 0) {
                alpha *= 0.5f;
            } else {
                accepted = true;
                if (gr < 0.25) alpha *= 0.5f;
                else if (gr > 0.75) alpha *= 3.0f;
            }
        }
        if (!accepted) {
            bxf->coeff = x;
            break;
        }
        bst->it++;
        x = bxf->coeff;
        old_score = ssd->total_score;
        if (bst->it >= parms->max_its || bst->feval >= parms->max_feval) break;
    }
    return PLM_HIP_OK;
}
