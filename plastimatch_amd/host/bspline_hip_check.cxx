/* bspline_hip_check: score, gradient and gradient check of the
 * threading=hip flavor from the command line.
 *
 * Counterpart of the reference's check_grad tool
 * (src/standalone/check_grad.cxx:81-273, options :309-345): loads a fixed
 * and a moving image (MetaImage .mha, float, uncompressed), builds the
 * B-spline grid, optionally random coefficients, and evaluates the score and
 * gradient through bspline_score().  --process fwd|ctr compares one
 * gradient component with a forward or central difference (the reference
 * parses "ctr" but always does a forward difference, check_grad.cxx:234-262;
 * here ctr is the central difference), --process line scans the score along
 * the gradient, --process optimize runs the steepest-descent loop.  Output
 * is one JSON object on stdout.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <fstream>
#include <random>
#include <sstream>
#include <string>

#include "bspline_hip_host.h"

static void
usage ()
{
    fprintf (stderr,
        "usage: bspline_hip_check [options] fixed.mha moving.mha\n"
        "       bspline_hip_check --xf-roundtrip in.txt out.txt   (MGH_GPUIT_BSP read / write)\n"
        "       bspline_hip_check --mha-roundtrip in.mha out.mha  (MetaImage read / write as float)\n"
        "  -M, --metric mse|mi          similarity metric (mse)\n"
        "  -f, --flavor g|p             MSE dc/dv: g = nearest-voxel gradient, p = exact (g)\n"
        "  -p, --process fwd|ctr|line|score|optimize   (score)\n"
        "  -e, --step h                 finite-difference step in mm (1e-4)\n"
        "  -i, --index k                coefficient index checked (514)\n"
        "  -l, --line-range a b         line scan range in steps (0 30)\n"
        "  -R, --random-start lo hi     uniform random coefficients\n"
        "      --seed s                 random seed (0)\n"
        "  -X, --input-xform file       coefficients, float32 raw, interleaved\n"
        "  -g, --grid-spac v            vox_per_rgn in all three axes (15)\n"
        "      --bins n                 MI histogram bins (32)\n"
        "      --hist eqsp|vopt         MI histogram type (eqsp)\n"
        "      --fixed-roi f.mha        fixed region of interest (uchar MetaImage)\n"
        "      --moving-roi m.mha       moving region of interest (uchar MetaImage)\n"
        "      --fixed-range lo hi      MI fixed intensity range (mi_fixed_image_minVal/maxVal)\n"
        "      --moving-range lo hi     MI moving intensity range\n"
        "      --bf16                   keep the images in HBM as bfloat16\n"
        "      --iterations n           optimize: max iterations (10)\n"
        "      --gpuid n                device (0)\n"
        "  -O, --output file            write the gradient (float64 raw)\n");
}

/* The input images: read_mha of the host library (the reference's
   MetaImage reading, widened to float). */
static bool
read_input (const std::string& fn, Volume** out, std::string* err)
{
    Volume* v = read_mha (fn, err);
    if (!v) return false;
    if (v->vox_planes != 1 || v->dim[0] < 2 || v->dim[1] < 2 || v->dim[2] < 2) {
        *err = fn + ": need a scalar image with every dim >= 2";
        delete v;
        return false;
    }
    *out = v;
    return true;
}

/* A region of interest: a MetaImage read as float, one byte a voxel. */
static Volume_uchar*
read_roi (const std::string& fn, std::string* err)
{
    Volume* v = read_mha (fn, err);
    if (!v) return nullptr;
    Volume_uchar* r = nullptr;
    if (v->vox_planes != 1) {
        *err = fn + ": a region of interest is a scalar image";
    } else {
        r = new Volume_uchar (v->dim, v->origin, v->spacing, v->direction_cosines);
        for (plm_long p = 0; p < r->npix; p++) r->img[(size_t) p] = (uint8_t) v->img[(size_t) p];
    }
    delete v;
    return r;
}

int
main (int argc, char** argv)
{
    /* --mha-roundtrip in out: read a MetaImage and write it back as float
       (no GPU needed) */
    if (argc == 4 && std::string (argv[1]) == "--mha-roundtrip") {
        std::string err;
        Volume* v = read_mha (argv[2], &err);
        if (!v) {
            fprintf (stderr, "%s\n", err.c_str ());
            return 1;
        }
        double sum = 0.0;
        for (float f : v->img) sum += f;
        printf ("dim %ld %ld %ld planes %d sum %.6f\n", (long) v->dim[0], (long) v->dim[1], (long) v->dim[2],
            v->vox_planes, sum);
        const bool ok = write_mha (argv[3], *v, &err);
        delete v;
        if (!ok) fprintf (stderr, "%s\n", err.c_str ());
        return ok ? 0 : 1;
    }
    /* --xf-roundtrip in out: read an MGH_GPUIT_BSP file and write it back
       (no GPU needed) */
    if (argc == 4 && std::string (argv[1]) == "--xf-roundtrip") {
        std::string err;
        Bspline_xform* bxf = bspline_xform_load (argv[2], &err);
        if (!bxf) {
            fprintf (stderr, "%s\n", err.c_str ());
            return 1;
        }
        printf ("img_dim %ld %ld %ld vox_per_rgn %ld %ld %ld num_coeff %ld\n", (long) bxf->img_dim[0],
            (long) bxf->img_dim[1], (long) bxf->img_dim[2], (long) bxf->vox_per_rgn[0],
            (long) bxf->vox_per_rgn[1], (long) bxf->vox_per_rgn[2], (long) bxf->num_coeff);
        const bool ok = bxf->save (argv[3]);
        delete bxf;
        return ok ? 0 : 1;
    }
    std::string metric = "mse", flavor = "g", process = "score", output, xform_fn, froi_fn, mroi_fn;
    float frange[2] = {0.f, 0.f}, mrange[2] = {0.f, 0.f};
    std::string hist_type = "eqsp";
    float step = 1e-4f, rlo = 0.f, rhi = 0.f;
    long index = 514;
    int line_a = 0, line_b = 30, seed = 0, vpr = 15, bins = 32, gpuid = 0, iters = 10;
    bool random_start = false, bf16 = false;
    std::string files[2];
    int nfiles = 0;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&] () -> const char* {
            if (i + 1 >= argc) { usage (); exit (2); }
            return argv[++i];
        };
        if (a == "-M" || a == "--metric") metric = next ();
        else if (a == "-f" || a == "--flavor") flavor = next ();
        else if (a == "-p" || a == "--process") process = next ();
        else if (a == "-e" || a == "--step") step = (float) atof (next ());
        else if (a == "-i" || a == "--index") index = atol (next ());
        else if (a == "-l" || a == "--line-range") { line_a = atoi (next ()); line_b = atoi (next ()); }
        else if (a == "-R" || a == "--random-start") { random_start = true; rlo = (float) atof (next ()); rhi = (float) atof (next ()); }
        else if (a == "--seed") seed = atoi (next ());
        else if (a == "-X" || a == "--input-xform") xform_fn = next ();
        else if (a == "-g" || a == "--grid-spac") vpr = atoi (next ());
        else if (a == "--bins") bins = atoi (next ());
        else if (a == "--fixed-roi") froi_fn = next ();
        else if (a == "--moving-roi") mroi_fn = next ();
        else if (a == "--hist") hist_type = next ();
        else if (a == "--fixed-range") { frange[0] = (float) atof (next ()); frange[1] = (float) atof (next ()); }
        else if (a == "--moving-range") { mrange[0] = (float) atof (next ()); mrange[1] = (float) atof (next ()); }
        else if (a == "--bf16") bf16 = true;
        else if (a == "--iterations") iters = atoi (next ());
        else if (a == "--gpuid") gpuid = atoi (next ());
        else if (a == "-O" || a == "--output") output = next ();
        else if (a == "-h" || a == "--help") { usage (); return 0; }
        else if (nfiles < 2) files[nfiles++] = a;
        else { usage (); return 2; }
    }
    if (nfiles != 2) { usage (); return 2; }

    std::string err;
    Volume *fixed = nullptr, *moving = nullptr;
    if (!read_input (files[0], &fixed, &err) || !read_input (files[1], &moving, &err)) {
        printf ("{\"status\": %d, \"error\": \"%s\"}\n", PLM_HIP_ERR_ARG, err.c_str ());
        return 1;
    }
    Bspline_parms parms;
    parms.gpuid = gpuid;
    parms.storage = bf16 ? PLM_STORE_BF16 : PLM_STORE_F32;
    parms.grad_mode = flavor == "p" ? PLM_GRAD_EXACT : PLM_GRAD_REF_NN;
    parms.implementation = flavor[0];
    parms.max_its = iters;
    parms.max_feval = 3 * iters + 3;
    parms.mi_hist_fixed_bins = parms.mi_hist_moving_bins = bins;
    parms.mi_fixed_image_minVal = frange[0];
    parms.mi_fixed_image_maxVal = frange[1];
    parms.mi_moving_image_minVal = mrange[0];
    parms.mi_moving_image_maxVal = mrange[1];
    /* the stage's regions of interest (src/register/bspline_stage.cxx:129-192) */
    Volume_uchar *froi = nullptr, *mroi = nullptr;
    std::vector<Volume_uchar*> made;
    if ((!froi_fn.empty () && !(froi = read_roi (froi_fn, &err)))
        || (!mroi_fn.empty () && !(mroi = read_roi (mroi_fn, &err)))) {
        printf ("{\"status\": %d, \"error\": \"%s\"}\n", PLM_HIP_ERR_ARG, err.c_str ());
        return 1;
    }
    if (!bspline_stage_rois (&parms, fixed, moving, &froi, &mroi, &made)) {
        printf ("{\"status\": %d, \"error\": \"a ROI and its image differ in dims\"}\n", PLM_HIP_ERR_ARG);
        return 1;
    }

    Bspline_xform bxf;
    const plm_long zero[3] = {0, 0, 0};
    const plm_long vprs[3] = {vpr, vpr, vpr};
    bxf.initialize (fixed->origin, fixed->spacing, fixed->dim, zero, fixed->dim, vprs,
        fixed->direction_cosines);
    if (random_start) {
        std::mt19937 gen (seed);
        std::uniform_real_distribution<float> u (rlo, rhi);
        for (auto& c : bxf.coeff) c = u (gen);
    }
    if (!xform_fn.empty ()) {
        FILE* fp = fopen (xform_fn.c_str (), "rb");
        size_t got = fp ? fread (bxf.coeff.data (), sizeof (float), bxf.coeff.size (), fp) : 0;
        if (fp) fclose (fp);
        if (got != bxf.coeff.size ()) {
            printf ("{\"status\": %d, \"error\": \"cannot read %zu coefficients from %s\"}\n",
                PLM_HIP_ERR_ARG, bxf.coeff.size (), xform_fn.c_str ());
            return 1;
        }
    }

    Bspline_state bst;
    bst.fixed = fixed;
    bst.moving = moving;
    bst.fixed_roi = froi;
    bst.moving_roi = mroi;
    Volume* mgrad = nullptr;
    Joint_histogram* jh = nullptr;
    if (metric == "mi") {
        bst.metric_type = SIMILARITY_METRIC_MI_MATTES;
        jh = new Joint_histogram (hist_type == "vopt" ? HIST_VOPT : HIST_EQSP, bins, bins);
        if (!jh->initialize (fixed, moving, gpuid)) {
            printf ("{\"status\": -1, \"error\": \"histogram: %s\"}\n", jh->last_error.c_str ());
            return 1;
        }
        bst.mi_hist = jh;
        bxf.jitter_if_zero ();          /* src/register/bspline_optimize.cxx:105-107 */
    } else if (parms.grad_mode == PLM_GRAD_REF_NN) {
        mgrad = new Volume (moving->dim, moving->origin, moving->spacing,
            moving->direction_cosines, 3);
        int rc = volume_gradient_hip (moving, mgrad, gpuid);
        if (rc < 0) {
            std::string e2;
            if (!hip_plugin_load (&e2)) err = e2;
            printf ("{\"status\": %d, \"error\": \"volume_gradient: %s\"}\n", rc, err.c_str ());
            return 1;
        }
        bst.moving_grad = mgrad;
    }
    int rc = bspline_state_initialize_hip (&bst, &bxf, &parms);
    if (rc < 0) {
        printf ("{\"status\": %d, \"error\": \"%s\"}\n", rc, bst.last_error.c_str ());
        return 1;
    }
    Bspline_optimize bod (&parms, &bst, &bxf);

    printf ("{\"metric\": \"%s\", \"flavor\": \"%s\", \"num_coeff\": %lld", metric.c_str (),
        flavor.c_str (), (long long) bxf.num_coeff);
    if (process == "optimize") {
        rc = bspline_score (&bod);
        if (rc < 0) { printf (", \"status\": %d, \"error\": \"%s\"}\n", rc, bst.last_error.c_str ()); return 1; }
        float s0 = bst.ssd.total_score;
        rc = bod.optimize ();
        if (rc < 0) { printf (", \"status\": %d, \"error\": \"%s\"}\n", rc, bst.last_error.c_str ()); return 1; }
        rc = bspline_score (&bod);
        printf (", \"initial_score\": %.9g, \"final_score\": %.9g, \"iterations\": %d, \"fevals\": %d",
            s0, bst.ssd.total_score, bst.it, bst.feval);
    } else {
        rc = bspline_score (&bod);
        if (rc < 0) { printf (", \"status\": %d, \"error\": \"%s\"}\n", rc, bst.last_error.c_str ()); return 1; }
        Bspline_score& ssd = bst.ssd;
        const double s0 = ssd.curr_smetric_f64;
        std::vector<double> g0 (ssd.curr_smetric_grad_f64);
        printf (", \"score\": %.17g, \"num_vox\": %lld", s0, (long long) ssd.metric_record[0].num_vox);
        double gn = 0.0;
        for (double v : g0) gn += v * v;
        printf (", \"grad_norm\": %.17g", sqrt (gn));
        if (!output.empty ()) {
            FILE* fp = fopen (output.c_str (), "wb");
            if (fp) { fwrite (g0.data (), sizeof (double), g0.size (), fp); fclose (fp); }
        }
        if ((process == "fwd" || process == "ctr") && index >= 0 && index < bxf.num_coeff) {
            const float c0 = bxf.coeff[index];
            bxf.coeff[index] = c0 + step;
            bspline_score (&bod);
            double sp = bst.ssd.curr_smetric_f64;
            double fd;
            if (process == "ctr") {
                bxf.coeff[index] = c0 - step;
                bspline_score (&bod);
                fd = (sp - bst.ssd.curr_smetric_f64) / (2.0 * step);
            } else {
                fd = (sp - s0) / step;
            }
            bxf.coeff[index] = c0;
            printf (", \"index\": %ld, \"analytic\": %.17g, \"finite_difference\": %.17g", index,
                g0[index], fd);
        } else if (process == "line") {
            printf (", \"line\": [");
            std::vector<float> c0 (bxf.coeff);
            for (int k = line_a; k <= line_b; k++) {
                for (plm_long i = 0; i < bxf.num_coeff; i++)
                    bxf.coeff[i] = (float) (c0[i] + k * step * g0[i] / sqrt (gn));
                bspline_score (&bod);
                printf ("%s%.17g", k == line_a ? "" : ", ", bst.ssd.curr_smetric_f64);
            }
            bxf.coeff = c0;
            printf ("]");
        }
    }
    printf (", \"status\": %d}\n", rc);
    delete froi;        /* (what bspline_stage_rois made is one of the two) */
    delete mroi;
    delete mgrad;
    delete jh;
    delete fixed;
    delete moving;
    return rc < 0 ? 1 : 0;
}
