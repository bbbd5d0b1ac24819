/* CPU oracle for the B-spline score/gradient path (TEST INFRASTRUCTURE).
 *
 * A plain-C restatement of the reference CPU flavors that the HIP product
 * path is checked against.  Only tests, the smoke check and the CPU
 * baseline of bench.py load it; the product library never links it.
 *
 * Every function names the reference file:line it follows (paths are
 * relative to the Plastimatch source root).  Float code keeps the
 * reference's operation order and is compiled with -ffp-contract=off, so
 * each float expression rounds exactly as the reference's does.  The
 * *_f64 functions redo the same algorithm in double precision.
 *
 * Pinning (what fixes each function independently of this file):
 *   orc_basis_lut            closed-form cubic B-spline values, partition
 *                            of unity, linear-polynomial reproduction
 *   orc_interp_vf / tiles    constant and linear coefficient fields
 *   orc_volume_gradient      exact on quadratic images, edge replicate
 *   orc_mse_score 'g'/'c'    voxel-order scatter ('c') equals tile
 *                            condense ('g'); identical images => 0;
 *                            constant offset => offset^2; linear ramps
 *                            => central finite differences
 *   orc_mse_score 'p', f64   central finite differences of the score
 *   orc_mi_score             histogram mass, diagonal joint histogram and
 *                            S = -H(F) for identical images, finite
 *                            differences of the FP64 score
 */
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "plm_oracle.h"

/* ---------------------------------------------------------------------
   Geometry
   --------------------------------------------------------------------- */

typedef struct vol_t {
    int64_t dim[3];
    float origin[3];
    float spacing[3];
    float step[9];
    float proj[9];
    float inv_dc[9];
    const float* img;
} vol_t;

typedef struct bxf_t {
    float img_origin[3];
    float img_spacing[3];
    int64_t img_dim[3];
    int64_t roi_offset[3];
    int64_t roi_dim[3];
    int64_t vpr[3];
    int64_t rdims[3];
    int64_t cdims[3];
    int64_t num_knots;
    int64_t num_coeff;
    int64_t num_tiles;
    float* bx_lut;      /* vpr*4, bspline_basis_eval rounding */
    float* by_lut;
    float* bz_lut;
    float* q_lut;       /* vpr^3*64, Bspline_xform::allocate rounding */
    int64_t* c_lut;     /* tiles*64 */
    const float* coeff;
} bxf_t;

/* 3x3 inverse by cofactors.  Follows volume_matrix3x3inverse,
   src/base/direction_cosines.cxx:134-159. */
static void
mat3_inverse (float* out, const float* m)
{
    float det = m[0]*(m[4]*m[8]-m[5]*m[7])
        - m[1]*(m[3]*m[8]-m[5]*m[6])
        + m[2]*(m[3]*m[7]-m[4]*m[6]);
    out[0] =  (m[4]*m[8]-m[5]*m[7]) / det;
    out[1] = -(m[1]*m[8]-m[2]*m[7]) / det;
    out[2] =  (m[1]*m[5]-m[2]*m[4]) / det;
    out[3] = -(m[3]*m[8]-m[5]*m[6]) / det;
    out[4] =  (m[0]*m[8]-m[2]*m[6]) / det;
    out[5] = -(m[0]*m[5]-m[2]*m[3]) / det;
    out[6] =  (m[3]*m[7]-m[4]*m[6]) / det;
    out[7] = -(m[0]*m[7]-m[1]*m[6]) / det;
    out[8] =  (m[0]*m[4]-m[1]*m[3]) / det;
}

/* step = dc*spacing (column scale), proj = inv(dc)/spacing (row scale).
   Follows compute_direction_matrices, src/base/direction_matrices.cxx:10-20. */
void
orc_direction_matrices (const float dc[9], const float spacing[3],
    float step[9], float proj[9], float inv_dc[9])
{
    mat3_inverse (inv_dc, dc);
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            step[3*i+j] = dc[3*i+j] * spacing[j];
            proj[3*i+j] = inv_dc[3*i+j] / spacing[i];
        }
    }
}

static void
vol_init (vol_t* v, const orc_vol_hdr* h, const float* img)
{
    for (int d = 0; d < 3; d++) {
        v->dim[d] = h->dim[d];
        v->origin[d] = h->origin[d];
        v->spacing[d] = h->spacing[d];
    }
    orc_direction_matrices (h->dc, h->spacing, v->step, v->proj, v->inv_dc);
    v->img = img;
}

/* rdims = 1 + (roi_dim-1)/vpr, cdims = rdims + 3.
   Follows Bspline_header::set, src/base/bspline_header.cxx:28-62. */
int64_t
orc_header (const orc_geom* g, int64_t rdims[3], int64_t cdims[3])
{
    for (int d = 0; d < 3; d++) {
        rdims[d] = 1 + (g->roi_dim[d] - 1) / g->vox_per_rgn[d];
        cdims[d] = 3 + rdims[d];
    }
    return cdims[0] * cdims[1] * cdims[2];
}

/* Per-axis cubic B-spline basis, lut[i*4+t] = B_t(i/vpr).  The argument is
   rounded to float and the polynomial evaluated in double, then rounded to
   float.  Follows bspline_basis_eval, src/base/bspline_xform.cxx:73-99. */
void
orc_basis_lut (int64_t vpr, float* lut)
{
    for (int64_t i = 0; i < vpr; i++) {
        float u = (float) i / (float) (int) vpr;
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

/* Per-axis basis as computed inside Bspline_xform::allocate, with t3, t2
   rounded to float before the double polynomial
   (src/base/bspline_xform.cxx:388-420).  Only q_lut uses this rounding. */
static void
alloc_basis (int64_t vpr, float* A)
{
    for (int64_t i = 0; i < vpr; i++) {
        float ii = ((float) i) / (float) vpr;
        float t3 = ii*ii*ii;
        float t2 = ii*ii;
        float t1 = ii;
        A[i*4+0] = (float) ((1.0/6.0) * (-1.0 * t3 + 3.0 * t2 - 3.0 * t1 + 1.0));
        A[i*4+1] = (float) ((1.0/6.0) * (+3.0 * t3 - 6.0 * t2 + 4.0));
        A[i*4+2] = (float) ((1.0/6.0) * (-3.0 * t3 + 3.0 * t2 + 3.0 * t1 + 1.0));
        A[i*4+3] = (float) ((1.0/6.0) * (+1.0 * t3));
    }
}

/* Builds the header fields and the q/c/b LUTs.
   Follows Bspline_xform::allocate, src/base/bspline_xform.cxx:361-488. */
static int
bxf_init (bxf_t* b, const orc_geom* g, const float* coeff)
{
    memset (b, 0, sizeof (*b));
    for (int d = 0; d < 3; d++) {
        b->img_origin[d] = g->img_origin[d];
        b->img_spacing[d] = g->img_spacing[d];
        b->img_dim[d] = g->img_dim[d];
        b->roi_offset[d] = g->roi_offset[d];
        b->roi_dim[d] = g->roi_dim[d];
        b->vpr[d] = g->vox_per_rgn[d];
        if (b->vpr[d] < 1 || b->roi_dim[d] < 1) return -1;
    }
    b->num_knots = orc_header (g, b->rdims, b->cdims);
    b->num_coeff = 3 * b->num_knots;
    b->num_tiles = b->rdims[0] * b->rdims[1] * b->rdims[2];
    b->coeff = coeff;

    b->bx_lut = (float*) malloc (sizeof(float) * 4 * b->vpr[0]);
    b->by_lut = (float*) malloc (sizeof(float) * 4 * b->vpr[1]);
    b->bz_lut = (float*) malloc (sizeof(float) * 4 * b->vpr[2]);
    orc_basis_lut (b->vpr[0], b->bx_lut);
    orc_basis_lut (b->vpr[1], b->by_lut);
    orc_basis_lut (b->vpr[2], b->bz_lut);

    float* A = (float*) malloc (sizeof(float) * 4 * b->vpr[0]);
    float* B = (float*) malloc (sizeof(float) * 4 * b->vpr[1]);
    float* C = (float*) malloc (sizeof(float) * 4 * b->vpr[2]);
    alloc_basis (b->vpr[0], A);
    alloc_basis (b->vpr[1], B);
    alloc_basis (b->vpr[2], C);
    int64_t nq = b->vpr[0] * b->vpr[1] * b->vpr[2];
    b->q_lut = (float*) malloc (sizeof(float) * 64 * nq);
    int64_t p = 0;
    for (int64_t k = 0; k < b->vpr[2]; k++)
        for (int64_t j = 0; j < b->vpr[1]; j++)
            for (int64_t i = 0; i < b->vpr[0]; i++)
                for (int tz = 0; tz < 4; tz++)
                    for (int ty = 0; ty < 4; ty++)
                        for (int tx = 0; tx < 4; tx++)
                            b->q_lut[p++] = A[i*4+tx] * B[j*4+ty] * C[k*4+tz];
    free (A); free (B); free (C);

    b->c_lut = (int64_t*) malloc (sizeof(int64_t) * 64 * b->num_tiles);
    p = 0;
    for (int64_t k = 0; k < b->rdims[2]; k++)
        for (int64_t j = 0; j < b->rdims[1]; j++)
            for (int64_t i = 0; i < b->rdims[0]; i++)
                for (int tz = 0; tz < 4; tz++)
                    for (int ty = 0; ty < 4; ty++)
                        for (int tx = 0; tx < 4; tx++)
                            b->c_lut[p++] = (k + tz) * b->cdims[0] * b->cdims[1]
                                + (j + ty) * b->cdims[0] + (i + tx);
    return 0;
}

static void
bxf_free (bxf_t* b)
{
    free (b->bx_lut); free (b->by_lut); free (b->bz_lut);
    free (b->q_lut); free (b->c_lut);
}

/* A uchar region-of-interest volume (Bspline_state::fixed_roi /
   moving_roi), with its own geometry. */
typedef struct roi_t {
    int64_t dim[3];
    float origin[3];
    float proj[9];
    double dorigin[3];
    double dproj[9];
    const uint8_t* img;
} roi_t;

static void dproj_of (double* proj, const orc_vol_hdr* h);

static const roi_t*
roi_init (roi_t* r, const orc_vol_hdr* h, const uint8_t* img)
{
    if (!h || !img) return 0;
    vol_t v;
    vol_init (&v, h, 0);
    for (int d = 0; d < 3; d++) {
        r->dim[d] = v.dim[d];
        r->origin[d] = v.origin[d];
        r->dorigin[d] = h->origin[d];
    }
    for (int i = 0; i < 9; i++) r->proj[i] = v.proj[i];
    dproj_of (r->dproj, h);
    r->img = img;
    return r;
}

/* 1 if xyz (mm) falls on a non-zero ROI voxel: index = floor of the
   projected position, clamped into the volume (below 0 -> 0, at or past dim
   -> dim - 1).  Follows inside_roi and float_to_plm_long_clamp,
   src/util/bspline_correspond.cxx:16-74, with PROJECT_X/Y/Z,
   src/base/volume_macros.h:75-80. */
static inline int
inside_roi (const float xyz[3], const roi_t* r)
{
    float tmp[3], pf[3];
    int64_t p[3];
    tmp[0] = xyz[0] - r->origin[0];
    tmp[1] = xyz[1] - r->origin[1];
    tmp[2] = xyz[2] - r->origin[2];
    for (int d = 0; d < 3; d++)
        pf[d] = tmp[0] * r->proj[3*d+0] + tmp[1] * r->proj[3*d+1] + tmp[2] * r->proj[3*d+2];
    for (int d = 0; d < 3; d++) {
        if (pf[d] < 0.f) p[d] = 0;
        else if (pf[d] >= (float) r->dim[d]) p[d] = r->dim[d] - 1;
        else p[d] = (int64_t) floor (pf[d]);
    }
    return (int) r->img[p[0] + r->dim[0] * (p[1] + r->dim[1] * p[2])];
}

/* The same test in double (the FP64 yardstick's positions). */
static inline int
inside_roi_d (const double xyz[3], const roi_t* r)
{
    double tmp[3], pf[3];
    int64_t p[3];
    for (int d = 0; d < 3; d++) tmp[d] = xyz[d] - r->dorigin[d];
    for (int d = 0; d < 3; d++)
        pf[d] = tmp[0] * r->dproj[3*d+0] + tmp[1] * r->dproj[3*d+1] + tmp[2] * r->dproj[3*d+2];
    for (int d = 0; d < 3; d++) {
        if (pf[d] < 0.0) p[d] = 0;
        else if (pf[d] >= (double) r->dim[d]) p[d] = r->dim[d] - 1;
        else p[d] = (int64_t) floor (pf[d]);
    }
    return (int) r->img[p[0] + r->dim[0] * (p[1] + r->dim[1] * p[2])];
}

/* ---------------------------------------------------------------------
   Per-voxel primitives (float)
   --------------------------------------------------------------------- */

/* Displacement d = sum_{k,j,i} (bx[i]*(by[j]*bz[k])) * coeff, summed in
   k-j-i order.  Follows bspline_interp_pix_c,
   src/base/bspline_interpolate.cxx:78-110. */
static inline void
interp_pix_c (float out[3], const bxf_t* b, int64_t pidx, const int64_t q[3])
{
    const int64_t* c_lut = &b->c_lut[pidx*64];
    const float* bx = &b->bx_lut[q[0]*4];
    const float* by = &b->by_lut[q[1]*4];
    const float* bz = &b->bz_lut[q[2]*4];
    out[0] = out[1] = out[2] = 0;
    int m = 0;
    for (int k = 0; k < 4; k++) {
        float C = bz[k];
        for (int j = 0; j < 4; j++) {
            float B = by[j] * C;
            for (int i = 0; i < 4; i++) {
                float A = bx[i] * B;
                int64_t cidx = 3 * c_lut[m++];
                out[0] += A * b->coeff[cidx+0];
                out[1] += A * b->coeff[cidx+1];
                out[2] += A * b->coeff[cidx+2];
            }
        }
    }
}

/* Same sum with the precomputed q_lut products.  Follows
   bspline_interp_pix_b, src/base/bspline_interpolate.cxx:50-76. */
static inline void
interp_pix_b (float out[3], const bxf_t* b, int64_t pidx, int64_t qidx)
{
    const float* q_lut = &b->q_lut[qidx*64];
    const int64_t* c_lut = &b->c_lut[pidx*64];
    out[0] = out[1] = out[2] = 0;
    for (int m = 0; m < 64; m++) {
        int64_t cidx = 3 * c_lut[m];
        out[0] += q_lut[m] * b->coeff[cidx+0];
        out[1] += q_lut[m] * b->coeff[cidx+1];
        out[2] += q_lut[m] * b->coeff[cidx+2];
    }
}

/* mxyz = fxyz + d; mijk = proj * (mxyz - origin); inside iff every
   coordinate lies in (-0.5, dim-0.5).  Follows
   bspline_find_correspondence_dcos, src/util/bspline_correspond.cxx:80-109,
   and Volume::is_inside, src/base/volume.cxx:899-906. */
static inline int
find_correspondence (float mijk[3], const float fxyz[3], const float d[3],
    const vol_t* mv)
{
    float mxyz[3], t[3];
    mxyz[0] = fxyz[0] + d[0];
    mxyz[1] = fxyz[1] + d[1];
    mxyz[2] = fxyz[2] + d[2];
    t[0] = mxyz[0] - mv->origin[0];
    t[1] = mxyz[1] - mv->origin[1];
    t[2] = mxyz[2] - mv->origin[2];
    mijk[0] = t[0] * mv->proj[0] + t[1] * mv->proj[1] + t[2] * mv->proj[2];
    mijk[1] = t[0] * mv->proj[3] + t[1] * mv->proj[4] + t[2] * mv->proj[5];
    mijk[2] = t[0] * mv->proj[6] + t[1] * mv->proj[7] + t[2] * mv->proj[8];
    for (int d2 = 0; d2 < 3; d2++) {
        if (mijk[d2] <= -0.5 || mijk[d2] >= mv->dim[d2] - 0.5) return 0;
    }
    return 1;
}

/* Clamped floor/round/fractions.  ROUND is round-half-away-from-zero.
   Follows li_clamp, src/base/interpolate.cxx:17-41, and ROUND_PLM_LONG,
   src/sys/plm_math.h:53-54. */
static inline void
li_clamp (float ma, int64_t dmax, int64_t* maf, int64_t* mar,
    float* fa1, float* fa2)
{
    if (ma < 0.f) {
        *maf = 0;
        *mar = 0;
        *fa2 = 0.0f;
    } else if (ma >= dmax) {
        *maf = dmax - 1;
        *mar = dmax;
        *fa2 = 1.0f;
    } else {
        *maf = (int64_t) floor (ma);
        *mar = (ma >= 0) ? (int64_t) (ma + 0.5) : (int64_t) (-(-ma + 0.5));
        *fa2 = ma - *maf;
    }
    *fa1 = 1.0f - *fa2;
}

static inline void
li_clamp_3d (const float mijk[3], int64_t f[3], int64_t r[3],
    float l1[3], float l2[3], const vol_t* mv)
{
    for (int d = 0; d < 3; d++)
        li_clamp (mijk[d], mv->dim[d] - 1, &f[d], &r[d], &l1[d], &l2[d]);
}

/* Trilinear value, corners in x-fastest order, products and sum left to
   right.  Follows li_value, src/base/interpolate.cxx:135-160. */
static inline float
li_value (const float f1[3], const float f2[3], int64_t mvf,
    const float* m, const int64_t dim[3])
{
    int64_t sx = dim[0], sxy = dim[0] * dim[1];
    float v0 = f1[0] * f1[1] * f1[2] * m[mvf];
    float v1 = f2[0] * f1[1] * f1[2] * m[mvf+1];
    float v2 = f1[0] * f2[1] * f1[2] * m[mvf+sx];
    float v3 = f2[0] * f2[1] * f1[2] * m[mvf+sx+1];
    float v4 = f1[0] * f1[1] * f2[2] * m[mvf+sxy];
    float v5 = f2[0] * f1[1] * f2[2] * m[mvf+sxy+1];
    float v6 = f1[0] * f2[1] * f2[2] * m[mvf+sxy+sx];
    float v7 = f2[0] * f2[1] * f2[2] * m[mvf+sxy+sx+1];
    return v0 + v1 + v2 + v3 + v4 + v5 + v6 + v7;
}

/* Analytic derivatives of the trilinear value times 1/spacing.  Follow
   li_value_dx/dy/dz, src/base/interpolate.cxx:162-244. */
static inline void
li_value_d (float out[3], const float f1[3], const float f2[3],
    const float inv[3], int64_t mvf, const float* m, const int64_t dim[3])
{
    int64_t sx = dim[0], sxy = dim[0] * dim[1];
    float a0 = m[mvf], a1 = m[mvf+1], a2 = m[mvf+sx], a3 = m[mvf+sx+1];
    float a4 = m[mvf+sxy], a5 = m[mvf+sxy+1], a6 = m[mvf+sxy+sx];
    float a7 = m[mvf+sxy+sx+1];
    float rx = inv[0], ry = inv[1], rz = inv[2];
    out[0] = -rx * f1[1] * f1[2] * a0 + rx * f1[1] * f1[2] * a1
        + -rx * f2[1] * f1[2] * a2 + rx * f2[1] * f1[2] * a3
        + -rx * f1[1] * f2[2] * a4 + rx * f1[1] * f2[2] * a5
        + -rx * f2[1] * f2[2] * a6 + rx * f2[1] * f2[2] * a7;
    out[1] = -f1[0] * ry * f1[2] * a0 + -f2[0] * ry * f1[2] * a1
        + f1[0] * ry * f1[2] * a2 + f2[0] * ry * f1[2] * a3
        + -f1[0] * ry * f2[2] * a4 + -f2[0] * ry * f2[2] * a5
        + f1[0] * ry * f2[2] * a6 + f2[0] * ry * f2[2] * a7;
    out[2] = -f1[0] * f1[1] * rz * a0 + -f2[0] * f1[1] * rz * a1
        + -f1[0] * f2[1] * rz * a2 + -f2[0] * f2[1] * rz * a3
        + f1[0] * f1[1] * rz * a4 + f2[0] * f1[1] * rz * a5
        + f1[0] * f2[1] * rz * a6 + f2[0] * f2[1] * rz * a7;
}

/* Condense dc_dv into the tile's 64 slots, weights bx[i]*(by[j]*bz[k]).
   Follows bspline_update_sets_b, src/register/bspline.cxx:347-374. */
static inline void
update_sets_b (float* sx, float* sy, float* sz, const int64_t q[3],
    const float dc_dv[3], const bxf_t* b)
{
    const float* bx = &b->bx_lut[q[0]*4];
    const float* by = &b->by_lut[q[1]*4];
    const float* bz = &b->bz_lut[q[2]*4];
    int m = 0;
    for (int k = 0; k < 4; k++) {
        float C = bz[k];
        for (int j = 0; j < 4; j++) {
            float B = by[j] * C;
            for (int i = 0; i < 4; i++) {
                float A = bx[i] * B;
                sx[m] += dc_dv[0] * A;
                sy[m] += dc_dv[1] * A;
                sz[m] += dc_dv[2] * A;
                m++;
            }
        }
    }
}

/* Knot index of slot (i,j,k) of tile pidx.  Same indices as find_knots,
   src/register/bspline.cxx:214-250, and the c_lut. */
static inline int64_t
tile_knot (const bxf_t* b, int64_t pidx, int slot)
{
    return b->c_lut[pidx*64 + slot];
}

/* Place the tile's slots at cond[64*knot+slot] (single writer per entry).
   Follows bspline_sort_sets, src/register/bspline.cxx:377-398. */
static inline void
sort_sets (float* cx, float* cy, float* cz, const float* sx, const float* sy,
    const float* sz, int64_t pidx, const bxf_t* b)
{
    for (int s = 0; s < 64; s++) {
        int64_t k = tile_knot (b, pidx, s);
        cx[64*k + s] = sx[s];
        cy[64*k + s] = sy[s];
        cz[64*k + s] = sz[s];
    }
}

/* grad[3k+d] += sum over 64 slots, serial.  Follows
   bspline_condense_smetric_grad, src/register/bspline.cxx:400-413. */
static void
condense_grad (float* grad, const float* cx, const float* cy, const float* cz,
    const bxf_t* b)
{
    for (int64_t k = 0; k < b->num_knots; k++) {
        for (int s = 0; s < 64; s++) {
            grad[3*k+0] += cx[64*k + s];
            grad[3*k+1] += cy[64*k + s];
            grad[3*k+2] += cz[64*k + s];
        }
    }
}

static void
set_threads (int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads (nthreads);
#else
    (void) nthreads;
#endif
}

/* ---------------------------------------------------------------------
   Moving-image gradient
   --------------------------------------------------------------------- */

/* Central difference (m[i+1]-m[i-1])/2/spacing with edge replicate,
   rotated by inv(dc); interleaved float3 output.  Follows
   volume_calc_grad, src/base/volume_grad.cxx:9-76. */
void
orc_volume_gradient (const orc_vol_hdr* h, const float* img, float* out)
{
    vol_t v;
    vol_init (&v, h, img);
    const int64_t* dim = v.dim;
    int64_t vv = 0;
    for (int64_t k = 0; k < dim[2]; k++) {
        int64_t kp = (k == 0) ? 0 : k - 1;
        int64_t kn = (k == dim[2]-1) ? dim[2]-1 : k + 1;
        for (int64_t j = 0; j < dim[1]; j++) {
            int64_t jp = (j == 0) ? 0 : j - 1;
            int64_t jn = (j == dim[1]-1) ? dim[1]-1 : j + 1;
            for (int64_t i = 0; i < dim[0]; i++, vv++) {
                int64_t ip = (i == 0) ? 0 : i - 1;
                int64_t in = (i == dim[0]-1) ? dim[0]-1 : i + 1;
                float g0 = 0.f, g1 = 0.f, g2 = 0.f, diff;
                const float* r = img;
                diff = (float) ((double) (r[in + dim[0]*(j + dim[1]*k)]
                        - r[ip + dim[0]*(j + dim[1]*k)]) / 2.0 / (double) v.spacing[0]);
                g0 += diff * v.inv_dc[0]; g1 += diff * v.inv_dc[3]; g2 += diff * v.inv_dc[6];
                diff = (float) ((double) (r[i + dim[0]*(jn + dim[1]*k)]
                        - r[i + dim[0]*(jp + dim[1]*k)]) / 2.0 / (double) v.spacing[1]);
                g0 += diff * v.inv_dc[1]; g1 += diff * v.inv_dc[4]; g2 += diff * v.inv_dc[7];
                diff = (float) ((double) (r[i + dim[0]*(j + dim[1]*kn)]
                        - r[i + dim[0]*(j + dim[1]*kp)]) / 2.0 / (double) v.spacing[2]);
                g0 += diff * v.inv_dc[2]; g1 += diff * v.inv_dc[5]; g2 += diff * v.inv_dc[8];
                out[3*vv+0] = g0;
                out[3*vv+1] = g1;
                out[3*vv+2] = g2;
            }
        }
    }
}

/* ---------------------------------------------------------------------
   Dense displacement field
   --------------------------------------------------------------------- */

/* vf[3*v+d] over the ROI, q_lut products.  Follows bspline_interpolate_vf
   and bspline_interp_pix, src/base/bspline_interpolate.cxx:21-48,112-142. */
void
orc_interp_vf (const orc_geom* g, const float* coeff, float* vf)
{
    bxf_t b;
    if (bxf_init (&b, g, coeff)) return;
    int64_t vox = 0;
    memset (vf, 0, sizeof(float) * 3 * g->img_dim[0] * g->img_dim[1] * g->img_dim[2]);
    for (int64_t k = 0; k < b.roi_dim[2]; k++) {
        int64_t p2 = k / b.vpr[2], q2 = k % b.vpr[2];
        for (int64_t j = 0; j < b.roi_dim[1]; j++) {
            int64_t p1 = j / b.vpr[1], q1 = j % b.vpr[1];
            for (int64_t i = 0; i < b.roi_dim[0]; i++) {
                int64_t p0 = i / b.vpr[0], q0 = i % b.vpr[0];
                int64_t qidx = q0 + b.vpr[0] * (q1 + b.vpr[1] * q2);
                const float* q_lut = &b.q_lut[qidx*64];
                float out[3] = {0, 0, 0};
                int m = 0;
                for (int kk = 0; kk < 4; kk++)
                    for (int jj = 0; jj < 4; jj++)
                        for (int ii = 0; ii < 4; ii++) {
                            int64_t cidx = 3 * ((p2 + kk) * b.cdims[1] * b.cdims[0]
                                + (p1 + jj) * b.cdims[0] + (p0 + ii));
                            out[0] += q_lut[m] * coeff[cidx+0];
                            out[1] += q_lut[m] * coeff[cidx+1];
                            out[2] += q_lut[m] * coeff[cidx+2];
                            m++;
                        }
                vox = (b.roi_offset[0] + i) + g->img_dim[0]
                    * ((b.roi_offset[1] + j) + g->img_dim[1] * (b.roi_offset[2] + k));
                vf[3*vox+0] = out[0];
                vf[3*vox+1] = out[1];
                vf[3*vox+2] = out[2];
            }
        }
    }
    bxf_free (&b);
}

/* ---------------------------------------------------------------------
   Warped moving image
   --------------------------------------------------------------------- */

/* out = the moving image at each fixed voxel moved by the B-spline
   displacement, trilinear (linear_interp) or nearest voxel; default_val
   where the moving position falls outside the moving image, and on the
   voxels outside the ROI.  Follows bspline_warp_dcos,
   src/base/bspline_warp.cxx:27-167, for float images: fixed positions as
   the LOOP_Z_OMP / LOOP_Y / LOOP_X macros form them
   (src/base/volume_macros.h:46-73: x accumulated step by step), the
   displacement of bspline_interp_pix_b, the correspondence and inside
   test, li_clamp_3d and LI_VALUE (li_value here).  The reference indexes
   regions from the image's first voxel (REGION_INDEX_*,
   src/base/bspline_macros.h:33-80), i.e. it assumes the ROI is the whole
   image; regions count from the ROI's first voxel here, the same thing in
   that case. */
int
orc_warp (const orc_geom* g, const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, int linear_interp, float default_val, int nthreads, float* out)
{
    bxf_t b;
    vol_t mv;
    if (bxf_init (&b, g, coeff)) return -1;
    vol_init (&mv, mh, m_img);
    set_threads (nthreads);
    float step[9], proj[9], inv_dc[9];
    orc_direction_matrices (g->dc, g->img_spacing, step, proj, inv_dc);
    const int64_t* dim = g->img_dim;
    const float* origin = g->img_origin;
    const int64_t npix = dim[0] * dim[1] * dim[2];
    for (int64_t v = 0; v < npix; v++) out[v] = default_val;
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < dim[2]; k++) {
        int64_t fijk[3];
        float fxyz[3];
        fijk[2] = k;
        fxyz[2] = origin[2] + fijk[2] * step[2*3+2];
        for (fijk[1] = 0, fxyz[1] = origin[1] + fijk[2] * step[1*3+2];
             fijk[1] < dim[1];
             ++fijk[1],
             fxyz[2] = origin[2] + fijk[2] * step[2*3+2] + fijk[1] * step[2*3+1],
             fxyz[1] = origin[1] + fijk[2] * step[1*3+2] + fijk[1] * step[1*3+1]) {
            for (fijk[0] = 0, fxyz[0] = origin[0] + fijk[2] * step[0*3+2] + fijk[1] * step[0*3+1];
                 fijk[0] < dim[0];
                 ++fijk[0], fxyz[0] += step[0*3+0], fxyz[1] += step[1*3+0],
                 fxyz[2] += step[2*3+0]) {
                int64_t r[3], p[3], q[3];
                int in_roi = 1;
                for (int d = 0; d < 3; d++) {
                    r[d] = fijk[d] - b.roi_offset[d];
                    if (r[d] < 0 || r[d] >= b.roi_dim[d]) in_roi = 0;
                }
                if (!in_roi) continue;
                for (int d = 0; d < 3; d++) {
                    p[d] = r[d] / b.vpr[d];
                    q[d] = r[d] % b.vpr[d];
                }
                const int64_t pidx = p[0] + b.rdims[0] * (p[1] + b.rdims[1] * p[2]);
                const int64_t qidx = q[0] + b.vpr[0] * (q[1] + b.vpr[1] * q[2]);
                float d[3], mijk[3], l1[3], l2[3];
                int64_t mf[3], mr[3];
                interp_pix_b (d, &b, pidx, qidx);
                const int64_t fv = fijk[0] + dim[0] * (fijk[1] + dim[1] * fijk[2]);
                if (!find_correspondence (mijk, fxyz, d, &mv)) continue;
                li_clamp_3d (mijk, mf, mr, l1, l2, &mv);
                if (linear_interp) {
                    const int64_t mvf = mf[0] + mv.dim[0] * (mf[1] + mv.dim[1] * mf[2]);
                    out[fv] = li_value (l1, l2, mvf, m_img, mv.dim);
                } else {
                    const int64_t mvr = mr[0] + mv.dim[0] * (mr[1] + mv.dim[1] * mr[2]);
                    out[fv] = m_img[mvr];
                }
            }
        }
    }
    bxf_free (&b);
    return 0;
}

/* ---------------------------------------------------------------------
   MSE
   --------------------------------------------------------------------- */

/* score = raw/N; grad = 2*grad/N; N < 1 => FLT_MAX and zero gradient.
   Follows bspline_score_normalize, src/register/bspline_mse.cxx:43-73. */
static void
mse_normalize (double raw, int64_t n, float* score, float* grad, int64_t nc)
{
    if (n < 1) {
        *score = FLT_MAX;
        for (int64_t i = 0; i < nc; i++) grad[i] = 0;
    } else {
        *score = (float) (raw / (double) n);
        for (int64_t i = 0; i < nc; i++) grad[i] = 2 * grad[i] / (float) n;
    }
}

/* Tile-condense MSE: flavor g (nearest-voxel m_grad),
   src/register/bspline_mse.cxx:519-720, and flavor p (analytic trilinear
   derivative), src/register/bspline_mse.cxx:1149-1384. */
static void
mse_tiles (const bxf_t* b, const vol_t* fx, const vol_t* mv,
    const float* m_grad, int analytic, const roi_t* froi, const roi_t* mroi,
    double* raw, int64_t* nvox, float* grad)
{
    int64_t csz = 64 * b->num_knots;
    float* cx = (float*) calloc (csz, sizeof(float));
    float* cy = (float*) calloc (csz, sizeof(float));
    float* cz = (float*) calloc (csz, sizeof(float));
    double score_tile = 0;
    int64_t num_vox = 0;
    float inv[3];
    inv[0] = (float) (1.0 / mv->spacing[0]);
    inv[1] = (float) (1.0 / mv->spacing[1]);
    inv[2] = (float) (1.0 / mv->spacing[2]);

#pragma omp parallel for reduction(+:num_vox,score_tile) schedule(dynamic, 4)
    for (int64_t pidx = 0; pidx < b->num_tiles; pidx++) {
        float sx[64], sy[64], sz[64];
        memset (sx, 0, sizeof sx); memset (sy, 0, sizeof sy); memset (sz, 0, sizeof sz);
        int64_t p[3], q[3], fijk[3];
        p[2] = pidx / (b->rdims[0] * b->rdims[1]);
        p[1] = (pidx - p[2] * b->rdims[0] * b->rdims[1]) / b->rdims[0];
        p[0] = pidx - p[2] * b->rdims[0] * b->rdims[1] - p[1] * b->rdims[0];
        for (q[2] = 0; q[2] < b->vpr[2]; q[2]++)
        for (q[1] = 0; q[1] < b->vpr[1]; q[1]++)
        for (q[0] = 0; q[0] < b->vpr[0]; q[0]++) {
            for (int d = 0; d < 3; d++)
                fijk[d] = b->roi_offset[d] + b->vpr[d] * p[d] + q[d];
            if (fijk[0] >= b->roi_offset[0] + b->roi_dim[0]) continue;
            if (fijk[1] >= b->roi_offset[1] + b->roi_dim[1]) continue;
            if (fijk[2] >= b->roi_offset[2] + b->roi_dim[2]) continue;
            float fxyz[3];
            for (int d = 0; d < 3; d++)
                fxyz[d] = b->img_origin[d] + fijk[0] * fx->step[3*d+0]
                    + fijk[1] * fx->step[3*d+1] + fijk[2] * fx->step[3*d+2];
            /* flavors i and r: the fixed ROI at the fixed position, the
               moving ROI at the moving one (src/register/bspline_mse.cxx:
               172-196, bspline_find_correspondence_dcos_roi,
               src/util/bspline_correspond.cxx:113-145) */
            if (froi && !inside_roi (fxyz, froi)) continue;
            int64_t fidx = fijk[0] + fx->dim[0] * (fijk[1] + fx->dim[1] * fijk[2]);
            float dxyz[3], mijk[3];
            interp_pix_c (dxyz, b, pidx, q);
            if (!find_correspondence (mijk, fxyz, dxyz, mv)) continue;
            if (mroi) {
                float mxyz[3] = {fxyz[0] + dxyz[0], fxyz[1] + dxyz[1], fxyz[2] + dxyz[2]};
                if (!inside_roi (mxyz, mroi)) continue;
            }
            int64_t mf[3], mr[3];
            float l1[3], l2[3];
            li_clamp_3d (mijk, mf, mr, l1, l2, mv);
            int64_t mvf = mf[0] + mv->dim[0] * (mf[1] + mv->dim[1] * mf[2]);
            int64_t mvr = mr[0] + mv->dim[0] * (mr[1] + mv->dim[1] * mr[2]);
            float m_val = li_value (l1, l2, mvf, mv->img, mv->dim);
            float diff = m_val - fx->img[fidx];
            score_tile += diff * diff;
            num_vox++;
            float dc_dv[3];
            if (analytic) {
                float md[3];
                li_value_d (md, l1, l2, inv, mvf, mv->img, mv->dim);
                dc_dv[0] = diff * md[0];
                dc_dv[1] = diff * md[1];
                dc_dv[2] = diff * md[2];
            } else {
                dc_dv[0] = diff * m_grad[3*mvr+0];
                dc_dv[1] = diff * m_grad[3*mvr+1];
                dc_dv[2] = diff * m_grad[3*mvr+2];
            }
            update_sets_b (sx, sy, sz, q, dc_dv, b);
        }
        sort_sets (cx, cy, cz, sx, sy, sz, pidx, b);
    }
    condense_grad (grad, cx, cy, cz, b);
    free (cx); free (cy); free (cz);
    *raw = score_tile;
    *nvox = num_vox;
}

/* Voxel-order MSE with per-voxel gradient scatter (the simplest form of
   the g semantics).  Follows bspline_score_c_mse,
   src/register/bspline_mse.cxx:728-885, with the scatter of
   Bspline_score::update_grad_b, src/register/bspline_score.cxx:145-170,
   and the LOOP_Z/Y/X position recurrences of
   src/base/volume_macros.h:28-70. */
static void
mse_voxel_order (const bxf_t* b, const vol_t* fx, const vol_t* mv,
    const float* m_grad, double* raw, int64_t* nvox, float* grad)
{
    double score_acc = 0.;
    int64_t num_vox = 0;
    int64_t fijk[3];
    float fxyz[3];
    const float* o = fx->origin;
    const float* s = fx->step;
    fxyz[2] = o[2];
    for (fijk[2] = 0; fijk[2] < fx->dim[2]; ) {
        int64_t p[3], q[3];
        p[2] = fijk[2] / b->vpr[2]; q[2] = fijk[2] % b->vpr[2];
        fxyz[1] = o[1] + fijk[2] * s[5];
        for (fijk[1] = 0; fijk[1] < fx->dim[1]; ) {
            p[1] = fijk[1] / b->vpr[1]; q[1] = fijk[1] % b->vpr[1];
            fxyz[0] = o[0] + fijk[2] * s[2] + fijk[1] * s[1];
            for (fijk[0] = 0; fijk[0] < fx->dim[0]; ) {
                p[0] = fijk[0] / b->vpr[0]; q[0] = fijk[0] % b->vpr[0];
                int64_t pidx = p[0] + b->rdims[0] * (p[1] + b->rdims[1] * p[2]);
                int64_t qidx = q[0] + b->vpr[0] * (q[1] + b->vpr[1] * q[2]);
                float dxyz[3], mxyz[3], mijk[3];
                interp_pix_b (dxyz, b, pidx, qidx);
                mxyz[2] = fxyz[2] + dxyz[2] - mv->origin[2];
                mxyz[1] = fxyz[1] + dxyz[1] - mv->origin[1];
                mxyz[0] = fxyz[0] + dxyz[0] - mv->origin[0];
                mijk[2] = mxyz[0] * mv->proj[6] + mxyz[1] * mv->proj[7] + mxyz[2] * mv->proj[8];
                mijk[1] = mxyz[0] * mv->proj[3] + mxyz[1] * mv->proj[4] + mxyz[2] * mv->proj[5];
                mijk[0] = mxyz[0] * mv->proj[0] + mxyz[1] * mv->proj[1] + mxyz[2] * mv->proj[2];
                int inside = 1;
                for (int d = 0; d < 3; d++)
                    if (mijk[d] <= -0.5 || mijk[d] >= mv->dim[d] - 0.5) inside = 0;
                if (inside) {
                    int64_t mf[3], mr[3];
                    float l1[3], l2[3];
                    li_clamp_3d (mijk, mf, mr, l1, l2, mv);
                    int64_t mvf = mf[0] + mv->dim[0] * (mf[1] + mv->dim[1] * mf[2]);
                    float m_val = li_value (l1, l2, mvf, mv->img, mv->dim);
                    int64_t fv = fijk[0] + fx->dim[0] * (fijk[1] + fx->dim[1] * fijk[2]);
                    float diff = m_val - fx->img[fv];
                    int64_t mvr = mr[0] + mv->dim[0] * (mr[1] + mv->dim[1] * mr[2]);
                    float dc_dv[3];
                    dc_dv[0] = diff * m_grad[3*mvr+0];
                    dc_dv[1] = diff * m_grad[3*mvr+1];
                    dc_dv[2] = diff * m_grad[3*mvr+2];
                    const float* q_lut = &b->q_lut[qidx*64];
                    for (int m = 0; m < 64; m++) {
                        int64_t cidx = 3 * b->c_lut[pidx*64 + m];
                        grad[cidx+0] += dc_dv[0] * q_lut[m];
                        grad[cidx+1] += dc_dv[1] * q_lut[m];
                        grad[cidx+2] += dc_dv[2] * q_lut[m];
                    }
                    score_acc += diff * diff;
                    num_vox++;
                }
                ++fijk[0];
                fxyz[0] += s[0]; fxyz[1] += s[3]; fxyz[2] += s[6];
            }
            ++fijk[1];
            fxyz[2] = o[2] + fijk[2] * s[8] + fijk[1] * s[7];
            fxyz[1] = o[1] + fijk[2] * s[5] + fijk[1] * s[4];
        }
        ++fijk[2];
        fxyz[2] = o[2] + fijk[2] * s[8];
    }
    *raw = score_acc;
    *nvox = num_vox;
}

int
orc_mse_score (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img, const float* m_grad,
    const float* coeff, int flavor, int nthreads, const orc_rois* rois,
    double* raw_score, float* score, float* grad, int64_t* num_vox)
{
    bxf_t b;
    vol_t fx, mv;
    roi_t fr_, mr_;
    const roi_t* froi = rois ? roi_init (&fr_, rois->fixed_hdr, rois->fixed) : 0;
    const roi_t* mroi = rois ? roi_init (&mr_, rois->moving_hdr, rois->moving) : 0;
    if (flavor == 'c' && (froi || mroi)) return -4;
    if (bxf_init (&b, g, coeff)) return -1;
    vol_init (&fx, fh, f_img);
    vol_init (&mv, mh, m_img);
    set_threads (nthreads);
    memset (grad, 0, sizeof(float) * b.num_coeff);
    double raw = 0;
    int64_t n = 0;
    if (flavor == 'g') {
        if (!m_grad) { bxf_free (&b); return -2; }
        mse_tiles (&b, &fx, &mv, m_grad, 0, froi, mroi, &raw, &n, grad);
    } else if (flavor == 'p') {
        mse_tiles (&b, &fx, &mv, 0, 1, froi, mroi, &raw, &n, grad);
    } else if (flavor == 'c') {
        if (!m_grad) { bxf_free (&b); return -2; }
        mse_voxel_order (&b, &fx, &mv, m_grad, &raw, &n, grad);
    } else {
        bxf_free (&b);
        return -3;
    }
    mse_normalize (raw, n, score, grad, b.num_coeff);
    *raw_score = raw;
    *num_vox = n;
    bxf_free (&b);
    return 0;
}

/* ---------------------------------------------------------------------
   MSE in double precision (the FP64 yardstick)
   --------------------------------------------------------------------- */

static void
basis_f64 (int64_t vpr, double* lut)
{
    for (int64_t i = 0; i < vpr; i++) {
        double u = (double) i / (double) vpr;
        lut[i*4+0] = (1.0/6.0) * (1.0 - u) * (1.0 - u) * (1.0 - u);
        lut[i*4+1] = (1.0/6.0) * (3.0*u*u*u - 6.0*u*u + 4.0);
        lut[i*4+2] = (1.0/6.0) * (-3.0*u*u*u + 3.0*u*u + 3.0*u + 1.0);
        lut[i*4+3] = (1.0/6.0) * u*u*u;
    }
}

typedef struct dvol_t {
    int64_t dim[3];
    double origin[3];
    double step[9];
    double proj[9];
    double inv_sp[3];
    const float* img;
} dvol_t;

/* proj = inv(dc) / spacing in double, as dvol_init forms it. */
static void
dproj_of (double* proj, const orc_vol_hdr* h)
{
    const float* m = h->dc;
    double a[9], inv[9];
    for (int i = 0; i < 9; i++) a[i] = m[i];
    double det = a[0]*(a[4]*a[8]-a[5]*a[7]) - a[1]*(a[3]*a[8]-a[5]*a[6])
        + a[2]*(a[3]*a[7]-a[4]*a[6]);
    inv[0] =  (a[4]*a[8]-a[5]*a[7]) / det;
    inv[1] = -(a[1]*a[8]-a[2]*a[7]) / det;
    inv[2] =  (a[1]*a[5]-a[2]*a[4]) / det;
    inv[3] = -(a[3]*a[8]-a[5]*a[6]) / det;
    inv[4] =  (a[0]*a[8]-a[2]*a[6]) / det;
    inv[5] = -(a[0]*a[5]-a[2]*a[3]) / det;
    inv[6] =  (a[3]*a[7]-a[4]*a[6]) / det;
    inv[7] = -(a[0]*a[7]-a[1]*a[6]) / det;
    inv[8] =  (a[0]*a[4]-a[1]*a[3]) / det;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            proj[3*i+j] = inv[3*i+j] / (double) h->spacing[i];
}

static void
dvol_init (dvol_t* v, const orc_vol_hdr* h, const float* img)
{
    /* inverse of the direction cosines in double */
    const float* m = h->dc;
    double a[9], inv[9];
    for (int i = 0; i < 9; i++) a[i] = m[i];
    double det = a[0]*(a[4]*a[8]-a[5]*a[7]) - a[1]*(a[3]*a[8]-a[5]*a[6])
        + a[2]*(a[3]*a[7]-a[4]*a[6]);
    inv[0] =  (a[4]*a[8]-a[5]*a[7]) / det;
    inv[1] = -(a[1]*a[8]-a[2]*a[7]) / det;
    inv[2] =  (a[1]*a[5]-a[2]*a[4]) / det;
    inv[3] = -(a[3]*a[8]-a[5]*a[6]) / det;
    inv[4] =  (a[0]*a[8]-a[2]*a[6]) / det;
    inv[5] = -(a[0]*a[5]-a[2]*a[3]) / det;
    inv[6] =  (a[3]*a[7]-a[4]*a[6]) / det;
    inv[7] = -(a[0]*a[7]-a[1]*a[6]) / det;
    inv[8] =  (a[0]*a[4]-a[1]*a[3]) / det;
    for (int d = 0; d < 3; d++) {
        v->dim[d] = h->dim[d];
        v->origin[d] = h->origin[d];
        v->inv_sp[d] = 1.0 / (double) h->spacing[d];
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            v->step[3*i+j] = a[3*i+j] * (double) h->spacing[j];
            v->proj[3*i+j] = inv[3*i+j] / (double) h->spacing[i];
        }
    v->img = img;
}

typedef struct dgeo_t {
    int64_t vpr[3], rdims[3], cdims[3], roi_offset[3], roi_dim[3];
    int64_t num_knots, num_tiles;
    double origin[3];
    double *bx, *by, *bz;
} dgeo_t;

static int
dgeo_init (dgeo_t* d, const orc_geom* g)
{
    for (int a = 0; a < 3; a++) {
        d->vpr[a] = g->vox_per_rgn[a];
        d->roi_offset[a] = g->roi_offset[a];
        d->roi_dim[a] = g->roi_dim[a];
        d->origin[a] = g->img_origin[a];
        if (d->vpr[a] < 1 || d->roi_dim[a] < 1) return -1;
    }
    d->num_knots = orc_header (g, d->rdims, d->cdims);
    d->num_tiles = d->rdims[0] * d->rdims[1] * d->rdims[2];
    d->bx = (double*) malloc (sizeof(double) * 4 * d->vpr[0]);
    d->by = (double*) malloc (sizeof(double) * 4 * d->vpr[1]);
    d->bz = (double*) malloc (sizeof(double) * 4 * d->vpr[2]);
    basis_f64 (d->vpr[0], d->bx);
    basis_f64 (d->vpr[1], d->by);
    basis_f64 (d->vpr[2], d->bz);
    return 0;
}

static void
dgeo_free (dgeo_t* d)
{
    free (d->bx); free (d->by); free (d->bz);
}

/* Per-voxel correspondence + trilinear state in double. */
typedef struct dsample_t {
    int64_t mvf, mvr;
    double l1[3], l2[3];
} dsample_t;

static inline int
dsample (dsample_t* s, const double fxyz[3], const double d[3], const dvol_t* mv)
{
    double t[3], mijk[3];
    for (int a = 0; a < 3; a++) t[a] = fxyz[a] + d[a] - mv->origin[a];
    for (int a = 0; a < 3; a++)
        mijk[a] = t[0] * mv->proj[3*a+0] + t[1] * mv->proj[3*a+1] + t[2] * mv->proj[3*a+2];
    for (int a = 0; a < 3; a++)
        if (mijk[a] <= -0.5 || mijk[a] >= mv->dim[a] - 0.5) return 0;
    int64_t f[3], r[3];
    for (int a = 0; a < 3; a++) {
        double ma = mijk[a];
        int64_t dmax = mv->dim[a] - 1;
        if (ma < 0) { f[a] = 0; r[a] = 0; s->l2[a] = 0; }
        else if (ma >= dmax) { f[a] = dmax - 1; r[a] = dmax; s->l2[a] = 1; }
        else {
            f[a] = (int64_t) floor (ma);
            r[a] = (int64_t) (ma + 0.5);
            s->l2[a] = ma - f[a];
        }
        s->l1[a] = 1.0 - s->l2[a];
    }
    s->mvf = f[0] + mv->dim[0] * (f[1] + mv->dim[1] * f[2]);
    s->mvr = r[0] + mv->dim[0] * (r[1] + mv->dim[1] * r[2]);
    return 1;
}

static inline void
dcorners (double a[8], const dsample_t* s, const dvol_t* mv)
{
    int64_t sx = mv->dim[0], sxy = mv->dim[0] * mv->dim[1];
    const float* m = mv->img;
    a[0] = m[s->mvf]; a[1] = m[s->mvf+1];
    a[2] = m[s->mvf+sx]; a[3] = m[s->mvf+sx+1];
    a[4] = m[s->mvf+sxy]; a[5] = m[s->mvf+sxy+1];
    a[6] = m[s->mvf+sxy+sx]; a[7] = m[s->mvf+sxy+sx+1];
}

/* Displacement of voxel q in tile p in double. */
static inline void
dinterp (double out[3], const dgeo_t* d, const int64_t p[3], const int64_t q[3],
    const float* coeff)
{
    out[0] = out[1] = out[2] = 0;
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) {
                double w = d->bx[q[0]*4+i] * d->by[q[1]*4+j] * d->bz[q[2]*4+k];
                int64_t kn = (p[0] + i) + d->cdims[0] * ((p[1] + j) + d->cdims[1] * (p[2] + k));
                out[0] += w * coeff[3*kn+0];
                out[1] += w * coeff[3*kn+1];
                out[2] += w * coeff[3*kn+2];
            }
}

static inline void
dscatter (double* tg, const dgeo_t* d, const int64_t p[3], const int64_t q[3],
    const double dc[3])
{
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) {
                double w = d->bx[q[0]*4+i] * d->by[q[1]*4+j] * d->bz[q[2]*4+k];
                int s = (k*4 + j)*4 + i;
                tg[3*s+0] += dc[0] * w;
                tg[3*s+1] += dc[1] * w;
                tg[3*s+2] += dc[2] * w;
            }
}

/* Tile partials (64 slots x 3) are placed per tile, then each knot sums
   its <=64 contributions in slot order; this is the same condense/reduce
   structure as the float flavors, carried out in double. */
static void
dreduce (double* grad, const double* tiles, const dgeo_t* d)
{
    int64_t nk = d->num_knots;
    memset (grad, 0, sizeof(double) * 3 * nk);
    for (int64_t t = 0; t < d->num_tiles; t++) {
        int64_t p[3];
        p[2] = t / (d->rdims[0] * d->rdims[1]);
        p[1] = (t / d->rdims[0]) % d->rdims[1];
        p[0] = t % d->rdims[0];
        for (int k = 0; k < 4; k++)
            for (int j = 0; j < 4; j++)
                for (int i = 0; i < 4; i++) {
                    int s = (k*4 + j)*4 + i;
                    int64_t kn = (p[0] + i) + d->cdims[0] * ((p[1] + j) + d->cdims[1] * (p[2] + k));
                    for (int c = 0; c < 3; c++)
                        grad[3*kn+c] += tiles[(t*64 + s)*3 + c];
                }
    }
}

int
orc_mse_score_f64 (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img, const float* m_grad,
    const float* coeff, int grad_mode, int nthreads, const orc_rois* rois,
    double* score, double* grad, int64_t* num_vox)
{
    dgeo_t d;
    dvol_t fx, mv;
    roi_t fr_, mr_;
    const roi_t* froi = rois ? roi_init (&fr_, rois->fixed_hdr, rois->fixed) : 0;
    const roi_t* mroi = rois ? roi_init (&mr_, rois->moving_hdr, rois->moving) : 0;
    if (dgeo_init (&d, g)) return -1;
    if (grad_mode == 0 && !m_grad) { dgeo_free (&d); return -2; }
    dvol_init (&fx, fh, f_img);
    dvol_init (&mv, mh, m_img);
    set_threads (nthreads);
    double* tiles = (double*) calloc (d.num_tiles * 64 * 3, sizeof(double));
    double raw = 0;
    int64_t n = 0;

#pragma omp parallel for reduction(+:n,raw) schedule(dynamic, 4)
    for (int64_t t = 0; t < d.num_tiles; t++) {
        int64_t p[3], q[3], fijk[3];
        p[2] = t / (d.rdims[0] * d.rdims[1]);
        p[1] = (t / d.rdims[0]) % d.rdims[1];
        p[0] = t % d.rdims[0];
        double* tg = &tiles[t * 64 * 3];
        for (q[2] = 0; q[2] < d.vpr[2]; q[2]++)
        for (q[1] = 0; q[1] < d.vpr[1]; q[1]++)
        for (q[0] = 0; q[0] < d.vpr[0]; q[0]++) {
            int skip = 0;
            for (int a = 0; a < 3; a++) {
                fijk[a] = d.roi_offset[a] + d.vpr[a] * p[a] + q[a];
                if (fijk[a] >= d.roi_offset[a] + d.roi_dim[a]) skip = 1;
            }
            if (skip) continue;
            double fxyz[3], dxyz[3];
            for (int a = 0; a < 3; a++)
                fxyz[a] = d.origin[a] + fijk[0] * fx.step[3*a+0]
                    + fijk[1] * fx.step[3*a+1] + fijk[2] * fx.step[3*a+2];
            if (froi && !inside_roi_d (fxyz, froi)) continue;
            dinterp (dxyz, &d, p, q, coeff);
            dsample_t s;
            if (!dsample (&s, fxyz, dxyz, &mv)) continue;
            if (mroi) {
                double mxyz[3] = {fxyz[0] + dxyz[0], fxyz[1] + dxyz[1], fxyz[2] + dxyz[2]};
                if (!inside_roi_d (mxyz, mroi)) continue;
            }
            double a8[8];
            dcorners (a8, &s, &mv);
            const double* l1 = s.l1;
            const double* l2 = s.l2;
            double m_val = l1[0]*l1[1]*l1[2]*a8[0] + l2[0]*l1[1]*l1[2]*a8[1]
                + l1[0]*l2[1]*l1[2]*a8[2] + l2[0]*l2[1]*l1[2]*a8[3]
                + l1[0]*l1[1]*l2[2]*a8[4] + l2[0]*l1[1]*l2[2]*a8[5]
                + l1[0]*l2[1]*l2[2]*a8[6] + l2[0]*l2[1]*l2[2]*a8[7];
            int64_t fidx = fijk[0] + fx.dim[0] * (fijk[1] + fx.dim[1] * fijk[2]);
            double diff = m_val - (double) fx.img[fidx];
            raw += diff * diff;
            n++;
            double dc[3];
            if (grad_mode == 1) {
                const double* r = mv.inv_sp;
                double gx = r[0] * (l1[1]*l1[2]*(a8[1]-a8[0]) + l2[1]*l1[2]*(a8[3]-a8[2])
                    + l1[1]*l2[2]*(a8[5]-a8[4]) + l2[1]*l2[2]*(a8[7]-a8[6]));
                double gy = r[1] * (l1[0]*l1[2]*(a8[2]-a8[0]) + l2[0]*l1[2]*(a8[3]-a8[1])
                    + l1[0]*l2[2]*(a8[6]-a8[4]) + l2[0]*l2[2]*(a8[7]-a8[5]));
                double gz = r[2] * (l1[0]*l1[1]*(a8[4]-a8[0]) + l2[0]*l1[1]*(a8[5]-a8[1])
                    + l1[0]*l2[1]*(a8[6]-a8[2]) + l2[0]*l2[1]*(a8[7]-a8[3]));
                dc[0] = diff * gx; dc[1] = diff * gy; dc[2] = diff * gz;
            } else {
                dc[0] = diff * (double) m_grad[3*s.mvr+0];
                dc[1] = diff * (double) m_grad[3*s.mvr+1];
                dc[2] = diff * (double) m_grad[3*s.mvr+2];
            }
            dscatter (tg, &d, p, q, dc);
        }
    }
    dreduce (grad, tiles, &d);
    free (tiles);
    if (n < 1) {
        *score = FLT_MAX;
        for (int64_t i = 0; i < 3 * d.num_knots; i++) grad[i] = 0;
    } else {
        *score = raw / (double) n;
        for (int64_t i = 0; i < 3 * d.num_knots; i++) grad[i] = 2.0 * grad[i] / (double) n;
    }
    *num_vox = n;
    dgeo_free (&d);
    return 0;
}

/* The FP64 sum of squared differences over the support of coefficient ci
   (the voxels of the up to 4x4x4 tiles whose knots include its knot), with
   that coefficient moved by delta in double; the other coefficients, and
   every voxel outside the support, are unchanged, so the central difference
   (raw(+h) - raw(-h)) / (2 h N) is that of the whole FP64 score whenever the
   overlap count N does not change.  n counts the support's voxels inside
   the moving image, nclamp those among them whose position li_clamp clamps
   (some coordinate below 0 or at or above dim - 1: there li_value_dx keeps
   its slope while the clamped value does not move).  Same arithmetic as
   orc_mse_score_f64. */
int
orc_mse_f64_support (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, int64_t ci, double delta, const orc_rois* rois,
    double* raw_out, int64_t* n_out, int64_t* nclamp_out)
{
    dgeo_t d;
    dvol_t fx, mv;
    roi_t fr_, mr_;
    const roi_t* froi = rois ? roi_init (&fr_, rois->fixed_hdr, rois->fixed) : 0;
    const roi_t* mroi = rois ? roi_init (&mr_, rois->moving_hdr, rois->moving) : 0;
    if (dgeo_init (&d, g)) return -1;
    if (ci < 0 || ci >= 3 * d.num_knots) { dgeo_free (&d); return -2; }
    dvol_init (&fx, fh, f_img);
    dvol_init (&mv, mh, m_img);
    const int64_t knot = ci / 3, comp = ci % 3;
    const int64_t kk[3] = {knot % d.cdims[0], (knot / d.cdims[0]) % d.cdims[1],
                           knot / (d.cdims[0] * d.cdims[1])};
    double raw = 0;
    int64_t n = 0, nclamp = 0;
    int64_t p[3], q[3], fijk[3];
    for (p[2] = kk[2] - 3; p[2] <= kk[2]; p[2]++)
    for (p[1] = kk[1] - 3; p[1] <= kk[1]; p[1]++)
    for (p[0] = kk[0] - 3; p[0] <= kk[0]; p[0]++) {
        int out = 0;
        for (int a = 0; a < 3; a++) if (p[a] < 0 || p[a] >= d.rdims[a]) out = 1;
        if (out) continue;
        for (q[2] = 0; q[2] < d.vpr[2]; q[2]++)
        for (q[1] = 0; q[1] < d.vpr[1]; q[1]++)
        for (q[0] = 0; q[0] < d.vpr[0]; q[0]++) {
            int skip = 0;
            for (int a = 0; a < 3; a++) {
                fijk[a] = d.roi_offset[a] + d.vpr[a] * p[a] + q[a];
                if (fijk[a] >= d.roi_offset[a] + d.roi_dim[a]) skip = 1;
            }
            if (skip) continue;
            double fxyz[3], dxyz[3];
            for (int a = 0; a < 3; a++)
                fxyz[a] = d.origin[a] + fijk[0] * fx.step[3*a+0]
                    + fijk[1] * fx.step[3*a+1] + fijk[2] * fx.step[3*a+2];
            dinterp (dxyz, &d, p, q, coeff);
            /* the moved coefficient's share: its basis weight times delta */
            const int64_t i = kk[0] - p[0], j = kk[1] - p[1], k = kk[2] - p[2];
            dxyz[comp] += d.bx[q[0]*4+i] * d.by[q[1]*4+j] * d.bz[q[2]*4+k] * delta;
            if (froi && !inside_roi_d (fxyz, froi)) continue;
            dsample_t s;
            if (!dsample (&s, fxyz, dxyz, &mv)) continue;
            if (mroi) {
                double mxyz[3] = {fxyz[0] + dxyz[0], fxyz[1] + dxyz[1], fxyz[2] + dxyz[2]};
                if (!inside_roi_d (mxyz, mroi)) continue;
            }
            double t[3], mijk[3];
            for (int a = 0; a < 3; a++) t[a] = fxyz[a] + dxyz[a] - mv.origin[a];
            for (int a = 0; a < 3; a++)
                mijk[a] = t[0] * mv.proj[3*a+0] + t[1] * mv.proj[3*a+1] + t[2] * mv.proj[3*a+2];
            for (int a = 0; a < 3; a++)
                if (mijk[a] < 0 || mijk[a] >= (double) (mv.dim[a] - 1)) { nclamp++; break; }
            double a8[8];
            dcorners (a8, &s, &mv);
            const double* l1 = s.l1;
            const double* l2 = s.l2;
            double m_val = l1[0]*l1[1]*l1[2]*a8[0] + l2[0]*l1[1]*l1[2]*a8[1]
                + l1[0]*l2[1]*l1[2]*a8[2] + l2[0]*l2[1]*l1[2]*a8[3]
                + l1[0]*l1[1]*l2[2]*a8[4] + l2[0]*l1[1]*l2[2]*a8[5]
                + l1[0]*l2[1]*l2[2]*a8[6] + l2[0]*l2[1]*l2[2]*a8[7];
            int64_t fidx = fijk[0] + fx.dim[0] * (fijk[1] + fx.dim[1] * fijk[2]);
            double diff = m_val - (double) fx.img[fidx];
            raw += diff * diff;
            n++;
        }
    }
    *raw_out = raw;
    *n_out = n;
    *nclamp_out = nclamp;
    dgeo_free (&d);
    return 0;
}

/* ---------------------------------------------------------------------
   Mattes mutual information
   --------------------------------------------------------------------- */

/* delta = (max-min)/(bins-1), offset = min - 0.5*delta.  Follows
   bspline_initialize_mi_hist_eqsp, src/register/joint_histogram.cxx:115-139. */
void
orc_hist_eqsp (const float* img, int64_t npix, int64_t bins, float* offset,
    float* delta)
{
    float mn = img[0], mx = img[0];
    for (int64_t i = 0; i < npix; i++) {
        if (img[i] < mn) mn = img[i];
        else if (img[i] > mx) mx = img[i];
    }
    *delta = (mx - mn) / (float) (bins - 1);
    *offset = (float) (mn - 0.5 * *delta);
}

/* V-optimal histogram.  Follows bspline_initialize_mi_hist_vopt and
   vopt_bin_error (src/register/joint_histogram.cxx:62-82, 195-324): the
   image range split into ORC_VOPT_KEYS equal keys (the grid of
   bspline_initialize_mi_hist_eqsp with keys for bins), each key's voxel
   count and average of (v - offset) (a key with no voxels takes the
   previous key's average), running sums of the averages, of their squares
   and of the counts, then a dynamic programme over (bins, keys) for the
   contiguous grouping of keys into bins of least total error, a bin's
   error being ssq - s^2 / n over its prefix-sum differences (and DBL_MAX
   when it holds less than one voxel), ties going to the last split
   examined; finally the key table by walking the splits back. */
static double
vopt_err (int64_t start, int64_t end, const double* s, const double* ssq, const double* cnt)
{
    const double sq = ssq[end] - ssq[start];
    const double d = s[end] - s[start];
    const double nk = (double) end - (double) start + 1.0;
    const double n = cnt[end] - cnt[start];
    const double v = sq - (d * d) / nk;
    return n < 1.0 ? DBL_MAX : v;
}

void
orc_vopt_partition (const double* s, const double* ssq, const double* cnt, int64_t keys,
    int64_t bins, int32_t* key_lut)
{
    double* err = (double*) malloc (sizeof (double) * bins * keys);
    int64_t* split = (int64_t*) calloc ((size_t) (bins * keys), sizeof (int64_t));
    for (int64_t i = 0; i < keys; i++) err[i] = vopt_err (0, i, s, ssq, cnt);
    for (int64_t j = 1; j < bins; j++) {
        for (int64_t i = 0; i < keys; i++) {
            double best = DBL_MAX;
            int64_t arg = 0;
            for (int64_t k = 0; k < i; k++) {
                const double c = err[(j - 1) * keys + k] + vopt_err (k + 1, i, s, ssq, cnt);
                if (c <= best) {
                    best = c;
                    arg = k;
                }
            }
            err[j * keys + i] = best;
            split[j * keys + i] = arg;
        }
    }
    for (int64_t i = 0; i < keys; i++) key_lut[i] = 0;
    int64_t hi = keys - 1;
    for (int64_t j = bins - 1; j >= 0; j--) {
        const int64_t nx = split[j * keys + hi];
        const int64_t lo = j == 0 ? 0 : nx + 1;
        for (int64_t i = lo; i <= hi; i++) key_lut[i] = (int32_t) j;
        hi = nx;
    }
    free (err);
    free (split);
}

int
orc_hist_vopt (const float* img, int64_t npix, int64_t bins, float* offset, float* delta,
    int32_t* key_lut, double* key_count, double* key_sum)
{
    const int64_t keys = ORC_VOPT_KEYS;
    if (bins < 2 || bins > keys || npix < 1) return -1;
    float mn = img[0], mx = img[0];
    for (int64_t i = 1; i < npix; i++) {
        if (img[i] < mn) mn = img[i];
        else if (img[i] > mx) mx = img[i];
    }
    if (!(mx > mn)) return -1;
    const float dl = (mx - mn) / (float) (keys - 1);
    const float off = (float) (mn - 0.5 * dl);
    double* cnt1 = (double*) calloc ((size_t) keys, sizeof (double));
    double* avg = (double*) calloc ((size_t) keys, sizeof (double));
    for (int64_t i = 0; i < npix; i++) {
        const float x = img[i] - off;
        int64_t k = (int64_t) floorf (x / dl);
        k = k < 0 ? 0 : (k >= keys ? keys - 1 : k);
        cnt1[k] += 1.0;
        avg[k] += x;
    }
    if (key_count) memcpy (key_count, cnt1, sizeof (double) * keys);
    if (key_sum) memcpy (key_sum, avg, sizeof (double) * keys);
    for (int64_t i = 0; i < keys; i++) {
        if (cnt1[i] > 0) avg[i] = avg[i] / cnt1[i];
        else if (i > 0) avg[i] = avg[i - 1];
    }
    double* s = (double*) malloc (sizeof (double) * keys);
    double* ssq = (double*) malloc (sizeof (double) * keys);
    double* cnt = (double*) malloc (sizeof (double) * keys);
    s[0] = avg[0];
    ssq[0] = avg[0] * avg[0];
    cnt[0] = cnt1[0];
    for (int64_t i = 1; i < keys; i++) {
        s[i] = s[i - 1] + avg[i];
        ssq[i] = ssq[i - 1] + avg[i] * avg[i];
        cnt[i] = cnt[i - 1] + cnt1[i];
    }
    orc_vopt_partition (s, ssq, cnt, keys, bins, key_lut);
    *offset = off;
    *delta = dl;
    free (cnt1);
    free (avg);
    free (s);
    free (ssq);
    free (cnt);
    return 0;
}

/* Fixed-point partial-volume weights: q = floor(w*2^32 + 0.5), and the
   remainder 2^32 - sum(q) is added to the first largest weight, so every
   voxel adds exactly 2^32 to the moving and joint histograms.  This
   quantization has no counterpart in the reference (whose histogram is a
   double sum); it makes the accumulation order-independent. */
void
orc_pv_quantize (const float w[8], int64_t q[8])
{
    int64_t sum = 0;
    int amax = 0;
    for (int n = 0; n < 8; n++) {
        q[n] = (int64_t) ((double) w[n] * 4294967296.0 + 0.5);
        sum += q[n];
        if (w[n] > w[amax]) amax = n;
    }
    q[amax] += ((int64_t) 1 << 32) - sum;
}

static inline int
hist_bin (float v, float offset, float delta)
{
    return (int) floorf ((v - offset) / delta);
}

/* The bin of a fixed / moving value: its equal-spacing bin, or with a key
   table the bin of its key (src/register/joint_histogram.cxx:458-481). */
static inline int
key_bin (int b, const int32_t* keys)
{
    if (!keys) return b;
    return keys[b < 0 ? 0 : (b >= ORC_VOPT_KEYS ? ORC_VOPT_KEYS - 1 : b)];
}

static inline int
fbin_of (const orc_hist* hp, float v)
{
    return key_bin (hist_bin (v, hp->f_offset, hp->f_delta), hp->f_keys);
}

static inline int
mbin_of (const orc_hist* hp, float v)
{
    return key_bin (hist_bin (v, hp->m_offset, hp->m_delta), hp->m_keys);
}

/* PV-8 weights w[n], n in x,y,z bit order.  Follows
   Joint_histogram::add_pvi_8, src/register/joint_histogram.cxx:423-487. */
static inline void
pv_weights (float w[8], const float l1[3], const float l2[3])
{
    w[0] = l1[0] * l1[1] * l1[2];
    w[1] = l2[0] * l1[1] * l1[2];
    w[2] = l1[0] * l2[1] * l1[2];
    w[3] = l2[0] * l2[1] * l1[2];
    w[4] = l1[0] * l1[1] * l2[2];
    w[5] = l2[0] * l1[1] * l2[2];
    w[6] = l1[0] * l2[1] * l2[2];
    w[7] = l2[0] * l2[1] * l2[2];
}

static inline void
pv_neighbors (int64_t n[8], int64_t mvf, const int64_t dim[3])
{
    n[0] = mvf;
    n[1] = n[0] + 1;
    n[2] = n[0] + dim[0];
    n[3] = n[2] + 1;
    n[4] = n[0] + dim[0] * dim[1];
    n[5] = n[4] + 1;
    n[6] = n[4] + dim[0];
    n[7] = n[6] + 1;
}

/* S = -sum_{j>thresh} j*logf(N*j/(m*f)) / N, thresh = 0.001/(Bf*Bm),
   logf on the double argument rounded to float.  Follows
   mi_hist_score_omp, src/register/bspline_mi.cxx:285-310. */
static float
mi_hist_score (const double* f_hist, const double* m_hist, const double* j_hist,
    int64_t fb, int64_t mb, int64_t num_vox)
{
    double fnv = (double) num_vox;
    double score = 0;
    double thresh = 0.001 / (double) (mb * fb);
    for (int64_t jb = 0; jb < fb * mb; jb++) {
        int64_t m = jb % mb, f = jb / mb;
        if (j_hist[jb] > thresh) {
            score -= j_hist[jb] * logf ((float) (fnv * j_hist[jb] / (m_hist[m] * f_hist[f])));
        }
    }
    score = score / fnv;
    return (float) score;
}

/* PV-8 derivative of the score w.r.t. the voxel position, divided by N and
   projected to mm (in place, as the reference does).  Follows
   bspline_mi_pvi_8_dc_dv_dcos, src/register/bspline_mi.txx:9-120. */
static inline void
mi_dc_dv (float dc_dv[3], const double* f_hist, const double* m_hist,
    const double* j_hist, const orc_hist* hp, float f_val, const float* m_img,
    const int64_t n[8], float num_vox_f, float score, const float l1[3],
    const float l2[3], const vol_t* mv)
{
    float dw[24];
    dc_dv[0] = dc_dv[1] = dc_dv[2] = 0.0f;
    dw[0] = (-1) * l1[1] * l1[2];  dw[1] = l1[0] * (-1) * l1[2];  dw[2] = l1[0] * l1[1] * (-1);
    dw[3] = (+1) * l1[1] * l1[2];  dw[4] = l2[0] * (-1) * l1[2];  dw[5] = l2[0] * l1[1] * (-1);
    dw[6] = (-1) * l2[1] * l1[2];  dw[7] = l1[0] * (+1) * l1[2];  dw[8] = l1[0] * l2[1] * (-1);
    dw[9] = (+1) * l2[1] * l1[2];  dw[10] = l2[0] * (+1) * l1[2]; dw[11] = l2[0] * l2[1] * (-1);
    dw[12] = (-1) * l1[1] * l2[2]; dw[13] = l1[0] * (-1) * l2[2]; dw[14] = l1[0] * l1[1] * (+1);
    dw[15] = (+1) * l1[1] * l2[2]; dw[16] = l2[0] * (-1) * l2[2]; dw[17] = l2[0] * l1[1] * (+1);
    dw[18] = (-1) * l2[1] * l2[2]; dw[19] = l1[0] * (+1) * l2[2]; dw[20] = l1[0] * l2[1] * (+1);
    dw[21] = (+1) * l2[1] * l2[2]; dw[22] = l2[0] * (+1) * l2[2]; dw[23] = l2[0] * l2[1] * (+1);
    int fb = fbin_of (hp, f_val);
    int64_t off = (int64_t) fb * hp->moving_bins;
    for (int pv = 0; pv < 8; pv++) {
        int mb = mbin_of (hp, m_img[n[pv]]);
        int64_t jb = off + mb;
        if (j_hist[jb] > 0.0001) {
            float ds = logf ((float) (((double) num_vox_f * j_hist[jb])
                    / (f_hist[fb] * m_hist[mb]))) - score;
            dc_dv[0] -= dw[3*pv+0] * ds;
            dc_dv[1] -= dw[3*pv+1] * ds;
            dc_dv[2] -= dw[3*pv+2] * ds;
        }
    }
    dc_dv[0] = dc_dv[0] / num_vox_f;
    dc_dv[1] = dc_dv[1] / num_vox_f;
    dc_dv[2] = dc_dv[2] / num_vox_f;
    const float* P = mv->proj;
    dc_dv[0] = dc_dv[0] * P[0] + dc_dv[1] * P[1] + dc_dv[2] * P[2];
    dc_dv[1] = dc_dv[0] * P[3] + dc_dv[1] * P[4] + dc_dv[2] * P[5];
    dc_dv[2] = dc_dv[0] * P[6] + dc_dv[1] * P[7] + dc_dv[2] * P[8];
}

/* Counts voxels whose pass-1 position (LOOP_Z/Y/X recurrences over the
   fixed image, src/base/volume_macros.h:28-70) differs from the pass-2
   position (POSITION_FROM_COORDS with the grid origin,
   src/base/volume_macros.h:82-96).  Zero for geometries whose positions
   are exact in float. */
int64_t
orc_pass1_position_mismatch (const orc_vol_hdr* fh, const orc_geom* g)
{
    vol_t fx;
    vol_init (&fx, fh, 0);
    const float* o = fx.origin;
    const float* s = fx.step;
    int64_t bad = 0;
    float fxyz[3];
    int64_t fijk[3];
    fxyz[2] = o[2];
    for (fijk[2] = 0; fijk[2] < fx.dim[2]; ) {
        fxyz[1] = o[1] + fijk[2] * s[5];
        for (fijk[1] = 0; fijk[1] < fx.dim[1]; ) {
            fxyz[0] = o[0] + fijk[2] * s[2] + fijk[1] * s[1];
            for (fijk[0] = 0; fijk[0] < fx.dim[0]; ) {
                for (int d = 0; d < 3; d++) {
                    float direct = g->img_origin[d] + fijk[0] * s[3*d+0]
                        + fijk[1] * s[3*d+1] + fijk[2] * s[3*d+2];
                    if (direct != fxyz[d]) { bad++; break; }
                }
                ++fijk[0];
                fxyz[0] += s[0]; fxyz[1] += s[3]; fxyz[2] += s[6];
            }
            ++fijk[1];
            fxyz[2] = o[2] + fijk[2] * s[8] + fijk[1] * s[7];
            fxyz[1] = o[1] + fijk[2] * s[5] + fijk[1] * s[4];
        }
        ++fijk[2];
        fxyz[2] = o[2] + fijk[2] * s[8];
    }
    return bad;
}

/* Mattes MI, flavor g: serial pass 1 builds the PV-8 histograms, the score
   is taken from them, and pass 2 (OpenMP over tiles) condenses the PV
   derivative.  Follows bspline_score_g_mi,
   src/register/bspline_mi.cxx:1505-1737.  The fixed-point histograms
   (fx_f_hist counts, fx_j_hist in units of 2^-32) are filled from the same
   float weights by orc_pv_quantize. */
int
orc_mi_score (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, const orc_hist* hp, int nthreads, const orc_rois* rois,
    float* score, float* grad, int64_t* num_vox,
    double* f_hist, double* m_hist, double* j_hist,
    int64_t* fx_f_hist, int64_t* fx_j_hist)
{
    bxf_t b;
    vol_t fx, mv;
    roi_t fr_, mr_;
    const roi_t* froi = rois ? roi_init (&fr_, rois->fixed_hdr, rois->fixed) : 0;
    const roi_t* mroi = rois ? roi_init (&mr_, rois->moving_hdr, rois->moving) : 0;
    if (bxf_init (&b, g, coeff)) return -1;
    vol_init (&fx, fh, f_img);
    vol_init (&mv, mh, m_img);
    set_threads (nthreads);
    int64_t fbins = hp->fixed_bins, mbins = hp->moving_bins;
    memset (f_hist, 0, sizeof(double) * fbins);
    memset (m_hist, 0, sizeof(double) * mbins);
    memset (j_hist, 0, sizeof(double) * fbins * mbins);
    if (fx_f_hist) memset (fx_f_hist, 0, sizeof(int64_t) * fbins);
    if (fx_j_hist) memset (fx_j_hist, 0, sizeof(int64_t) * fbins * mbins);
    memset (grad, 0, sizeof(float) * b.num_coeff);
    int64_t nv = 0;

    /* PASS 1 (serial), LOOP_Z/Y/X position recurrences */
    {
        const float* o = fx.origin;
        const float* s = fx.step;
        float fxyz[3];
        int64_t fijk[3], p[3], q[3];
        fxyz[2] = o[2];
        for (fijk[2] = 0; fijk[2] < fx.dim[2]; ) {
            p[2] = fijk[2] / b.vpr[2]; q[2] = fijk[2] % b.vpr[2];
            fxyz[1] = o[1] + fijk[2] * s[5];
            for (fijk[1] = 0; fijk[1] < fx.dim[1]; ) {
                p[1] = fijk[1] / b.vpr[1]; q[1] = fijk[1] % b.vpr[1];
                fxyz[0] = o[0] + fijk[2] * s[2] + fijk[1] * s[1];
                for (fijk[0] = 0; fijk[0] < fx.dim[0]; ) {
                    p[0] = fijk[0] / b.vpr[0]; q[0] = fijk[0] % b.vpr[0];
                    int64_t pidx = p[0] + b.rdims[0] * (p[1] + b.rdims[1] * p[2]);
                    float dxyz[3], mijk[3];
                    /* flavor h: the ROIs of bspline_score_h_mi,
                       src/register/bspline_mi.cxx:1304-1326 */
                    int keep = !froi || inside_roi (fxyz, froi);
                    if (keep) interp_pix_c (dxyz, &b, pidx, q);
                    if (keep) keep = find_correspondence (mijk, fxyz, dxyz, &mv);
                    if (keep && mroi) {
                        float mxyz[3] = {fxyz[0] + dxyz[0], fxyz[1] + dxyz[1], fxyz[2] + dxyz[2]};
                        keep = inside_roi (mxyz, mroi);
                    }
                    if (keep) {
                        int64_t mf[3], mr[3];
                        float l1[3], l2[3];
                        li_clamp_3d (mijk, mf, mr, l1, l2, &mv);
                        int64_t fidx = fijk[0] + fx.dim[0] * (fijk[1] + fx.dim[1] * fijk[2]);
                        int64_t mvf = mf[0] + mv.dim[0] * (mf[1] + mv.dim[1] * mf[2]);
                        float w[8];
                        int64_t n[8], qw[8];
                        pv_weights (w, l1, l2);
                        pv_neighbors (n, mvf, mv.dim);
                        orc_pv_quantize (w, qw);
                        int fb = fbin_of (hp, f_img[fidx]);
                        f_hist[fb]++;
                        if (fx_f_hist) fx_f_hist[fb]++;
                        int64_t off = (int64_t) fb * mbins;
                        for (int pv = 0; pv < 8; pv++) {
                            int mb = mbin_of (hp, m_img[n[pv]]);
                            m_hist[mb] += w[pv];
                            j_hist[off + mb] += w[pv];
                            if (fx_j_hist) fx_j_hist[off + mb] += qw[pv];
                        }
                        nv++;
                    }
                    ++fijk[0];
                    fxyz[0] += s[0]; fxyz[1] += s[3]; fxyz[2] += s[6];
                }
                ++fijk[1];
                fxyz[2] = o[2] + fijk[2] * s[8] + fijk[1] * s[7];
                fxyz[1] = o[1] + fijk[2] * s[5] + fijk[1] * s[4];
            }
            ++fijk[2];
            fxyz[2] = o[2] + fijk[2] * s[8];
        }
    }

    float sc = mi_hist_score (f_hist, m_hist, j_hist, fbins, mbins, nv);
    float num_vox_f = (float) nv;

    /* PASS 2 (tiles) */
    int64_t csz = 64 * b.num_knots;
    float* cx = (float*) calloc (csz, sizeof(float));
    float* cy = (float*) calloc (csz, sizeof(float));
    float* cz = (float*) calloc (csz, sizeof(float));
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t pidx = 0; pidx < b.num_tiles; pidx++) {
        float sx[64], sy[64], sz[64];
        memset (sx, 0, sizeof sx); memset (sy, 0, sizeof sy); memset (sz, 0, sizeof sz);
        int64_t p[3], q[3], fijk[3];
        p[2] = pidx / (b.rdims[0] * b.rdims[1]);
        p[1] = (pidx - p[2] * b.rdims[0] * b.rdims[1]) / b.rdims[0];
        p[0] = pidx - p[2] * b.rdims[0] * b.rdims[1] - p[1] * b.rdims[0];
        for (q[2] = 0; q[2] < b.vpr[2]; q[2]++)
        for (q[1] = 0; q[1] < b.vpr[1]; q[1]++)
        for (q[0] = 0; q[0] < b.vpr[0]; q[0]++) {
            for (int d = 0; d < 3; d++)
                fijk[d] = b.roi_offset[d] + b.vpr[d] * p[d] + q[d];
            if (fijk[0] >= b.roi_offset[0] + b.roi_dim[0]) continue;
            if (fijk[1] >= b.roi_offset[1] + b.roi_dim[1]) continue;
            if (fijk[2] >= b.roi_offset[2] + b.roi_dim[2]) continue;
            float fxyz[3];
            for (int d = 0; d < 3; d++)
                fxyz[d] = b.img_origin[d] + fijk[0] * fx.step[3*d+0]
                    + fijk[1] * fx.step[3*d+1] + fijk[2] * fx.step[3*d+2];
            if (froi && !inside_roi (fxyz, froi)) continue;
            float dxyz[3], mijk[3];
            interp_pix_c (dxyz, &b, pidx, q);
            if (!find_correspondence (mijk, fxyz, dxyz, &mv)) continue;
            if (mroi) {
                float mxyz[3] = {fxyz[0] + dxyz[0], fxyz[1] + dxyz[1], fxyz[2] + dxyz[2]};
                if (!inside_roi (mxyz, mroi)) continue;
            }
            int64_t mf[3], mr[3];
            float l1[3], l2[3];
            li_clamp_3d (mijk, mf, mr, l1, l2, &mv);
            int64_t fidx = fijk[0] + fx.dim[0] * (fijk[1] + fx.dim[1] * fijk[2]);
            int64_t mvf = mf[0] + mv.dim[0] * (mf[1] + mv.dim[1] * mf[2]);
            int64_t n[8];
            pv_neighbors (n, mvf, mv.dim);
            float dc_dv[3];
            mi_dc_dv (dc_dv, f_hist, m_hist, j_hist, hp, f_img[fidx], m_img, n,
                num_vox_f, sc, l1, l2, &mv);
            update_sets_b (sx, sy, sz, q, dc_dv, &b);
        }
        sort_sets (cx, cy, cz, sx, sy, sz, pidx, &b);
    }
    condense_grad (grad, cx, cy, cz, &b);
    free (cx); free (cy); free (cz);
    *score = sc;
    *num_vox = nv;
    bxf_free (&b);
    return 0;
}

/* The same MI algorithm in double precision: double weights and
   histograms, natural log in double, double condense and reduce.
   Positions use the direct formula in both passes.  With hist_in (f, m and
   joint histograms, in that order, and their voxel count) pass 1 is
   skipped and the score and pass 2 run on those histograms. */
static int
mi_f64 (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, const orc_hist* hp, int nthreads, const orc_rois* rois,
    const double* hist_in, int64_t nv_in,
    double* score, double* grad, int64_t* num_vox)
{
    dgeo_t d;
    dvol_t fx, mv;
    roi_t fr_, mr_;
    const roi_t* froi = rois ? roi_init (&fr_, rois->fixed_hdr, rois->fixed) : 0;
    const roi_t* mroi = rois ? roi_init (&mr_, rois->moving_hdr, rois->moving) : 0;
    if (dgeo_init (&d, g)) return -1;
    dvol_init (&fx, fh, f_img);
    dvol_init (&mv, mh, m_img);
    set_threads (nthreads);
    int64_t fbins = hp->fixed_bins, mbins = hp->moving_bins;
    double* fh_ = (double*) calloc (fbins, sizeof(double));
    double* mh_ = (double*) calloc (mbins, sizeof(double));
    double* jh_ = (double*) calloc (fbins * mbins, sizeof(double));
    double* tiles = (double*) calloc (d.num_tiles * 64 * 3, sizeof(double));
    int64_t nv = 0;
    if (hist_in) {
        memcpy (fh_, hist_in, sizeof(double) * fbins);
        memcpy (mh_, hist_in + fbins, sizeof(double) * mbins);
        memcpy (jh_, hist_in + fbins + mbins, sizeof(double) * fbins * mbins);
        nv = nv_in;
    }

    for (int pass = hist_in ? 2 : 1; pass <= 2; pass++) {
        double sc = 0;
        if (pass == 2) {
            double fnv = (double) nv, thresh = 0.001 / (double) (fbins * mbins);
            for (int64_t jb = 0; jb < fbins * mbins; jb++) {
                if (jh_[jb] > thresh)
                    sc -= jh_[jb] * log (fnv * jh_[jb] / (mh_[jb % mbins] * fh_[jb / mbins]));
            }
            sc /= fnv;
            *score = sc;
        }
#pragma omp parallel for schedule(dynamic, 4) if (pass == 2)
        for (int64_t t = 0; t < d.num_tiles; t++) {
            int64_t p[3], q[3], fijk[3];
            p[2] = t / (d.rdims[0] * d.rdims[1]);
            p[1] = (t / d.rdims[0]) % d.rdims[1];
            p[0] = t % d.rdims[0];
            double* tg = &tiles[t * 64 * 3];
            for (q[2] = 0; q[2] < d.vpr[2]; q[2]++)
            for (q[1] = 0; q[1] < d.vpr[1]; q[1]++)
            for (q[0] = 0; q[0] < d.vpr[0]; q[0]++) {
                int skip = 0;
                for (int a = 0; a < 3; a++) {
                    fijk[a] = d.roi_offset[a] + d.vpr[a] * p[a] + q[a];
                    if (fijk[a] >= d.roi_offset[a] + d.roi_dim[a]) skip = 1;
                }
                if (skip) continue;
                double fxyz[3], dxyz[3];
                for (int a = 0; a < 3; a++)
                    fxyz[a] = d.origin[a] + fijk[0] * fx.step[3*a+0]
                        + fijk[1] * fx.step[3*a+1] + fijk[2] * fx.step[3*a+2];
                if (froi && !inside_roi_d (fxyz, froi)) continue;
                dinterp (dxyz, &d, p, q, coeff);
                dsample_t s;
                if (!dsample (&s, fxyz, dxyz, &mv)) continue;
                if (mroi) {
                    double mxyz[3] = {fxyz[0] + dxyz[0], fxyz[1] + dxyz[1], fxyz[2] + dxyz[2]};
                    if (!inside_roi_d (mxyz, mroi)) continue;
                }
                int64_t fidx = fijk[0] + fx.dim[0] * (fijk[1] + fx.dim[1] * fijk[2]);
                const double* l1 = s.l1;
                const double* l2 = s.l2;
                double w[8] = {
                    l1[0]*l1[1]*l1[2], l2[0]*l1[1]*l1[2], l1[0]*l2[1]*l1[2], l2[0]*l2[1]*l1[2],
                    l1[0]*l1[1]*l2[2], l2[0]*l1[1]*l2[2], l1[0]*l2[1]*l2[2], l2[0]*l2[1]*l2[2]};
                int64_t n[8];
                pv_neighbors (n, s.mvf, mv.dim);
                int fb = fbin_of (hp, f_img[fidx]);
                int64_t off = (int64_t) fb * mbins;
                if (pass == 1) {
                    fh_[fb] += 1;
                    for (int pv = 0; pv < 8; pv++) {
                        int mb = mbin_of (hp, m_img[n[pv]]);
                        mh_[mb] += w[pv];
                        jh_[off + mb] += w[pv];
                    }
                    nv++;
                } else {
                    double dw[24] = {
                        -l1[1]*l1[2], -l1[0]*l1[2], -l1[0]*l1[1],
                        +l1[1]*l1[2], -l2[0]*l1[2], -l2[0]*l1[1],
                        -l2[1]*l1[2], +l1[0]*l1[2], -l1[0]*l2[1],
                        +l2[1]*l1[2], +l2[0]*l1[2], -l2[0]*l2[1],
                        -l1[1]*l2[2], -l1[0]*l2[2], +l1[0]*l1[1],
                        +l1[1]*l2[2], -l2[0]*l2[2], +l2[0]*l1[1],
                        -l2[1]*l2[2], +l1[0]*l2[2], +l1[0]*l2[1],
                        +l2[1]*l2[2], +l2[0]*l2[2], +l2[0]*l2[1]};
                    double dc[3] = {0, 0, 0};
                    for (int pv = 0; pv < 8; pv++) {
                        int mb = mbin_of (hp, m_img[n[pv]]);
                        int64_t jb = off + mb;
                        if (jh_[jb] > 0.0001) {
                            double ds = log ((double) nv * jh_[jb] / (fh_[fb] * mh_[mb])) - *score;
                            dc[0] -= dw[3*pv+0] * ds;
                            dc[1] -= dw[3*pv+1] * ds;
                            dc[2] -= dw[3*pv+2] * ds;
                        }
                    }
                    for (int a = 0; a < 3; a++) dc[a] /= (double) nv;
                    /* projected in place, component by component, as the
                       reference does (src/register/bspline_mi.txx:107-109) */
                    for (int a = 0; a < 3; a++)
                        dc[a] = dc[0] * mv.proj[3*a+0] + dc[1] * mv.proj[3*a+1] + dc[2] * mv.proj[3*a+2];
                    dscatter (tg, &d, p, q, dc);
                }
            }
        }
    }
    dreduce (grad, tiles, &d);
    *num_vox = nv;
    free (fh_); free (mh_); free (jh_); free (tiles);
    dgeo_free (&d);
    return 0;
}

int
orc_mi_score_f64 (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, const orc_hist* hp, int nthreads, const orc_rois* rois,
    double* score, double* grad, int64_t* num_vox)
{
    return mi_f64 (g, fh, f_img, mh, m_img, coeff, hp, nthreads, rois, 0, 0, score, grad, num_vox);
}

/* FP64 score and pass 2 on given histograms (the float flavor's: the PV
   histogram is fixed to the reference's, bin for bin, so this isolates the
   arithmetic of the score and the gradient pass from the float rounding of
   the histogram weights). */
int
orc_mi_score_f64_hist (const orc_geom* g,
    const orc_vol_hdr* fh, const float* f_img,
    const orc_vol_hdr* mh, const float* m_img,
    const float* coeff, const orc_hist* hp, int nthreads, const orc_rois* rois,
    const double* hist_in, int64_t nv_in,
    double* score, double* grad, int64_t* num_vox)
{
    return mi_f64 (g, fh, f_img, mh, m_img, coeff, hp, nthreads, rois, hist_in, nv_in, score, grad,
        num_vox);
}

/* ---------------------------------------------------------------------
   Analytic regularization (flavor c)
   --------------------------------------------------------------------- */

/* V[4j+i] = sum_{a,b} I[a+b] Qn1[4j+a] Qn2[4i+b], I[n] = gs^(n+1)/(n+1):
   the integral over one region of the product of two basis polynomials.
   Follows eval_integral, src/register/bspline_regularize_analytic.cxx:167-199
   (the outer product of vec_outer, src/sys/plm_math.h:204-212). */
static void
reg_eval_integral (double* V, const double* Qn1, const double* Qn2, double gs)
{
    double I[7];
    I[0] = gs;
    I[1] = (1.0/2.0) * (gs * gs);
    I[2] = (1.0/3.0) * (gs * gs * gs);
    I[3] = (1.0/4.0) * (gs * gs * gs * gs);
    I[4] = (1.0/5.0) * (gs * gs * gs * gs * gs);
    I[5] = (1.0/6.0) * (gs * gs * gs * gs * gs * gs);
    I[6] = (1.0/7.0) * (gs * gs * gs * gs * gs * gs * gs);
    for (int j = 0; j < 4; j++) {
        for (int i = 0; i < 4; i++) {
            double S[16];
            for (int a = 0; a < 4; a++)
                for (int b = 0; b < 4; b++)
                    S[4*a + b] = Qn1[4*j + a] * Qn2[4*i + b];
            V[4*j + i] = (I[0] * S[0])
                + (I[1] * (S[1] + S[4]))
                + (I[2] * (S[2] + S[5] + S[8]))
                + (I[3] * (S[3] + S[6] + S[9] + S[12]))
                + (I[4] * (S[7] + S[10] + S[13]))
                + (I[5] * (S[11] + S[14]))
                + (I[6] * (S[15]));
        }
    }
}

static void
reg_mat_mult (double* m1, const double* m2, const double* m3)
{
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double acc = 0.0;
            for (int k = 0; k < 4; k++) acc += m2[4*i + k] * m3[4*k + j];
            m1[4*i + j] = acc;
        }
}

/* Q[d][n] = B R_d delta_n: the basis polynomial coefficients scaled by the
   reciprocal grid spacing, and their n-th derivatives.  Follows
   init_analytic, src/register/bspline_regularize_analytic.cxx:201-281. */
static void
reg_init_q (double Q[3][4][16], const double gs[3])
{
    static const double B[16] = {
        1.0/6.0, -1.0/2.0,  1.0/2.0, -1.0/6.0,
        2.0/3.0,  0.0    , -1.0    ,  1.0/2.0,
        1.0/6.0,  1.0/2.0,  1.0/2.0, -1.0/2.0,
        0.0    ,  0.0    ,  0.0    ,  1.0/6.0
    };
    static const double delta[3][16] = {
        {0,0,0,0, 1,0,0,0, 0,2,0,0, 0,0,3,0},
        {0,0,0,0, 0,0,0,0, 2,0,0,0, 0,6,0,0},
        {0,0,0,0, 0,0,0,0, 0,0,0,0, 6,0,0,0}
    };
    for (int d = 0; d < 3; d++) {
        const double r = 1.0 / gs[d];
        double R[16] = {1.0, 0, 0, 0, 0, r, 0, 0, 0, 0, r*r, 0, 0, 0, 0, r*r*r};
        reg_mat_mult (Q[d][0], B, R);
        for (int n = 1; n < 4; n++) reg_mat_mult (Q[d][n], Q[d][0], delta[n-1]);
    }
}

/* V = X (x) Y (x) Z as get_Vmatrix lays it out: row 16a + 4c + j takes X's
   row a, Y's row c and Z's row j (src/register/bspline_regularize_analytic.cxx:283-338),
   scaled by lambda (scale_Vmatrix, :340-346). */
static void
reg_vmatrix (double* V, const double* X, const double* Y, const double* Z, double lambda)
{
    for (int a = 0; a < 4; a++)
        for (int c = 0; c < 4; c++)
            for (int j = 0; j < 4; j++)
                for (int b = 0; b < 4; b++)
                    for (int d = 0; d < 4; d++)
                        for (int i = 0; i < 4; i++)
                            V[64*(16*a + 4*c + j) + (16*b + 4*d + i)]
                                = lambda * (X[4*a + b] * (Y[4*c + d] * Z[4*j + i]));
}

/* The 32 V matrices of Bspline_regularize::analytic_init
   (src/register/bspline_regularize_analytic.cxx:638-940): (x order,
   y order, z order) of each factor's two Q matrices and the weight. */
static void
reg_analytic_init (double* V, const double gs[3], const orc_reg_parms* rp)
{
    double Q[3][4][16];
    reg_init_q (Q, gs);
    /* the weights as the reference forms them, in float */
    const float cp = rp->curvature_penalty, mix = rp->curvature_mixed_weight;
    const float dp = rp->diffusion_penalty, m = rp->linear_elastic_multiplier;
    const float l1 = rp->lame_coefficient_1, l2 = rp->lame_coefficient_2;
    const float tdp = rp->total_displacement_penalty, top = rp->third_order_penalty;
    const float cpm = cp * mix, e1 = l1 * m, e2 = l2 * m;
    const float ed = (m) * (2*l1 + l2)/2, eh = (m) * (l1)/2;
    /* {x1, x2, y1, y2, z1, z2} derivative orders, weight */
    const struct { int o[6]; float w; } T[32] = {
        {{2,2, 0,0, 0,0}, cp}, {{0,0, 2,2, 0,0}, cp}, {{0,0, 0,0, 2,2}, cp},
        {{1,1, 1,1, 0,0}, cpm}, {{1,1, 0,0, 1,1}, cpm}, {{0,0, 1,1, 1,1}, cpm},
        {{1,1, 0,0, 0,0}, dp}, {{0,0, 1,1, 0,0}, dp}, {{0,0, 0,0, 1,1}, dp},
        {{1,0, 0,1, 0,0}, e1}, {{0,1, 1,0, 0,0}, e2},
        {{1,0, 0,0, 0,1}, e1}, {{0,1, 0,0, 1,0}, e2},
        {{0,0, 1,0, 0,1}, e1}, {{0,0, 0,1, 1,0}, e2},
        {{1,1, 0,0, 0,0}, ed}, {{0,0, 1,1, 0,0}, ed}, {{0,0, 0,0, 1,1}, ed},
        {{1,1, 0,0, 0,0}, eh}, {{0,0, 1,1, 0,0}, eh}, {{0,0, 0,0, 1,1}, eh},
        {{0,0, 0,0, 0,0}, tdp},
        {{3,3, 0,0, 0,0}, top}, {{0,0, 3,3, 0,0}, top}, {{0,0, 0,0, 3,3}, top},
        {{2,2, 1,1, 0,0}, top}, {{2,2, 0,0, 1,1}, top}, {{0,0, 2,2, 1,1}, top},
        {{1,1, 2,2, 0,0}, top}, {{1,1, 0,0, 2,2}, top}, {{0,0, 1,1, 2,2}, top},
        {{1,1, 1,1, 1,1}, top}
    };
    for (int n = 0; n < 32; n++) {
        double X[16], Y[16], Z[16];
        reg_eval_integral (X, Q[0][T[n].o[0]], Q[0][T[n].o[1]], gs[0]);
        reg_eval_integral (Y, Q[1][T[n].o[2]], Q[1][T[n].o[3]], gs[1]);
        reg_eval_integral (Z, Q[2][T[n].o[4]], Q[2][T[n].o[5]], gs[2]);
        reg_vmatrix (&V[4096 * n], X, Y, Z, (double) T[n].w);
    }
}

/* The 64 knots of region r, z outer then y then x.  Follows find_knots_3,
   src/register/bspline_regularize_analytic.cxx:129-165. */
static void
reg_find_knots (int64_t knots[64], int64_t r, const int64_t rdims[3], const int64_t cdims[3])
{
    int64_t t0 = r % rdims[0], t1 = (r / rdims[0]) % rdims[1], t2 = r / (rdims[0] * rdims[1]);
    int idx = 0;
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++)
                knots[idx++] = cdims[0] * cdims[1] * (t2 + k) + cdims[0] * (t1 + j) + (t0 + i);
}

/* S = sum over the symmetric V of p'Vp per component, sets += 2 V p.
   Follows region_smoothness_omp, :363-399. */
static double
reg_region_smoothness (double* sets, const float* coeff, const double* V, const int64_t knots[64])
{
    double S = 0.0;
    for (int j = 0; j < 64; j++) {
        double X = 0.0, Y = 0.0, Z = 0.0;
        for (int i = 0; i < 64; i++) {
            X += coeff[3*knots[i]+0] * V[64*j + i];
            Y += coeff[3*knots[i]+1] * V[64*j + i];
            Z += coeff[3*knots[i]+2] * V[64*j + i];
        }
        S += X * coeff[3*knots[j]+0];
        S += Y * coeff[3*knots[j]+1];
        S += Z * coeff[3*knots[j]+2];
        sets[3*j+0] += 2 * X;
        sets[3*j+1] += 2 * Y;
        sets[3*j+2] += 2 * Z;
    }
    return S;
}

/* The linear-elastic terms V9..V20 with their component couplings and the
   reference's gradient expression.  Follows region_smoothness_elastic_omp,
   :501-601. */
static double
reg_region_elastic (double* sets, const float* coeff, const double* V0, const int64_t knots[64])
{
    const double* V9 = V0 + 9*4096;  const double* V10 = V0 + 10*4096;
    const double* V11 = V0 + 11*4096; const double* V12 = V0 + 12*4096;
    const double* V13 = V0 + 13*4096; const double* V14 = V0 + 14*4096;
    const double* V15 = V0 + 15*4096; const double* V16 = V0 + 16*4096;
    const double* V17 = V0 + 17*4096; const double* V18 = V0 + 18*4096;
    const double* V19 = V0 + 19*4096; const double* V20 = V0 + 20*4096;
    double S = 0.0;
    for (int j = 0; j < 64; j++) {
        double X1 = 0, Y1 = 0, X2 = 0, Y2 = 0, X3 = 0, Z3 = 0, X4 = 0, Z4 = 0, Y5 = 0, Z5 = 0;
        double Y6 = 0, Z6 = 0, X7 = 0, Y7 = 0, Z7 = 0, X8 = 0, Y8 = 0, Z8 = 0, X9 = 0, Y9 = 0, Z9 = 0;
        for (int i = 0; i < 64; i++) {
            const double cx = coeff[3*knots[i]+0], cy = coeff[3*knots[i]+1], cz = coeff[3*knots[i]+2];
            const int e = 64*j + i;
            X1 += cx * V9[e];  X2 += cx * V10[e]; X3 += cx * V11[e]; X4 += cx * V12[e];
            Y5 += cy * V13[e]; Y6 += cy * V14[e]; Y1 += cy * V9[e];  Y2 += cy * V10[e];
            Z3 += cz * V11[e]; Z4 += cz * V12[e]; Z5 += cz * V13[e]; Z6 += cz * V14[e];
            X7 += cx * V15[e]; Y8 += cy * V16[e]; Z9 += cz * V17[e]; Y7 += cy * V18[e];
            Z7 += cz * V18[e]; X8 += cx * V19[e]; Z8 += cz * V19[e]; X9 += cx * V20[e];
            Y9 += cy * V20[e];
        }
        const double px = coeff[3*knots[j]+0], py = coeff[3*knots[j]+1], pz = coeff[3*knots[j]+2];
        S += X1 * py; S += X2 * py; S += X3 * pz; S += X4 * pz; S += Y5 * pz; S += Y6 * pz;
        S += X7 * px; S += Y8 * py; S += Z9 * pz; S += Y7 * py; S += Z7 * pz; S += X8 * px;
        S += Z8 * pz; S += X9 * px; S += Y9 * py;
        sets[3*j+0] += Y1 + Y2 + Z3 + Z4 + 2 * X7 + 2 * X8 + 2 * X9;
        sets[3*j+1] += X1 + X2 + Z5 + Z6 + 2 * Y8 + 2 * Y7 + 2 * Y9;
        sets[3*j+2] += X3 + X4 + Y5 + Y6 + 2 * Z9 + 2 * Z7 + 2 * Z8;
    }
    return S;
}

/* Regularization score and gradient, flavor c: per region the 21
   symmetric terms and the elastic terms, slots placed per knot and summed
   in slot order.  Follows Bspline_regularize::compute_score_analytic_omp,
   src/register/bspline_regularize_analytic.cxx:943-1007 (with
   reg_sort_sets :87-105 and reg_update_grad :110-127); the region sums are
   added in region order here (the reference's OpenMP reduction has no fixed
   order).  grid_spac = vox_per_rgn * |img_spacing| in float
   (src/base/bspline_header.cxx:50). */
int
orc_reg_analytic (const orc_geom* g, const orc_reg_parms* rp, const float* coeff, int nthreads,
    double* score, double* grad)
{
    int64_t rdims[3], cdims[3];
    const int64_t nk = orc_header (g, rdims, cdims);
    const int64_t nr = rdims[0] * rdims[1] * rdims[2];
    double gs[3];
    for (int d = 0; d < 3; d++)
        gs[d] = (double) ((float) g->vox_per_rgn[d] * fabsf (g->img_spacing[d]));
    double* V = (double*) malloc (sizeof (double) * 32 * 4096);
    double* cond = (double*) calloc ((size_t) 3 * 64 * nk, sizeof (double));
    double* Sr = (double*) malloc (sizeof (double) * nr);
    if (!V || !cond || !Sr) { free (V); free (cond); free (Sr); return -1; }
    reg_analytic_init (V, gs, rp);
    set_threads (nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < nr; r++) {
        int64_t knots[64];
        double sets[3*64];
        memset (sets, 0, sizeof (sets));
        reg_find_knots (knots, r, rdims, cdims);
        double S = 0.0;
        for (int n = 0; n < 9; n++) S += reg_region_smoothness (sets, coeff, &V[4096*n], knots);
        S += reg_region_elastic (sets, coeff, V, knots);
        for (int n = 21; n < 32; n++) S += reg_region_smoothness (sets, coeff, &V[4096*n], knots);
        for (int s = 0; s < 64; s++)
            for (int d = 0; d < 3; d++) cond[3 * (64 * knots[s] + s) + d] = sets[3*s + d];
        Sr[r] = S;
    }
    double S = 0.0;
    for (int64_t r = 0; r < nr; r++) S += Sr[r];
    for (int64_t k = 0; k < nk; k++)
        for (int d = 0; d < 3; d++) {
            double a = 0.0;
            for (int s = 0; s < 64; s++) a += cond[3 * (64 * k + s) + d];
            grad[3*k + d] = a;
        }
    *score = S;
    free (V); free (cond); free (Sr);
    return 0;
}

/* The 32 V matrices, for tests. */
int
orc_reg_matrices (const orc_geom* g, const orc_reg_parms* rp, double* V)
{
    double gs[3];
    for (int d = 0; d < 3; d++)
        gs[d] = (double) ((float) g->vox_per_rgn[d] * fabsf (g->img_spacing[d]));
    reg_analytic_init (V, gs, rp);
    return 0;
}
