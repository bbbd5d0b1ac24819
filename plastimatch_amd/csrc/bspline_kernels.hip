// B-spline score/gradient kernels for gfx950 (MI355X).
//
// One workgroup owns one B-spline tile (vpr0 x vpr1 x vpr2 fixed voxels
// sharing the same 4x4x4 control knots).  Inside the tile:
//   1. the 64x3 tile coefficients and the per-axis basis LUTs go to LDS;
//   2. the displacement is contracted separably: T1 = sum_i bx*c, then
//      T2 = sum_j by*T1, so each voxel only does the last 4-term sum in z;
//   3. each voxel warps, samples the moving image trilinearly and forms
//      dc/dv, which is parked in LDS;
//   4. the condense into the tile's 64 knot slots is the transposed
//      separable contraction (x, then y, then z), written once per tile to
//      a tile-major buffer: no global atomics, a single writer per slot;
//   5. a reduce kernel sums, per knot, its <= 64 slots in FP64 in a fixed
//      order, so the gradient is deterministic run to run.
// The MI histogram pass instead evaluates the displacement in the
// reference's k-j-i order without FMA so its partial-volume weights are
// bit-identical to the CPU flavor's, and accumulates them as 2^-32
// fixed-point integers (order independent) with LDS atomics.
//
// The file is compiled with -ffp-contract=off: every a*b+c below rounds
// twice unless fmaf() is written, which is done only where the reference's
// rounding is not reproduced anyway (the separable contractions).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>

#include "bspline_kernels.h"

namespace plmhip {

// ---------------------------------------------------------------------------
// image loads (FP32 or BF16 storage, FP32 arithmetic)

__device__ __forceinline__ float ld (const float* __restrict__ p, int i) { return p[i]; }
__device__ __forceinline__ float ld (const uint16_t* __restrict__ p, int i)
{
    return __uint_as_float (((uint32_t) p[i]) << 16);
}

// ---------------------------------------------------------------------------
// per-voxel primitives

struct Sample {
    int mvf;            // linear index of the floor corner
    int mvr;            // linear index of the rounded voxel
    float l1[3];        // 1 - frac
    float l2[3];        // frac
};

// Correspondence, inside test and clamped floor/round/fractions.
// Follows bspline_find_correspondence_dcos
// (src/util/bspline_correspond.cxx:80-109), Volume::is_inside
// (src/base/volume.cxx:899-906) and li_clamp (src/base/interpolate.cxx:17-41).
__device__ __forceinline__ bool
correspond (const KGeom& g, const float fxyz[3], const float d[3], Sample& s)
{
    float t[3], mijk[3];
    t[0] = (fxyz[0] + d[0]) - g.morigin[0];
    t[1] = (fxyz[1] + d[1]) - g.morigin[1];
    t[2] = (fxyz[2] + d[2]) - g.morigin[2];
#pragma unroll
    for (int a = 0; a < 3; a++)
        mijk[a] = t[0] * g.mproj[3*a+0] + t[1] * g.mproj[3*a+1] + t[2] * g.mproj[3*a+2];
    bool inside = true;
#pragma unroll
    for (int a = 0; a < 3; a++)
        inside = inside && (mijk[a] > -0.5f) && (mijk[a] < (float) g.mdim[a] - 0.5f);
    if (!inside) return false;
    int f[3], r[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        float ma = mijk[a];
        int dmax = g.mdim[a] - 1;
        float fa2;
        if (ma < 0.f) {
            f[a] = 0; r[a] = 0; fa2 = 0.f;
        } else if (ma >= (float) dmax) {
            f[a] = dmax - 1; r[a] = dmax; fa2 = 1.f;
        } else {
            float fl = floorf (ma);
            f[a] = (int) fl;
            fa2 = ma - fl;                              // exact
            r[a] = f[a] + (fa2 >= 0.5f ? 1 : 0);        // == trunc(ma + 0.5) in double
        }
        s.l2[a] = fa2;
        s.l1[a] = 1.0f - fa2;
    }
    s.mvf = f[0] + g.mdim[0] * (f[1] + g.mdim[1] * f[2]);
    s.mvr = r[0] + g.mdim[0] * (r[1] + g.mdim[1] * r[2]);
    return true;
}

template <typename TI>
__device__ __forceinline__ void
corners (float a[8], const TI* __restrict__ m, const KGeom& g, int mvf)
{
    const int sx = g.mdim[0];
    const int sxy = g.mdim[0] * g.mdim[1];
    a[0] = ld (m, mvf);
    a[1] = ld (m, mvf + 1);
    a[2] = ld (m, mvf + sx);
    a[3] = ld (m, mvf + sx + 1);
    a[4] = ld (m, mvf + sxy);
    a[5] = ld (m, mvf + sxy + 1);
    a[6] = ld (m, mvf + sxy + sx);
    a[7] = ld (m, mvf + sxy + sx + 1);
}

// Trilinear value, products and sum in the reference's order.
// Follows li_value (src/base/interpolate.cxx:135-160).
__device__ __forceinline__ float
li_value (const Sample& s, const float a[8])
{
    const float* f1 = s.l1;
    const float* f2 = s.l2;
    float v0 = f1[0] * f1[1] * f1[2] * a[0];
    float v1 = f2[0] * f1[1] * f1[2] * a[1];
    float v2 = f1[0] * f2[1] * f1[2] * a[2];
    float v3 = f2[0] * f2[1] * f1[2] * a[3];
    float v4 = f1[0] * f1[1] * f2[2] * a[4];
    float v5 = f2[0] * f1[1] * f2[2] * a[5];
    float v6 = f1[0] * f2[1] * f2[2] * a[6];
    float v7 = f2[0] * f2[1] * f2[2] * a[7];
    return v0 + v1 + v2 + v3 + v4 + v5 + v6 + v7;
}

// Analytic derivative of the trilinear value times 1/spacing.
// Follows li_value_dx/dy/dz (src/base/interpolate.cxx:162-244).
__device__ __forceinline__ void
li_value_d (float out[3], const Sample& s, const float a[8], const float inv[3])
{
    const float* f1 = s.l1;
    const float* f2 = s.l2;
    const float rx = inv[0], ry = inv[1], rz = inv[2];
    out[0] = -rx * f1[1] * f1[2] * a[0] + rx * f1[1] * f1[2] * a[1]
        + -rx * f2[1] * f1[2] * a[2] + rx * f2[1] * f1[2] * a[3]
        + -rx * f1[1] * f2[2] * a[4] + rx * f1[1] * f2[2] * a[5]
        + -rx * f2[1] * f2[2] * a[6] + rx * f2[1] * f2[2] * a[7];
    out[1] = -f1[0] * ry * f1[2] * a[0] + -f2[0] * ry * f1[2] * a[1]
        + f1[0] * ry * f1[2] * a[2] + f2[0] * ry * f1[2] * a[3]
        + -f1[0] * ry * f2[2] * a[4] + -f2[0] * ry * f2[2] * a[5]
        + f1[0] * ry * f2[2] * a[6] + f2[0] * ry * f2[2] * a[7];
    out[2] = -f1[0] * f1[1] * rz * a[0] + -f2[0] * f1[1] * rz * a[1]
        + -f1[0] * f2[1] * rz * a[2] + -f2[0] * f2[1] * rz * a[3]
        + f1[0] * f1[1] * rz * a[4] + f2[0] * f1[1] * rz * a[5]
        + f1[0] * f2[1] * rz * a[6] + f2[0] * f2[1] * rz * a[7];
}

// Fixed-image position, origin + step*ijk in the reference's order.
// Follows POSITION_FROM_COORDS (src/base/volume_macros.h:82-96).
__device__ __forceinline__ void
position (float fxyz[3], const KGeom& g, int fi, int fj, int fk)
{
    const float x = (float) fi, y = (float) fj, z = (float) fk;
#pragma unroll
    for (int a = 0; a < 3; a++)
        fxyz[a] = g.img_origin[a] + x * g.fstep[3*a+0] + y * g.fstep[3*a+1] + z * g.fstep[3*a+2];
}

// ---------------------------------------------------------------------------
// tile bookkeeping

__device__ __forceinline__ int
xcd_tile (const KGeom& g, int b)
{
    // Workgroups are dealt to the 8 XCDs round-robin; give each XCD a
    // contiguous run of tiles so x-neighbours share that XCD's L2.
    return (b & 7) * g.tiles_per_xcd + (b >> 3);
}

__device__ __forceinline__ void
tile_coords (const KGeom& g, int tile, int p[3])
{
    p[0] = tile % g.rdims[0];
    int t = tile / g.rdims[0];
    p[1] = t % g.rdims[1];
    p[2] = t / g.rdims[1];
}

// LDS carve-up shared by the tile kernels (sizes in floats).
struct TileLds {
    float* bx; float* by; float* bz;   // 4*vpr each
    float* coef;                       // [k][j][i][c], 192
    float* t1;                         // [k][j][x][c] then S2 [z][j][i][c]
    float* t2;                         // [y][x][c][k] then S1 [z][y][i][c]
    float* dv;                         // [z][y][x][c]
    double* red;                       // 2 * waves
};

__device__ __forceinline__ TileLds
carve (float* sm, const KGeom& g)
{
    TileLds L;
    const int vx = g.vpr[0], vy = g.vpr[1], vz = g.vpr[2];
    L.bx = sm;
    L.by = L.bx + 4 * vx;
    L.bz = L.by + 4 * vy;
    L.coef = L.bz + 4 * vz;
    L.t1 = L.coef + 192;
    L.t2 = L.t1 + 48 * (vx > vz ? vx : vz);
    L.dv = L.t2 + 12 * vy * (vx > vz ? vx : vz);
    float* end = L.dv + 3 * vx * vy * vz;
    uintptr_t e = ((uintptr_t) end + 7) & ~(uintptr_t) 7;
    L.red = (double*) e;
    return L;
}

// Stage LUTs and the tile's 64 knot coefficients.
__device__ __forceinline__ void
load_tile (const TileLds& L, const KGeom& g, const int p[3],
    const float* __restrict__ coeff, const float* __restrict__ blut)
{
    const int vx = g.vpr[0], vy = g.vpr[1], vz = g.vpr[2];
    const int nl = 4 * (vx + vy + vz);
    for (int i = threadIdx.x; i < nl; i += blockDim.x) L.bx[i] = blut[i];
    for (int e = threadIdx.x; e < 192; e += blockDim.x) {
        int c = e % 3, m = e / 3;
        int i = m & 3, j = (m >> 2) & 3, k = m >> 4;
        int knot = (p[0] + i) + g.cdims[0] * ((p[1] + j) + g.cdims[1] * (p[2] + k));
        L.coef[e] = coeff[3 * knot + c];
    }
}

// T1[k][j][x][c] = sum_i bx[x][i] coef[k][j][i][c];
// T2[y][x][c][k] = sum_j by[y][j] T1[k][j][x][c].
__device__ __forceinline__ void
separable_displacement (const TileLds& L, const KGeom& g)
{
    const int vx = g.vpr[0], vy = g.vpr[1];
    const int n1 = 48 * vx;
    for (int e = threadIdx.x; e < n1; e += blockDim.x) {
        int c = e % 3, x = (e / 3) % vx, kj = e / (3 * vx);
        const float* cf = &L.coef[(kj * 4) * 3 + c];
        const float* b = &L.bx[4 * x];
        float v = b[0] * cf[0];
        v = fmaf (b[1], cf[3], v);
        v = fmaf (b[2], cf[6], v);
        v = fmaf (b[3], cf[9], v);
        L.t1[e] = v;   // index ((kj*vx + x)*3 + c)
    }
    __syncthreads ();
    const int n2 = 12 * vx * vy;
    for (int e = threadIdx.x; e < n2; e += blockDim.x) {
        int k = e & 3, c = (e >> 2) % 3, xy = e / 12;
        int x = xy % vx, y = xy / vx;
        const float* b = &L.by[4 * y];
        float v = 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++)
            v = fmaf (b[j], L.t1[(((k * 4 + j) * vx) + x) * 3 + c], v);
        L.t2[e] = v;   // index ((y*vx + x)*3 + c)*4 + k
    }
    __syncthreads ();
}

__device__ __forceinline__ void
voxel_displacement (float d[3], const TileLds& L, int qx, int qy, int qz, int vx)
{
    const float4* t = reinterpret_cast<const float4*> (&L.t2[((qy * vx + qx) * 3) * 4]);
    const float* b = &L.bz[4 * qz];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float4 v = t[c];
        float s = b[0] * v.x;
        s = fmaf (b[1], v.y, s);
        s = fmaf (b[2], v.z, s);
        s = fmaf (b[3], v.w, s);
        d[c] = s;
    }
}

// Condense the parked dc/dv of the tile into its 64 knot slots and write
// them to cond[tile][slot][c].  S1 over x, S2 over y, S3 over z.
__device__ __forceinline__ void
condense_tile (const TileLds& L, const KGeom& g, int tile, float* __restrict__ cond)
{
    const int vx = g.vpr[0], vy = g.vpr[1], vz = g.vpr[2];
    float* s1 = L.t2;   // [z][y][i][c]
    float* s2 = L.t1;   // [z][j][i][c]
    const int n1 = 3 * vy * vz;
    for (int e = threadIdx.x; e < n1; e += blockDim.x) {
        int c = e % 3, zy = e / 3;
        const float* dv = &L.dv[zy * vx * 3 + c];
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        for (int x = 0; x < vx; x++) {
            float v = dv[3 * x];
            const float* b = &L.bx[4 * x];
            a0 = fmaf (b[0], v, a0);
            a1 = fmaf (b[1], v, a1);
            a2 = fmaf (b[2], v, a2);
            a3 = fmaf (b[3], v, a3);
        }
        float* o = &s1[zy * 12 + c];
        o[0] = a0; o[3] = a1; o[6] = a2; o[9] = a3;
    }
    __syncthreads ();
    const int n2 = 12 * vz;
    for (int e = threadIdx.x; e < n2; e += blockDim.x) {
        int ic = e % 12, z = e / 12;
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        for (int y = 0; y < vy; y++) {
            float v = s1[(z * vy + y) * 12 + ic];
            const float* b = &L.by[4 * y];
            a0 = fmaf (b[0], v, a0);
            a1 = fmaf (b[1], v, a1);
            a2 = fmaf (b[2], v, a2);
            a3 = fmaf (b[3], v, a3);
        }
        float* o = &s2[z * 48 + ic];
        o[0] = a0; o[12] = a1; o[24] = a2; o[36] = a3;
    }
    __syncthreads ();
    if (threadIdx.x < 48) {
        int jic = threadIdx.x;          // (j*4 + i)*3 + c
        float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
        for (int z = 0; z < vz; z++) {
            float v = s2[z * 48 + jic];
            const float* b = &L.bz[4 * z];
            a0 = fmaf (b[0], v, a0);
            a1 = fmaf (b[1], v, a1);
            a2 = fmaf (b[2], v, a2);
            a3 = fmaf (b[3], v, a3);
        }
        float* o = &cond[(size_t) tile * 192 + jic];
        o[0] = a0; o[48] = a1; o[96] = a2; o[144] = a3;   // slot = k*16 + j*4 + i
    }
}

// Block sum of a double and an int in a fixed order (deterministic).
__device__ __forceinline__ void
block_sum (const TileLds& L, double& s, int& n)
{
    double nd = (double) n;
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down (s, off);
        nd += __shfl_down (nd, off);
    }
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads ();
    if ((threadIdx.x & 63) == 0) {
        L.red[2 * w] = s;
        L.red[2 * w + 1] = nd;
    }
    __syncthreads ();
    if (threadIdx.x == 0) {
        double a = 0, b = 0;
        for (int i = 0; i < nw; i++) { a += L.red[2 * i]; b += L.red[2 * i + 1]; }
        s = a;
        n = (int) b;
    }
}

// ---------------------------------------------------------------------------
// MSE tile kernel (flavor g: MODE 0, flavor p: MODE 1)

template <int MODE, typename TI>
__global__ __launch_bounds__ (TILE_THREADS)
void k_mse_tiles (KGeom g, const TI* __restrict__ fimg, const TI* __restrict__ mimg,
    const TI* __restrict__ mgrad, const float* __restrict__ coeff,
    const float* __restrict__ blut, float* __restrict__ cond,
    double* __restrict__ tile_score, int* __restrict__ tile_count)
{
    extern __shared__ float sm[];
    const int tile = xcd_tile (g, blockIdx.x);
    if (tile >= g.num_tiles) return;
    const KGeom& G = g;
    TileLds L = carve (sm, G);
    int p[3];
    tile_coords (G, tile, p);
    load_tile (L, G, p, coeff, blut);
    __syncthreads ();
    separable_displacement (L, G);

    const int vx = G.vpr[0], vy = G.vpr[1], vz = G.vpr[2];
    const int nvox = vx * vy * vz;
    double score = 0.0;
    int count = 0;
    for (int v = threadIdx.x; v < nvox; v += blockDim.x) {
        int qx = v % vx, t = v / vx, qy = t % vy, qz = t / vy;
        int fi = G.roi_off[0] + vx * p[0] + qx;
        int fj = G.roi_off[1] + vy * p[1] + qy;
        int fk = G.roi_off[2] + vz * p[2] + qz;
        float dc0 = 0.f, dc1 = 0.f, dc2 = 0.f;
        if (fi < G.roi_end[0] && fj < G.roi_end[1] && fk < G.roi_end[2]) {
            float fxyz[3], d[3];
            position (fxyz, G, fi, fj, fk);
            voxel_displacement (d, L, qx, qy, qz, vx);
            Sample s;
            if (correspond (G, fxyz, d, s)) {
                float a[8];
                corners (a, mimg, G, s.mvf);
                float m_val = li_value (s, a);
                int fidx = fi + G.fdim[0] * (fj + G.fdim[1] * fk);
                float diff = m_val - ld (fimg, fidx);
                score += (double) (diff * diff);
                count++;
                if (MODE == 0) {
                    dc0 = diff * ld (mgrad, 3 * s.mvr + 0);
                    dc1 = diff * ld (mgrad, 3 * s.mvr + 1);
                    dc2 = diff * ld (mgrad, 3 * s.mvr + 2);
                } else {
                    float md[3];
                    li_value_d (md, s, a, G.inv_msp);
                    dc0 = diff * md[0];
                    dc1 = diff * md[1];
                    dc2 = diff * md[2];
                }
            }
        }
        L.dv[3 * v + 0] = dc0;
        L.dv[3 * v + 1] = dc1;
        L.dv[3 * v + 2] = dc2;
    }
    __syncthreads ();
    condense_tile (L, G, tile, cond);
    block_sum (L, score, count);
    if (threadIdx.x == 0) {
        tile_score[tile] = score;
        tile_count[tile] = count;
    }
}

// Sum the per-tile partial scores in a fixed order; score = raw/N,
// gradient scale 2/N, or FLT_MAX and 0 when N < 1
// (src/register/bspline_mse.cxx:43-73).
__global__ __launch_bounds__ (1024)
void k_mse_finalize (const double* __restrict__ tile_score,
    const int* __restrict__ tile_count, int num_tiles,
    double* __restrict__ out_score, int64_t* __restrict__ out_nvox,
    double* __restrict__ out_gscale)
{
    __shared__ double s_s[1024];
    __shared__ long long s_n[1024];
    double s = 0.0;
    long long n = 0;
    for (int i = threadIdx.x; i < num_tiles; i += blockDim.x) {
        s += tile_score[i];
        n += tile_count[i];
    }
    s_s[threadIdx.x] = s;
    s_n[threadIdx.x] = n;
    __syncthreads ();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if ((int) threadIdx.x < h) {
            s_s[threadIdx.x] += s_s[threadIdx.x + h];
            s_n[threadIdx.x] += s_n[threadIdx.x + h];
        }
        __syncthreads ();
    }
    if (threadIdx.x == 0) {
        long long N = s_n[0];
        if (N < 1) {
            *out_score = (double) FLT_MAX;
            *out_gscale = 0.0;
        } else {
            *out_score = s_s[0] / (double) N;
            *out_gscale = 2.0 / (double) N;
        }
        *out_nvox = N;
    }
}

// Per knot, sum its <= 64 tile slots in slot order in FP64 and scale.
// Same result as bspline_condense_smetric_grad
// (src/register/bspline.cxx:400-413) followed by the normalization.
__global__ __launch_bounds__ (256)
void k_grad_reduce (KGeom g, const float* __restrict__ cond,
    const double* __restrict__ gscale, double* __restrict__ grad)
{
    const int knot = blockIdx.x * blockDim.x + threadIdx.x;
    if (knot >= g.num_knots) return;
    const int kx = knot % g.cdims[0];
    const int kt = knot / g.cdims[0];
    const int ky = kt % g.cdims[1];
    const int kz = kt / g.cdims[1];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int k = 0; k < 4; k++) {
        int tz = kz - k;
        if (tz < 0 || tz >= g.rdims[2]) continue;
        for (int j = 0; j < 4; j++) {
            int ty = ky - j;
            if (ty < 0 || ty >= g.rdims[1]) continue;
            for (int i = 0; i < 4; i++) {
                int tx = kx - i;
                if (tx < 0 || tx >= g.rdims[0]) continue;
                int tile = tx + g.rdims[0] * (ty + g.rdims[1] * tz);
                const float* c = &cond[(size_t) tile * 192 + ((k * 4 + j) * 4 + i) * 3];
                a0 += (double) c[0];
                a1 += (double) c[1];
                a2 += (double) c[2];
            }
        }
    }
    const double s = *gscale;
    grad[3 * knot + 0] = a0 * s;
    grad[3 * knot + 1] = a1 * s;
    grad[3 * knot + 2] = a2 * s;
}

// ---------------------------------------------------------------------------
// Mattes MI

__device__ __forceinline__ int
hist_bin (float v, float offset, float delta, int bins)
{
    int b = (int) floorf ((v - offset) / delta);
    // In-range intensities always give 0..bins-1; the clamp only keeps a
    // corrupted input from writing outside the histogram.
    return b < 0 ? 0 : (b >= bins ? bins - 1 : b);
}

// Displacement in the reference's k-j-i order, bx[i]*(by[j]*bz[k]) weights,
// no FMA: bit-identical to bspline_interp_pix_c
// (src/base/bspline_interpolate.cxx:78-110).
__device__ __forceinline__ void
exact_displacement (float out[3], const TileLds& L, int qx, int qy, int qz)
{
    const float* bx = &L.bx[4 * qx];
    const float* by = &L.by[4 * qy];
    const float* bz = &L.bz[4 * qz];
    float o0 = 0.f, o1 = 0.f, o2 = 0.f;
    int m = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float C = bz[k];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float B = by[j] * C;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                float A = bx[i] * B;
                const float* cf = &L.coef[3 * m];
                o0 += A * cf[0];
                o1 += A * cf[1];
                o2 += A * cf[2];
                m++;
            }
        }
    }
    out[0] = o0; out[1] = o1; out[2] = o2;
}

// Fixed-point PV weights (units of 2^-32): q = floor(w*2^32 + 0.5), the
// remainder to the first largest weight, so each voxel adds exactly 2^32.
__device__ __forceinline__ void
pv_quantize (const float w[8], long long q[8])
{
    long long sum = 0;
    int amax = 0;
#pragma unroll
    for (int n = 0; n < 8; n++) {
        q[n] = (long long) ((double) w[n] * 4294967296.0 + 0.5);
        sum += q[n];
        if (w[n] > w[amax]) amax = n;
    }
    const long long r = (1ll << 32) - sum;
#pragma unroll
    for (int n = 0; n < 8; n++) if (n == amax) q[n] += r;
}

// Pass 1: PV-8 joint histogram.  Follows the pass-1 loop of
// bspline_score_g_mi (src/register/bspline_mi.cxx:1567-1615) with
// Joint_histogram::add_pvi_8 (src/register/joint_histogram.cxx:423-487).
// Persistent: each workgroup walks tiles with a grid stride and keeps a
// private int64 histogram in LDS, flushed once with integer atomics.
template <typename TI>
__global__ __launch_bounds__ (TILE_THREADS)
void k_mi_hist (KGeom g, HGeom h, const TI* __restrict__ fimg,
    const TI* __restrict__ mimg, const float* __restrict__ coeff,
    const float* __restrict__ blut, unsigned long long* __restrict__ jhist,
    unsigned long long* __restrict__ fhist)
{
    extern __shared__ float sm[];
    const KGeom& G = g;
    const int vx = G.vpr[0], vy = G.vpr[1], vz = G.vpr[2];
    float* s_bx = sm;
    float* s_coef = s_bx + 4 * (vx + vy + vz);
    uintptr_t hp = ((uintptr_t) (s_coef + 192) + 7) & ~(uintptr_t) 7;
    unsigned long long* s_j = (unsigned long long*) hp;
    unsigned long long* s_f = s_j + h.fbins * h.mbins;
    TileLds L;
    L.bx = s_bx; L.by = s_bx + 4 * vx; L.bz = L.by + 4 * vy; L.coef = s_coef;

    const int nbins = h.fbins * h.mbins + h.fbins;
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) s_j[i] = 0ull;
    const int nl = 4 * (vx + vy + vz);
    for (int i = threadIdx.x; i < nl; i += blockDim.x) s_bx[i] = blut[i];

    const int nvox = vx * vy * vz;
    for (int tile = blockIdx.x; tile < G.num_tiles; tile += gridDim.x) {
        int p[3];
        tile_coords (G, tile, p);
        __syncthreads ();
        for (int e = threadIdx.x; e < 192; e += blockDim.x) {
            int c = e % 3, m = e / 3;
            int i = m & 3, j = (m >> 2) & 3, k = m >> 4;
            int knot = (p[0] + i) + G.cdims[0] * ((p[1] + j) + G.cdims[1] * (p[2] + k));
            s_coef[e] = coeff[3 * knot + c];
        }
        __syncthreads ();
        for (int v = threadIdx.x; v < nvox; v += blockDim.x) {
            int qx = v % vx, t = v / vx, qy = t % vy, qz = t / vy;
            int fi = vx * p[0] + qx, fj = vy * p[1] + qy, fk = vz * p[2] + qz;
            if (fi >= G.fdim[0] || fj >= G.fdim[1] || fk >= G.fdim[2]) continue;
            float fxyz[3], d[3];
            position (fxyz, G, fi, fj, fk);
            exact_displacement (d, L, qx, qy, qz);
            Sample s;
            if (!correspond (G, fxyz, d, s)) continue;
            float a[8];
            corners (a, mimg, G, s.mvf);
            const float* l1 = s.l1;
            const float* l2 = s.l2;
            float w[8];
            w[0] = l1[0] * l1[1] * l1[2];
            w[1] = l2[0] * l1[1] * l1[2];
            w[2] = l1[0] * l2[1] * l1[2];
            w[3] = l2[0] * l2[1] * l1[2];
            w[4] = l1[0] * l1[1] * l2[2];
            w[5] = l2[0] * l1[1] * l2[2];
            w[6] = l1[0] * l2[1] * l2[2];
            w[7] = l2[0] * l2[1] * l2[2];
            long long q[8];
            pv_quantize (w, q);
            int fidx = fi + G.fdim[0] * (fj + G.fdim[1] * fk);
            int fb = hist_bin (ld (fimg, fidx), h.f_offset, h.f_delta, h.fbins);
            atomicAdd (&s_f[fb], 1ull);
            unsigned long long* row = &s_j[fb * h.mbins];
#pragma unroll
            for (int n = 0; n < 8; n++) {
                int mb = hist_bin (a[n], h.m_offset, h.m_delta, h.mbins);
                atomicAdd (&row[mb], (unsigned long long) q[n]);
            }
        }
    }
    __syncthreads ();
    for (int i = threadIdx.x; i < h.fbins * h.mbins; i += blockDim.x)
        if (s_j[i]) atomicAdd (&jhist[i], s_j[i]);
    for (int i = threadIdx.x; i < h.fbins; i += blockDim.x)
        if (s_f[i]) atomicAdd (&fhist[i], s_f[i]);
}

// Score from the integer histograms, and the per-bin dS/dP table for
// pass 2.  Follows mi_hist_score_omp (src/register/bspline_mi.cxx:285-310)
// and the dS_dP expression of bspline_mi_pvi_8_dc_dv_dcos
// (src/register/bspline_mi.txx:96-101).  One workgroup of 1024 threads.
__global__ __launch_bounds__ (1024)
void k_mi_score (HGeom h, const unsigned long long* __restrict__ jhist,
    const unsigned long long* __restrict__ fhist, float* __restrict__ table,
    double* __restrict__ out_score, int64_t* __restrict__ out_nvox,
    float* __restrict__ out_nvox_f, int64_t* __restrict__ hist_out)
{
    __shared__ double s_m[MAX_BINS_1D];
    __shared__ double s_f[MAX_BINS_1D];
    __shared__ double s_red[1024];
    __shared__ long long s_n;
    const int fb = h.fbins, mb = h.mbins, nj = fb * mb;
    const double unit = 1.0 / 4294967296.0;
    if (threadIdx.x == 0) {
        long long n = 0;
        for (int i = 0; i < fb; i++) n += (long long) fhist[i];
        s_n = n;
    }
    for (int b = threadIdx.x; b < mb; b += blockDim.x) {
        unsigned long long m = 0;
        for (int a = 0; a < fb; a++) m += jhist[a * mb + b];
        s_m[b] = (double) m * unit;
    }
    for (int a = threadIdx.x; a < fb; a += blockDim.x) s_f[a] = (double) fhist[a];
    __syncthreads ();
    const long long N = s_n;
    const double fnv = (double) N;
    const double thresh = 0.001 / (double) nj;
    double acc = 0.0;
    for (int jb = threadIdx.x; jb < nj; jb += blockDim.x) {
        double j = (double) jhist[jb] * unit;
        if (j > thresh)
            acc -= j * (double) logf ((float) (fnv * j / (s_m[jb % mb] * s_f[jb / mb])));
    }
    s_red[threadIdx.x] = acc;
    __syncthreads ();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int) threadIdx.x < s) s_red[threadIdx.x] += s_red[threadIdx.x + s];
        __syncthreads ();
    }
    const double score = s_red[0] / fnv;
    const float score_f = (float) score;
    const float nvf = (float) N;
    for (int jb = threadIdx.x; jb < nj; jb += blockDim.x) {
        double j = (double) jhist[jb] * unit;
        float ds = 0.f;
        if (j > 0.0001)
            ds = logf ((float) (((double) nvf * j) / (s_f[jb / mb] * s_m[jb % mb]))) - score_f;
        table[jb] = ds;
        if (hist_out) hist_out[jb] = (int64_t) jhist[jb];
    }
    if (hist_out)
        for (int a = threadIdx.x; a < fb; a += blockDim.x) hist_out[nj + a] = (int64_t) fhist[a];
    if (threadIdx.x == 0) {
        *out_score = score;
        *out_nvox = N;
        *out_nvox_f = nvf;
    }
}

// Pass 2: PV-8 derivative per voxel, condensed per tile.  Follows the
// pass-2 loop of bspline_score_g_mi (src/register/bspline_mi.cxx:1627-1716)
// and bspline_mi_pvi_8_dc_dv_dcos (src/register/bspline_mi.txx:9-120).
template <typename TI>
__global__ __launch_bounds__ (TILE_THREADS)
void k_mi_tiles (KGeom g, HGeom h, const TI* __restrict__ fimg,
    const TI* __restrict__ mimg, const float* __restrict__ coeff,
    const float* __restrict__ blut, const float* __restrict__ table,
    const float* __restrict__ nvox_f, float* __restrict__ cond)
{
    extern __shared__ float sm[];
    const int tile = xcd_tile (g, blockIdx.x);
    if (tile >= g.num_tiles) return;
    const KGeom& G = g;
    TileLds L = carve (sm, G);
    float* s_tab = (float*) (L.red + 16);
    const int nj = h.fbins * h.mbins;
    for (int i = threadIdx.x; i < nj; i += blockDim.x) s_tab[i] = table[i];
    int p[3];
    tile_coords (G, tile, p);
    load_tile (L, G, p, coeff, blut);
    __syncthreads ();
    separable_displacement (L, G);

    const float N = *nvox_f;
    const int vx = G.vpr[0], vy = G.vpr[1], vz = G.vpr[2];
    const int nvox = vx * vy * vz;
    const float* P = G.mproj;
    for (int v = threadIdx.x; v < nvox; v += blockDim.x) {
        int qx = v % vx, t = v / vx, qy = t % vy, qz = t / vy;
        int fi = G.roi_off[0] + vx * p[0] + qx;
        int fj = G.roi_off[1] + vy * p[1] + qy;
        int fk = G.roi_off[2] + vz * p[2] + qz;
        float dc0 = 0.f, dc1 = 0.f, dc2 = 0.f;
        if (fi < G.roi_end[0] && fj < G.roi_end[1] && fk < G.roi_end[2]) {
            float fxyz[3], d[3];
            position (fxyz, G, fi, fj, fk);
            voxel_displacement (d, L, qx, qy, qz, vx);
            Sample s;
            if (correspond (G, fxyz, d, s)) {
                float a[8];
                corners (a, mimg, G, s.mvf);
                const float* l1 = s.l1;
                const float* l2 = s.l2;
                float dw[24];
                dw[0] = -l1[1] * l1[2];  dw[1] = -l1[0] * l1[2];  dw[2] = -(l1[0] * l1[1]);
                dw[3] = l1[1] * l1[2];   dw[4] = -l2[0] * l1[2];  dw[5] = -(l2[0] * l1[1]);
                dw[6] = -l2[1] * l1[2];  dw[7] = l1[0] * l1[2];   dw[8] = -(l1[0] * l2[1]);
                dw[9] = l2[1] * l1[2];   dw[10] = l2[0] * l1[2];  dw[11] = -(l2[0] * l2[1]);
                dw[12] = -l1[1] * l2[2]; dw[13] = -l1[0] * l2[2]; dw[14] = l1[0] * l1[1];
                dw[15] = l1[1] * l2[2];  dw[16] = -l2[0] * l2[2]; dw[17] = l2[0] * l1[1];
                dw[18] = -l2[1] * l2[2]; dw[19] = l1[0] * l2[2];  dw[20] = l1[0] * l2[1];
                dw[21] = l2[1] * l2[2];  dw[22] = l2[0] * l2[2];  dw[23] = l2[0] * l2[1];
                int fidx = fi + G.fdim[0] * (fj + G.fdim[1] * fk);
                int fb = hist_bin (ld (fimg, fidx), h.f_offset, h.f_delta, h.fbins);
                const float* row = &s_tab[fb * h.mbins];
#pragma unroll
                for (int n = 0; n < 8; n++) {
                    int mb = hist_bin (a[n], h.m_offset, h.m_delta, h.mbins);
                    float ds = row[mb];
                    dc0 -= dw[3 * n + 0] * ds;
                    dc1 -= dw[3 * n + 1] * ds;
                    dc2 -= dw[3 * n + 2] * ds;
                }
                dc0 = dc0 / N;
                dc1 = dc1 / N;
                dc2 = dc2 / N;
                // projected in place, component by component, as the
                // reference does (src/register/bspline_mi.txx:107-109)
                dc0 = dc0 * P[0] + dc1 * P[1] + dc2 * P[2];
                dc1 = dc0 * P[3] + dc1 * P[4] + dc2 * P[5];
                dc2 = dc0 * P[6] + dc1 * P[7] + dc2 * P[8];
            }
        }
        L.dv[3 * v + 0] = dc0;
        L.dv[3 * v + 1] = dc1;
        L.dv[3 * v + 2] = dc2;
    }
    __syncthreads ();
    condense_tile (L, G, tile, cond);
}

// Central-difference gradient with edge replicate, rotated by inv(dc),
// float3 interleaved.  The difference is formed in float, divided by 2 and
// by the spacing in double and rounded to float, then accumulated x, y, z as
// volume_calc_grad does (src/base/volume_grad.cxx:9-76).
__global__ __launch_bounds__ (256)
void k_volume_gradient (const float* __restrict__ img, float* __restrict__ out,
    int nx, int ny, int nz, float sx, float sy, float sz, Mat3 inv)
{
    const int64_t n = (int64_t) nx * ny * nz;
    for (int64_t v = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; v < n;
         v += (int64_t) gridDim.x * blockDim.x) {
        const int i = (int) (v % nx);
        const int64_t t = v / nx;
        const int j = (int) (t % ny);
        const int k = (int) (t / ny);
        const int ip = i == 0 ? 0 : i - 1, in = i == nx - 1 ? nx - 1 : i + 1;
        const int jp = j == 0 ? 0 : j - 1, jn = j == ny - 1 ? ny - 1 : j + 1;
        const int kp = k == 0 ? 0 : k - 1, kn = k == nz - 1 ? nz - 1 : k + 1;
        const int64_t row = (int64_t) nx * (j + (int64_t) ny * k);
        const int64_t sxy = (int64_t) nx * ny;
        float g0 = 0.f, g1 = 0.f, g2 = 0.f, diff;
        diff = (float) ((double) (img[row + in] - img[row + ip]) / 2.0 / (double) sx);
        g0 += diff * inv.m[0]; g1 += diff * inv.m[3]; g2 += diff * inv.m[6];
        diff = (float) ((double) (img[v + (int64_t) (jn - j) * nx] - img[v + (int64_t) (jp - j) * nx])
            / 2.0 / (double) sy);
        g0 += diff * inv.m[1]; g1 += diff * inv.m[4]; g2 += diff * inv.m[7];
        diff = (float) ((double) (img[v + (int64_t) (kn - k) * sxy] - img[v + (int64_t) (kp - k) * sxy])
            / 2.0 / (double) sz);
        g0 += diff * inv.m[2]; g1 += diff * inv.m[5]; g2 += diff * inv.m[8];
        out[3 * v + 0] = g0;
        out[3 * v + 1] = g1;
        out[3 * v + 2] = g2;
    }
}

__global__ void k_fill_one (double* p) { *p = 1.0; }

__global__ void
k_zero_u64 (unsigned long long* __restrict__ p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// float32 -> bfloat16, round to nearest even (NaN kept quiet).
__global__ void
k_to_bf16 (const float* __restrict__ src, uint16_t* __restrict__ dst, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        uint32_t u = __float_as_uint (src[i]);
        if ((u & 0x7fffffffu) > 0x7f800000u) {
            dst[i] = (uint16_t) ((u >> 16) | 0x40u);
        } else {
            uint32_t r = 0x7fffu + ((u >> 16) & 1u);
            dst[i] = (uint16_t) ((u + r) >> 16);
        }
    }
}

// ---------------------------------------------------------------------------
// launchers

size_t
tile_lds_bytes (const KGeom& g)
{
    const size_t vx = g.vpr[0], vy = g.vpr[1], vz = g.vpr[2];
    const size_t mx = vx > vz ? vx : vz;
    size_t f = 4 * (vx + vy + vz) + 192 + 48 * mx + 12 * vy * mx + 3 * vx * vy * vz;
    size_t bytes = ((f * 4 + 7) & ~(size_t) 7) + 16 * sizeof (double);
    return bytes;
}

// Records ev[i] on st when stage timing is on (ev != nullptr).
static inline hipError_t
mark (hipEvent_t* ev, int i, hipStream_t st)
{
    return ev ? hipEventRecord (ev[i], st) : hipSuccess;
}

#define LAUNCH_CHECK()                                                  \
    do {                                                                \
        hipError_t e_ = hipGetLastError ();                             \
        if (e_ != hipSuccess) return e_;                                \
    } while (0)

hipError_t
launch_mse (const KGeom& g, int mode, int storage, const void* f, const void* m,
    const void* mg, const float* coeff, const float* blut, float* cond,
    double* tile_score, int* tile_count, double* score, int64_t* nvox,
    double* gscale, double* grad, hipStream_t st, hipEvent_t* ev)
{
    dim3 grid (8 * g.tiles_per_xcd), block (TILE_THREADS);
    size_t lds = tile_lds_bytes (g);
    hipError_t e = mark (ev, 0, st);
    if (e != hipSuccess) return e;
    if (storage == 0) {
        const float *F = (const float*) f, *M = (const float*) m, *MG = (const float*) mg;
        if (mode == 0)
            hipLaunchKernelGGL ((k_mse_tiles<0, float>), grid, block, lds, st, g, F, M, MG,
                coeff, blut, cond, tile_score, tile_count);
        else
            hipLaunchKernelGGL ((k_mse_tiles<1, float>), grid, block, lds, st, g, F, M, MG,
                coeff, blut, cond, tile_score, tile_count);
    } else {
        const uint16_t *F = (const uint16_t*) f, *M = (const uint16_t*) m,
            *MG = (const uint16_t*) mg;
        if (mode == 0)
            hipLaunchKernelGGL ((k_mse_tiles<0, uint16_t>), grid, block, lds, st, g, F, M, MG,
                coeff, blut, cond, tile_score, tile_count);
        else
            hipLaunchKernelGGL ((k_mse_tiles<1, uint16_t>), grid, block, lds, st, g, F, M, MG,
                coeff, blut, cond, tile_score, tile_count);
    }
    LAUNCH_CHECK ();
    if ((e = mark (ev, 1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL (k_mse_finalize, dim3 (1), dim3 (1024), 0, st, tile_score,
        tile_count, g.num_tiles, score, nvox, gscale);
    LAUNCH_CHECK ();
    if ((e = mark (ev, 2, st)) != hipSuccess) return e;
    hipLaunchKernelGGL (k_grad_reduce, dim3 ((g.num_knots + 255) / 256), dim3 (256), 0, st,
        g, cond, gscale, grad);
    LAUNCH_CHECK ();
    return mark (ev, 3, st);
}

hipError_t
launch_mi (const KGeom& g, const HGeom& h, int storage, int hist_blocks,
    const void* f, const void* m, const float* coeff, const float* blut,
    unsigned long long* jhist, unsigned long long* fhist, float* table,
    float* nvox_f, float* cond, const double* one, double* score, int64_t* nvox,
    double* grad, int64_t* hist_out, hipStream_t st, hipEvent_t* ev)
{
    const int nbins = h.fbins * h.mbins + h.fbins;
    hipError_t e = mark (ev, 0, st);
    if (e != hipSuccess) return e;
    // jhist and fhist are one allocation: joint bins then fixed bins.
    hipLaunchKernelGGL (k_zero_u64, dim3 ((nbins + 255) / 256), dim3 (256), 0, st, jhist, nbins);
    LAUNCH_CHECK ();
    const size_t lds1 = ((4 * (g.vpr[0] + g.vpr[1] + g.vpr[2]) + 192) * 4 + 8)
        + sizeof (unsigned long long) * nbins;
    const size_t lds2 = tile_lds_bytes (g) + sizeof (float) * h.fbins * h.mbins;
    dim3 grid (8 * g.tiles_per_xcd), block (TILE_THREADS);
    if ((e = mark (ev, 1, st)) != hipSuccess) return e;
    if (storage == 0)
        hipLaunchKernelGGL ((k_mi_hist<float>), dim3 (hist_blocks), block, lds1, st, g, h,
            (const float*) f, (const float*) m, coeff, blut, jhist, fhist);
    else
        hipLaunchKernelGGL ((k_mi_hist<uint16_t>), dim3 (hist_blocks), block, lds1, st, g, h,
            (const uint16_t*) f, (const uint16_t*) m, coeff, blut, jhist, fhist);
    LAUNCH_CHECK ();
    if ((e = mark (ev, 2, st)) != hipSuccess) return e;
    hipLaunchKernelGGL (k_mi_score, dim3 (1), dim3 (1024), 0, st, h, jhist, fhist, table,
        score, nvox, nvox_f, hist_out);
    LAUNCH_CHECK ();
    if ((e = mark (ev, 3, st)) != hipSuccess) return e;
    if (storage == 0)
        hipLaunchKernelGGL ((k_mi_tiles<float>), grid, block, lds2, st, g, h, (const float*) f,
            (const float*) m, coeff, blut, table, nvox_f, cond);
    else
        hipLaunchKernelGGL ((k_mi_tiles<uint16_t>), grid, block, lds2, st, g, h,
            (const uint16_t*) f, (const uint16_t*) m, coeff, blut, table, nvox_f, cond);
    LAUNCH_CHECK ();
    if ((e = mark (ev, 4, st)) != hipSuccess) return e;
    hipLaunchKernelGGL (k_grad_reduce, dim3 ((g.num_knots + 255) / 256), dim3 (256), 0, st,
        g, cond, one, grad);
    LAUNCH_CHECK ();
    return mark (ev, 5, st);
}

hipError_t
launch_fill_one (double* p, hipStream_t st)
{
    hipLaunchKernelGGL (k_fill_one, dim3 (1), dim3 (1), 0, st, p);
    return hipGetLastError ();
}

hipError_t
launch_volume_gradient (const float* img, float* out, const int dim[3], const float sp[3],
    const Mat3& inv, hipStream_t st)
{
    hipLaunchKernelGGL (k_volume_gradient, dim3 (2048), dim3 (256), 0, st, img, out, dim[0],
        dim[1], dim[2], sp[0], sp[1], sp[2], inv);
    return hipGetLastError ();
}

hipError_t
launch_to_bf16 (const float* src, uint16_t* dst, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL (k_to_bf16, dim3 (1024), dim3 (256), 0, st, src, dst, n);
    return hipGetLastError ();
}

} // namespace plmhip
