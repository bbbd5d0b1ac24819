// B-spline score/gradient kernels for gfx950 (MI355X).
//
// A persistent tile kernel walks the B-spline tiles (vpr0 x vpr1 x vpr2
// fixed voxels sharing the same 4x4x4 control knots).  Per tile:
//   1. the 64x3 tile coefficients (projected onto the moving index axes)
//      and the per-axis basis LUTs are in LDS;
//   2. the displacement is contracted separably: T1 = sum_i bx*c, then the
//      per-column terms A = sum_j by*T1, so each voxel only does the last
//      4-term sum in z;
//   3. each thread takes G voxels at a time (G in {1, 2, 4}, chosen by the
//      autotuner) and issues all of their loads before using any: the
//      eight corners as two 16-byte loads from a z-pair copy of the moving
//      image, the moving gradient sample; the fixed values come from LDS,
//      where the tile's rows were staged with wide loads;
//   4. dc/dv is parked in LDS and condensed into the tile's 64 knot slots by
//      the transposed separable contraction (x, then y, then z), run on the
//      matrix cores as single-term FMA chains;
//   5. the 192 slot values go to a slot-major buffer cond[slot][knot][c]
//      (one writer per entry, no atomics on floats), and a reduce kernel
//      sums each gradient component's 64 slots in FP64 in a fixed order.
// The score is a fixed-order sum of per-tile FP64 partials and the overlap
// count N a sum of per-workgroup integer partials; the only atomics on the
// MSE path are integer ones (the tile counters, the counts), so the
// results are bitwise reproducible run to run and do not depend on which
// workgroup took which tile.
//
// The MI histogram pass instead evaluates the displacement in the
// reference's k-j-i order without FMA so its partial-volume weights are
// bit-identical to the CPU flavor's, and accumulates them as 2^-32
// fixed-point integers (order independent) with LDS atomics.
//
// The file is compiled with -ffp-contract=off: every a*b+c below rounds
// twice unless fmaf() (or a matrix-core FMA) is written, which is done only
// where the reference's rounding is not reproduced anyway (the separable
// contractions).

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <float.h>
#include <algorithm>
#include <mutex>
#include <type_traits>

#include "bspline_kernels.h"

namespace plmhip {


// ---------------------------------------------------------------------------
// small helpers

__device__ __forceinline__ float ld (const float* __restrict__ p, int i) { return p[i]; }
__device__ __forceinline__ float ld (const uint16_t* __restrict__ p, int i)
{
    return __uint_as_float (((uint32_t) p[i]) << 16);
}

// A wave-uniform float moved to a scalar register.
__device__ __forceinline__ float
uniform (float v)
{
    return __int_as_float (__builtin_amdgcn_readfirstlane (__float_as_int (v)));
}

// n / d for n < 2^27, d <= 32, with m = ceil(2^32 / d).
__device__ __forceinline__ int
fdiv (int n, uint32_t m)
{
    return (int) __umulhi ((uint32_t) n, m);
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
// ALIGNED (direction cosines diagonal): the off-diagonal terms of proj are
// exact zeros, t*0 adds +-0 to a float, so dropping them leaves every bit
// unchanged.  The clamp is branch free: f = clamp(floor(ma), 0, dmax-1),
// frac = clamp(ma - f, 0, 1) and round = f + (frac >= 0.5) give exactly the
// three branches of li_clamp (and trunc(ma + 0.5) of ROUND_PLM_LONG).
// Outside voxels get in-range indices so their loads can be issued blindly.
template <bool ALIGNED>
__device__ __forceinline__ bool
correspond (const KGeom& g, const float fxyz[3], const float d[3], Sample& s)
{
    float t[3], mijk[3];
    t[0] = (fxyz[0] + d[0]) - g.morigin[0];
    t[1] = (fxyz[1] + d[1]) - g.morigin[1];
    t[2] = (fxyz[2] + d[2]) - g.morigin[2];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (ALIGNED)
            mijk[a] = t[a] * g.mproj[4*a];
        else
            mijk[a] = t[0] * g.mproj[3*a+0] + t[1] * g.mproj[3*a+1] + t[2] * g.mproj[3*a+2];
    }
    bool inside = true;
#pragma unroll
    for (int a = 0; a < 3; a++)
        inside = inside && (mijk[a] > -0.5f) && (mijk[a] < (float) g.mdim[a] - 0.5f);
    int f[3], r[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float ma = inside ? mijk[a] : 0.f;
        const float fl = fminf (fmaxf (floorf (ma), 0.f), (float) (g.mdim[a] - 2));
        const float fa2 = fminf (fmaxf (ma - fl, 0.f), 1.f);
        f[a] = (int) fl;
        r[a] = f[a] + (fa2 >= 0.5f ? 1 : 0);
        s.l2[a] = fa2;
        s.l1[a] = 1.0f - fa2;
    }
    s.mvf = f[0] + g.mdim[0] * (f[1] + g.mdim[1] * f[2]);
    s.mvr = r[0] + g.mdim[0] * (r[1] + g.mdim[1] * r[2]);
    return inside;
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

// Linear interpolation t*b + (1-t)*a as two FMAs; exact at t = 0 and
// t = 1, which the clamped border samples of li_clamp rely on.
__device__ __forceinline__ float
lerp (float a, float b, float t)
{
    return fmaf (t, b, fmaf (-t, a, a));
}

// Trilinear value and, optionally, its analytic derivative in moving index
// units (li_value and li_value_dx/dy/dz, src/base/interpolate.cxx:135-244,
// without their 1/spacing factor, which the condense applies per component)
// as nested lerps: the same polynomial in 14 FMAs (+ 15 operations for the
// derivative), rounding within a few ulp of the reference's
// product-and-sum form and exact on voxel centres and clamped borders.
// SHARED (fractions below 1: no clamped upper border) forms each lerp as
// a + t (b - a), still exact at t = 0; the differences are those the
// derivative needs, so value and derivative take 25 operations.  Every MSE
// kind uses the same form for the same voxels, so their scores agree bit
// for bit.
template <bool DERIV, bool SHARED = false>
__device__ __forceinline__ float
trilinear (const Sample& s, const float a[8], float dm[3])
{
    const float fx = s.l2[0], fy = s.l2[1], fz = s.l2[2];
    if (SHARED) {
        const float d0 = a[1] - a[0], d1 = a[3] - a[2], d2 = a[5] - a[4], d3 = a[7] - a[6];
        const float x00 = fmaf (fx, d0, a[0]);
        const float x10 = fmaf (fx, d1, a[2]);
        const float x01 = fmaf (fx, d2, a[4]);
        const float x11 = fmaf (fx, d3, a[6]);
        const float e0 = x10 - x00, e1 = x11 - x01;
        const float y0 = fmaf (fy, e0, x00);
        const float y1 = fmaf (fy, e1, x01);
        const float f0 = y1 - y0;
        if (DERIV) {
            const float gx0 = fmaf (fy, d1 - d0, d0);
            const float gx1 = fmaf (fy, d3 - d2, d2);
            dm[0] = fmaf (fz, gx1 - gx0, gx0);
            dm[1] = fmaf (fz, e1 - e0, e0);
            dm[2] = f0;
        }
        return fmaf (fz, f0, y0);
    }
    const float x00 = lerp (a[0], a[1], fx);
    const float x10 = lerp (a[2], a[3], fx);
    const float x01 = lerp (a[4], a[5], fx);
    const float x11 = lerp (a[6], a[7], fx);
    const float y0 = lerp (x00, x10, fy);
    const float y1 = lerp (x01, x11, fy);
    if (DERIV) {
        const float gx0 = lerp (a[1] - a[0], a[3] - a[2], fy);
        const float gx1 = lerp (a[5] - a[4], a[7] - a[6], fy);
        dm[0] = lerp (gx0, gx1, fz);
        dm[1] = lerp (x10 - x00, x11 - x01, fz);
        dm[2] = y1 - y0;
    }
    return lerp (y0, y1, fz);
}

// Fixed-image position, origin + step*ijk in the reference's order
// (POSITION_FROM_COORDS, src/base/volume_macros.h:82-96); ALIGNED drops the
// exact-zero off-diagonal products, bit for bit the same.
template <bool ALIGNED>
__device__ __forceinline__ void
position (float fxyz[3], const KGeom& g, int fi, int fj, int fk)
{
    const float x = (float) fi, y = (float) fj, z = (float) fk;
    if (ALIGNED) {
        fxyz[0] = g.img_origin[0] + x * g.fstep[0];
        fxyz[1] = g.img_origin[1] + y * g.fstep[4];
        fxyz[2] = g.img_origin[2] + z * g.fstep[8];
    } else {
#pragma unroll
        for (int a = 0; a < 3; a++)
            fxyz[a] = g.img_origin[a] + x * g.fstep[3*a+0] + y * g.fstep[3*a+1] + z * g.fstep[3*a+2];
    }
}

// inside_roi (src/util/bspline_correspond.cxx:52-74): the byte offset in r
// of the voxel holding world position xyz, tmp = xyz - origin projected in
// the reference's order, each index the floor, below 0 taken as 0 and at or
// past dim as dim - 1 (float_to_plm_long_clamp, :16-27).
__device__ __forceinline__ int
roi_offset (const RoiVol& r, const float xyz[3])
{
    const float t0 = xyz[0] - r.origin[0], t1 = xyz[1] - r.origin[1], t2 = xyz[2] - r.origin[2];
    int i[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const float pf = t0 * r.proj[3*a] + t1 * r.proj[3*a+1] + t2 * r.proj[3*a+2];
        i[a] = pf < 0.f ? 0 : (pf >= (float) r.dim[a] ? r.dim[a] - 1 : (int) floorf (pf));
    }
    return i[0] + r.dim[0] * (i[1] + r.dim[1] * i[2]);
}

__device__ __forceinline__ int
hist_bin (float v, float offset, float delta, int bins)
{
    int b = (int) floorf ((v - offset) / delta);
    // In-range intensities always give 0..bins-1; the clamp only keeps a
    // corrupted input from indexing outside the histogram.
    return b < 0 ? 0 : (b >= bins ? bins - 1 : b);
}

// ---------------------------------------------------------------------------
// tile bookkeeping

__device__ __forceinline__ void
tile_coords (const KGeom& g, int tile, int p[3])
{
    p[0] = tile % g.rdims[0];
    int t = tile / g.rdims[0];
    p[1] = t % g.rdims[1];
    p[2] = t / g.rdims[1];
}

// A voxel slot of a tile, packed in one register: x (bits 0-4), y (5-9),
// its column index x + vx y (10-19) and z (24-31), every field below 32
// (MAX_VPR) or 1024.  The voxel pass needs only the column and z (z as a
// float comes from its byte in one conversion); x, y and z as integers only
// at tiles that mask voxels.
__device__ __forceinline__ int slot_pack (int x, int y, int r, int z) { return x | (y << 5) | (r << 10) | (z << 24); }
__device__ __forceinline__ int slot_x (int w) { return w & 31; }
__device__ __forceinline__ int slot_y (int w) { return (w >> 5) & 31; }
__device__ __forceinline__ int slot_col (int w) { return (w >> 10) & 1023; }
__device__ __forceinline__ int slot_z (int w) { return (int) ((uint32_t) w >> 24); }
__device__ __forceinline__ float slot_zf (int w) { return (float) ((uint32_t) w >> 24); }

// (int) floor (x) in one instruction, for |x| < 2^31.
__device__ __forceinline__ int
floor_int (float x)
{
    int r;
    asm ("v_cvt_flr_i32_f32 %0, %1" : "=v" (r) : "v" (x));
    return r;
}

// ---------------------------------------------------------------------------
// raw buffer loads: 32-bit byte offsets from a wave-uniform base, so the
// per-voxel address arithmetic stays in 32-bit VALU; a load at or past the
// resource's byte size returns 0 without touching memory.

typedef __amdgpu_buffer_rsrc_t Rsrc;

__device__ __forceinline__ Rsrc
make_rsrc (const void* p, uint32_t bytes)
{
    // word 3 = 0x00020000: untyped raw buffer on gfx9 (no swizzle, no format)
    return __builtin_amdgcn_make_buffer_rsrc (const_cast<void*> (p), (short) 0, (int) bytes,
        0x00020000);
}

__device__ __forceinline__ int
load_u8 (Rsrc r, int off, int soff)
{
    return (int) __builtin_amdgcn_raw_buffer_load_b8 (r, off, soff, 0);
}

// Image element loads by storage type: one value, two x-neighbours, and an
// interleaved float3 (the moving gradient).
template <typename TI> struct Img;

template <> struct Img<float> {
    static constexpr int ES = 4;
    static __device__ __forceinline__ float one (Rsrc r, int off, int soff)
    {
        return __uint_as_float (__builtin_amdgcn_raw_buffer_load_b32 (r, off, soff, 0));
    }
    static __device__ __forceinline__ void vec3 (float o[3], Rsrc r, int off, int soff)
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b96 (r, off, soff, 0);
        o[0] = __uint_as_float (v[0]);
        o[1] = __uint_as_float (v[1]);
        o[2] = __uint_as_float (v[2]);
    }
    // 16 bytes of the z-pair image: (x, z), (x, z + 1), (x + 1, z), (x + 1, z + 1)
    static __device__ __forceinline__ void quad (float& a, float& b, float& c, float& d, Rsrc r,
        int off, int soff)
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128 (r, off, soff, 0);
        a = __uint_as_float (v[0]);
        b = __uint_as_float (v[1]);
        c = __uint_as_float (v[2]);
        d = __uint_as_float (v[3]);
    }
};

// BF16: a gradient sample (6 bytes) starts on an even byte; one 64-bit
// load from the dword boundary at or below it always covers it, and a
// funnel shift brings it down.  The 64-bit load may reach up to 4 bytes
// past the last element, so BF16 images carry 16 bytes of tail padding and
// their resources count 8 of them (img_bytes).
template <> struct Img<uint16_t> {
    static constexpr int ES = 2;
    static __device__ __forceinline__ float one (Rsrc r, int off, int soff)
    {
        return __uint_as_float (((uint32_t) __builtin_amdgcn_raw_buffer_load_b16 (r, off, soff, 0)) << 16);
    }
    // the 64 bits at byte o - (o & 2), o = off + soff, shifted down by
    // (o & 2) bytes (the alignment is that of the whole offset, so soff is
    // folded in)
    static __device__ __forceinline__ void load_shifted (uint32_t& lo, uint32_t& hi, Rsrc r,
        int off, int soff)
    {
        const int o = off + soff;
        const auto v = __builtin_amdgcn_raw_buffer_load_b64 (r, o & ~3, 0, 0);
        const uint32_t sh = (uint32_t) (o & 2) << 3;
        lo = __builtin_amdgcn_alignbit (v[1], v[0], sh);
        hi = v[1] >> sh;
    }
    static __device__ __forceinline__ void vec3 (float o[3], Rsrc r, int off, int soff)
    {
        uint32_t lo, hi;
        load_shifted (lo, hi, r, off, soff);
        o[0] = __uint_as_float (lo << 16);
        o[1] = __uint_as_float (lo & 0xffff0000u);
        o[2] = __uint_as_float (hi << 16);
    }
    // 8 bytes of the z-pair image (4-byte records, so always aligned):
    // (x, z), (x, z + 1), (x + 1, z), (x + 1, z + 1)
    static __device__ __forceinline__ void quad (float& a, float& b, float& c, float& d, Rsrc r,
        int off, int soff)
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64 (r, off, soff, 0);
        a = __uint_as_float (v[0] << 16);
        b = __uint_as_float (v[0] & 0xffff0000u);
        c = __uint_as_float (v[1] << 16);
        d = __uint_as_float (v[1] & 0xffff0000u);
    }
};

// The eight trilinear corner bins from the moving bin-quad image (one
// 32-bit record per voxel: the bins of (x, y), (x+1, y), (x, y+1),
// (x+1, y+1), low byte first): the record at byte offset off (slice z) and
// the one dz bytes on (slice z + 1), in the corner order x, then y, then z.
__device__ __forceinline__ void
bin_quads (int b[8], Rsrc r, int off, int dz)
{
    const uint32_t lo = __builtin_amdgcn_raw_buffer_load_b32 (r, off, 0, 0);
    const uint32_t hi = __builtin_amdgcn_raw_buffer_load_b32 (r, off, dz, 0);
#pragma unroll
    for (int n = 0; n < 4; n++) {
        b[n] = (int) ((lo >> (8 * n)) & 0xffu);
        b[n + 4] = (int) ((hi >> (8 * n)) & 0xffu);
    }
}

// Byte size of a resource over n elements of es bytes (BF16 images count
// part of their tail padding, see Img<uint16_t>).
__device__ __forceinline__ uint32_t
img_bytes (uint32_t n, int es)
{
    return n * (uint32_t) es + (es == 2 ? 8u : 0u);
}

// n / D for 0 <= n < 4096 and D < 256 with one 24-bit multiply:
// floor(n * ceil(2^20 / D) / 2^20) is exact there (the excess is below
// n / 2^20 < 1/256 < 1/D).
template <int D>
__device__ __forceinline__ int
cdiv (int n)
{
    static_assert (D > 0 && D < 256, "cdiv divisor out of range");
    constexpr uint32_t M = ((1u << 20) + D - 1) / D;
    return (int) (__umul24 ((uint32_t) n, M) >> 20);
}

// Tile edge lengths: compile-time for the isotropic edges the kernels are
// specialized for (V > 0), else the runtime vox_per_rgn.
template <int V>
struct Shape {
    int vx, vy, vz;
    uint32_t mx, my, mxy;
    __device__ __forceinline__ explicit Shape (const KGeom& g)
        : vx (V ? V : g.vpr[0]), vy (V ? V : g.vpr[1]), vz (V ? V : g.vpr[2]),
          mx (g.mag_vx), my (g.mag_vy), mxy (g.mag_vxy) {}
    __device__ __forceinline__ int div_x (int n) const
    {
        if constexpr (V > 0) return cdiv<V> (n); else return fdiv (n, mx);
    }
    __device__ __forceinline__ int div_y (int n) const
    {
        if constexpr (V > 0) return cdiv<V> (n); else return fdiv (n, my);
    }
    __device__ __forceinline__ int div_xy (int n) const
    {
        if constexpr (V > 0) return cdiv<V * V> (n); else return fdiv (n, mxy);
    }
};

// LDS carve-up of the tile kernel, in floats; every array read as float4
// starts on a 16-byte boundary (all offsets are multiples of 4 floats).
// Arrays whose lifetimes within a tile do not overlap share memory: T1
// (stages 2-3) lies under the dc/dv planes (stages 4-5), the condensed
// S2 (stage 6) under the column terms (stages 3-4).
struct TileLds {
    float* bx;      // [x][4] basis LUT along x
    float* by;      // [y][4]
    float* bzz;     // [z][8]: bz[0..3], pad
    float* coef;    // [k][j][i][c]: the tile's 64 knots
    float* col;     // [3][y][x][4]: per component a, the Horner coefficients q0..q3 of ml[a](z)
    float* s2;      // [z][c][j][i]: dc/dv condensed over x and y (reuses col)
    float* t1;      // [k][j][x][4]: sum_i bx coef (reuses dv)
    float* dv;      // [c][slot]: dc/dv of every voxel slot, one plane per component
    int dvp;        // plane stride (slots)
    double* red;    // block-sum scratch, 16 doubles
    float* tab;     // MI: dS/dP table, Bf*Bm
};

__device__ __forceinline__ TileLds
carve (float* sm, int vx, int vy, int vz, int slots)
{
    TileLds L;
    L.bx = sm;
    L.by = L.bx + 4 * vx;
    L.bzz = L.by + 4 * vy;
    L.coef = L.bzz + 8 * vz;
    L.col = L.coef + 192;
    L.s2 = L.col;
    L.t1 = L.col + tile_col_floats (vx, vy, vz);
    L.dv = L.t1;
    L.dvp = slots;
    // both sizes are multiples of 4 floats, so red stays 16-byte aligned
    L.red = (double*) (L.dv + tile_dv_floats (vx, slots));
    L.tab = (float*) (L.red + 16);
    return L;
}

// Basis LUTs, once per workgroup: bx, by, and bz into the first half of
// each bzz row.
template <int V>
__device__ __forceinline__ void
tile_luts (const TileLds& L, const Shape<V>& S, const float* __restrict__ blut)
{
    const int nxy = 4 * (S.vx + S.vy);
    for (int e = threadIdx.x; e < nxy + 4 * S.vz; e += TILE_THREADS) {
        const float b = blut[e];
        if (e < nxy) {
            L.bx[e] = b;                  // bx and by are contiguous in both
        } else {
            const int q = e - nxy;        // z*4 + k -> z*8 + k
            L.bzz[2 * q - (q & 3)] = b;
        }
    }
}

// Tile position and the knot coordinates of its first knot.
struct TilePos {
    int bx0, by0, bz0;      // first fixed voxel of the tile
    int knot0;              // knot index of its (0, 0, 0) control point
};

__device__ __forceinline__ TilePos
tile_pos (const KGeom& g, int tile)
{
    int p[3];
    tile_coords (g, tile, p);
    TilePos q;
    q.bx0 = g.roi_off[0] + g.vpr[0] * p[0];
    q.by0 = g.roi_off[1] + g.vpr[1] * p[1];
    q.bz0 = g.roi_off[2] + g.vpr[2] * p[2];
    q.knot0 = p[0] + g.cdims[0] * (p[1] + g.cdims[1] * p[2]);
    return q;
}

// The 3 coefficients of knot slot t < 64 of a tile, t = (k*4 + j)*4 + i,
// the slot order of coef.
__device__ __forceinline__ void
tile_coef_fetch (float c[3], const KGeom& g, int knot0, int t, const float* __restrict__ coeff)
{
    const int cd0 = g.cdims[0], cd01 = g.cdims[0] * g.cdims[1];
    const int knot = knot0 + (t & 3) + cd0 * ((t >> 2) & 3) + cd01 * (t >> 4);
    c[0] = coeff[3 * knot + 0];
    c[1] = coeff[3 * knot + 1];
    c[2] = coeff[3 * knot + 2];
}

// How a tile kernel forms the per-voxel coordinate m from the column
// terms: PROJECTED, the moving index coordinate (A = P T2, C and Z the
// position terms); DISPLACEMENT, the displacement itself in mm (A = T2,
// C = Z = 0); OFFSETS, the position terms only (A unused, the displacement
// comes from elsewhere).
enum ColumnMode { PROJECTED = 0, DISPLACEMENT = 1, OFFSETS = 2 };

// Stage 1 of a tile: its 64 knot coefficients (fetched ahead into
// registers; PROJECTED: projected onto the moving index axes, P c, so the
// contractions that follow carry moving-index units and the column terms
// need no projection).
template <int V, int MODE>
__device__ __forceinline__ void
tile_stage1 (const TileLds& L, const Shape<V>& S, const KGeom& g, const float c[3], int t)
{
    if (MODE != OFFSETS && t < 64) {
        float c0 = c[0], c1 = c[1], c2 = c[2];
        if (MODE == PROJECTED) {
            const float* P = g.mproj;
            const float a0 = fmaf (P[2], c2, fmaf (P[1], c1, P[0] * c0));
            const float a1 = fmaf (P[5], c2, fmaf (P[4], c1, P[3] * c0));
            const float a2 = fmaf (P[8], c2, fmaf (P[7], c1, P[6] * c0));
            c0 = a0;
            c1 = a1;
            c2 = a2;
        }
        L.coef[3 * t + 0] = c0;
        L.coef[3 * t + 1] = c1;
        L.coef[3 * t + 2] = c2;
    }
}

// Stage 2: T1[k][j][x][c] = sum_i bx[x][i] coef[k][j][i][c], one (k, j, x)
// per thread.  A coefficient row (k, j) is 12 floats: c fastest, then i.
template <int V>
__device__ __forceinline__ void
tile_t1 (const TileLds& L, const Shape<V>& S, int t)
{
    for (int e = t; e < 16 * S.vx; e += TILE_THREADS) {
        const int kj = S.div_x (e), x = e - kj * S.vx;
        const float4* q = reinterpret_cast<const float4*> (&L.coef[12 * kj]);
        const float4 q0 = q[0], q1 = q[1], q2 = q[2];
        const float4 b = reinterpret_cast<const float4*> (L.bx)[x];
        float4 o;
        o.x = fmaf (b.w, q2.y, fmaf (b.z, q1.z, fmaf (b.y, q0.w, b.x * q0.x)));
        o.y = fmaf (b.w, q2.z, fmaf (b.z, q1.w, fmaf (b.y, q1.x, b.x * q0.y)));
        o.z = fmaf (b.w, q2.w, fmaf (b.z, q2.x, fmaf (b.y, q1.y, b.x * q0.z)));
        o.w = 0.f;
        reinterpret_cast<float4*> (L.t1)[e] = o;     // e = kj*vx + x
    }
}

// Stage 3: the per-column terms of the moving-voxel coordinate.  With
// A[a][k] = sum_j by[y][j] T1[k][j][x][a] (the projected displacement still
// to be contracted along z) and C[a] = Bf[a] + M[a][0] x + M[a][1] y (x, y,
// z local to the tile, Bf the fraction of the tile's base, TileBase), a
// voxel (x, y, z) of the tile maps to the local moving index coordinates
//   ml[a] = sum_k bz_k(z / vz) A[a][k] + C[a] + M[a][2] z,  mijk = Bi + ml,
// which is P * (fixed position + displacement - moving origin) regrouped
// (bspline_find_correspondence_dcos, src/util/bspline_correspond.cxx:80-109,
// with the position of POSITION_FROM_COORDS, src/base/volume_macros.h:82-96).
// The cubic B-spline basis along z (bspline_basis_eval,
// src/base/bspline_xform.cxx:73-99) makes ml a cubic polynomial in z for
// each column, so the column stores its Horner coefficients: one float4
// plane per component, ml[a] = ((q3 z + q2) z + q1) z + q0, three FMAs a
// component for each voxel.  (MODE DISPLACEMENT: A only, the displacement
// in mm; OFFSETS: C and M[a][2] z only.)
//
// Returns whether every column this thread formed stays in the interior of
// the moving image over all the tile's slices: with bz >= 0, sum_k bz = 1
// and the last term linear in z, ml lies in C + [min_k A, max_k A] +
// [0 | M[a][2] (vz - 1)]; 0.01 voxel of margin covers the rounding of ml.
// Inside that range li_clamp's clamps are identities and Volume::is_inside
// holds.
template <int V, int MODE>
__device__ __forceinline__ bool
tile_columns (const TileLds& L, const Shape<V>& S, const KGeom& g, const TileBase& tb, int t)
{
    const int vx = S.vx;
    const float4* t1 = reinterpret_cast<const float4*> (L.t1);
    // 1/vz, 1/vz^2, 1/vz^3 times the basis polynomial's factors
    const float ivz = 1.f / (float) S.vz;
    const float h1 = 0.5f * ivz, h2 = 0.5f * ivz * ivz, h3 = (1.f / 6.f) * ivz * ivz * ivz;
    bool ok = true;
    for (int e = t; e < vx * S.vy; e += TILE_THREADS) {
        const int y = S.div_x (e), x = e - y * vx;
        // one float4 plane per component, so the voxel pass reads
        // consecutive 16-byte records (conflict-free) for consecutive columns
        float4* o = reinterpret_cast<float4*> (L.col) + e;
        const int np = vx * S.vy;
        float C[3] = {0.f, 0.f, 0.f}, Mz[3] = {0.f, 0.f, 0.f};
        if (MODE != DISPLACEMENT) {
            const float fi = (float) x, fj = (float) y;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                C[a] = (tb.bf[a] + g.mstep[3 * a] * fi) + g.mstep[3 * a + 1] * fj;
                Mz[a] = g.mstep[3 * a + 2];
            }
        }
        if (MODE == OFFSETS) {
#pragma unroll
            for (int a = 0; a < 3; a++) o[a * np] = make_float4 (C[a], Mz[a], 0.f, 0.f);
            continue;
        }
        const float4 b = reinterpret_cast<const float4*> (L.by)[y];
        float T[3][4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float4 r0 = t1[(4 * k + 0) * vx + x];
            const float4 r1 = t1[(4 * k + 1) * vx + x];
            const float4 r2 = t1[(4 * k + 2) * vx + x];
            const float4 r3 = t1[(4 * k + 3) * vx + x];
            T[0][k] = fmaf (b.w, r3.x, fmaf (b.z, r2.x, fmaf (b.y, r1.x, b.x * r0.x)));
            T[1][k] = fmaf (b.w, r3.y, fmaf (b.z, r2.y, fmaf (b.y, r1.y, b.x * r0.y)));
            T[2][k] = fmaf (b.w, r3.z, fmaf (b.z, r2.z, fmaf (b.y, r1.z, b.x * r0.z)));
        }
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float A0 = T[a][0], A1 = T[a][1], A2 = T[a][2], A3 = T[a][3];
            // sum_k bz_k(u) A_k = q0 + q1' u + q2' u^2 + q3' u^3 with
            // 6 bz = ((1-u)^3, 3u^3 - 6u^2 + 4, -3u^3 + 3u^2 + 3u + 1, u^3),
            // u = z / vz
            const float s02 = A0 + A2;
            const float q0 = fmaf (A1, 2.f / 3.f, s02 * (1.f / 6.f)) + C[a];
            const float q1 = fmaf (A2 - A0, h1, Mz[a]);
            const float q2 = fmaf (A1, -2.f, s02) * h2;
            const float q3 = fmaf (A1 - A2, 3.f, A3 - A0) * h3;
            o[a * np] = make_float4 (q0, q1, q2, q3);
            if (MODE == PROJECTED) {
                const float zb = Mz[a] * (float) (S.vz - 1);
                const float lo = C[a] + fminf (0.f, zb) + fminf (fminf (A0, A1), fminf (A2, A3));
                const float hi = C[a] + fmaxf (0.f, zb) + fmaxf (fmaxf (A0, A1), fmaxf (A2, A3));
                ok = ok & (lo >= tb.cl_lo[a] + 0.01f) & (hi <= tb.cl_hi[a] + 0.99f);
            }
        }
    }
    return ok;
}

// Stages 5-7: condense the tile's parked dc/dv into its 64 knot slots, the
// transposed separable contraction (x, then y, then z), and write them
// slot-major: cond[slot][knot][c], slot = (k*4 + j)*4 + i.  Each (knot,
// slot) entry has exactly one writer, as in bspline_sort_sets
// (src/register/bspline.cxx:377-398); slot-major keeps the reduce's reads
// coalesced.
//
// The three contractions run on the matrix cores as chains of
// v_mfma_f32_4x4x1_16b_f32: sixteen 4x4 blocks per instruction, block b
// adding the outer product of lanes 4b..4b+3 of its two operands, D[r][q] +=
// A[4b + r] * B[4b + q], the result in register r of lane 4b + q.  With one
// term per instruction each element is a single fused multiply-add, so every
// sum is the same FMA chain, in the same order, as the scalar form
// acc = b0 v0, acc = fma (b_n, v_n, acc).
//   S1 (over x): block = one (c, z) pair, row r = y - 4q for the q-th group
//      of four rows, column = the x basis index i; A = dv, B = bx.
//   S2 (over y): the S1 results stay in registers and feed the next chain
//      directly: register r of lane (b, i) is S1[c][z][4q + r][i], exactly the
//      A operand of row i of the same block; B = by.  Lane (b, j) ends with
//      S2[c][z][j][i] in register i.
//   S3 (over z): one wave, block = (c, j), row = i, column = k; A = S2 from
//      LDS ([z][c][j][i]: consecutive per lane), B = bz.
// Every operand read is clamped into its array, so lanes past the tile's
// pairs or rows (whose results are never used) read in bounds.
//
// SCALE: multiply component c by g.inv_msp[c] on the way out (the analytic
// derivative's 1/spacing, left out per voxel).
typedef float mfma_f4 __attribute__ ((ext_vector_type (4)));

__device__ __forceinline__ mfma_f4
mfma_4x4x1 (float a, float b, mfma_f4 c)
{
    return __builtin_amdgcn_mfma_f32_4x4x1f32 (a, b, c, 0, 0, 0);
}

template <int V, bool SCALE>
__device__ __forceinline__ void
tile_condense (const TileLds& L, const Shape<V>& S, const KGeom& g, int knot0,
    float* __restrict__ cond, int t)
{
    const int vx = S.vx, vy = S.vy, vz = S.vz;
    const int lane = t & 63, w = t >> 6, q4 = lane & 3;
    // S1 and S2: sixteen (c, z) pairs per wave, pair index c * vz + z
    const int npairs = 3 * vz;
    for (int h = w; 16 * h < npairs; h += TILE_THREADS / 64) {
        const int pr = min (16 * h + (lane >> 2), npairs - 1);
        const int c = pr >= 2 * vz ? 2 : (pr >= vz ? 1 : 0);
        const int z = pr - c * vz;
        const float* dpair = L.dv + c * L.dvp + z * vy * vx;
        mfma_f4 s2 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < (vy + 3) / 4; q++) {
            const float* drow = dpair + min (4 * q + q4, vy - 1) * vx;
            mfma_f4 s1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int x = 0; x < vx; x++) s1 = mfma_4x4x1 (drow[x], L.bx[4 * x + q4], s1);
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (4 * q + r < vy) s2 = mfma_4x4x1 (s1[r], L.by[4 * (4 * q + r) + q4], s2);
        }
        if (16 * h + (lane >> 2) < npairs)
            reinterpret_cast<float4*> (L.s2)[(z * 3 + c) * 4 + q4] = make_float4 (s2[0], s2[1], s2[2], s2[3]);
    }
    __syncthreads ();
    // S3 on the last wave: lane (c, j, k) = 16 c + 4 j + k, register i
    if (w == TILE_THREADS / 64 - 1) {
        const int row = min (lane, 47);
        mfma_f4 s3 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int z = 0; z < vz; z++) s3 = mfma_4x4x1 (L.s2[48 * z + row], L.bzz[8 * z + q4], s3);
        if (lane < 48) {
            const int c = lane >> 4, j = (lane >> 2) & 3, k = q4;
            const int cd0 = g.cdims[0], cd01 = g.cdims[0] * g.cdims[1];
            // 32-bit byte offsets: the buffer holds 192 * num_knots floats,
            // below 2^31 bytes (checked at context creation)
            const uint32_t plane = 12u * (uint32_t) g.num_knots;
            const Rsrc cr = make_rsrc (cond, 64u * plane);
            const float sc = SCALE ? g.inv_msp[c] : 1.f;
            const uint32_t off = (uint32_t) (16 * k + 4 * j) * plane
                + 4u * (uint32_t) (3 * (knot0 + cd0 * j + cd01 * k) + c);
#pragma unroll
            for (int i = 0; i < 4; i++)
                __builtin_amdgcn_raw_buffer_store_b32 (__float_as_uint (SCALE ? s3[i] * sc : s3[i]), cr,
                    (int) (off + (uint32_t) i * (plane + 12u)), 0, 0);
        }
    }
}

// Block sum of a double and an int in a fixed order (deterministic);
// the result is valid in thread 0.
__device__ __forceinline__ void
block_sum (const TileLds& L, double& s, int& n)
{
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down (s, off);
        n += __shfl_down (n, off);
    }
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads ();
    if ((threadIdx.x & 63) == 0) {
        L.red[2 * w] = s;
        reinterpret_cast<int*> (&L.red[2 * w + 1])[0] = n;
    }
    __syncthreads ();
    if (threadIdx.x == 0) {
        double a = 0;
        int b = 0;
        for (int i = 0; i < nw; i++) {
            a += L.red[2 * i];
            b += reinterpret_cast<const int*> (&L.red[2 * i + 1])[0];
        }
        s = a;
        n = b;
    }
}

// Sums over a wavefront by an xor butterfly: every lane ends with the same
// value (each step adds a pair of equal partner sums, and addition is
// commutative), formed in an order that depends on nothing but the lanes'
// values.
__device__ __forceinline__ double
wave_sum (double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor (v, off);
    return v;
}

__device__ __forceinline__ long long
wave_sum_ll (long long v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor (v, off);
    return v;
}

// ---------------------------------------------------------------------------
// Tile kernel.  KIND 0: MSE with the nearest-voxel moving gradient
// (bspline_score_g_mse, src/register/bspline_mse.cxx:519-720); KIND 1: MSE
// with the analytic trilinear derivative (bspline_score_p_mse,
// src/register/bspline_mse.cxx:1149-1384); KIND 2: MI pass 2, the PV-8
// derivative (pass 2 of bspline_score_g_mi,
// src/register/bspline_mi.cxx:1627-1716, and bspline_mi_pvi_8_dc_dv_dcos,
// src/register/bspline_mi.txx:9-120).

struct MiArgs {
    HGeom h;
    const float* table;        // dS/dP per joint bin (0 where j <= 1e-4)
    const float* nvox_f;       // (float) N
    const uint8_t* fbin;       // per-voxel fixed bin (bin image of the stage)
    const uint32_t* mquad;     // per-voxel moving bins of (x, y), (x+1, y), (x, y+1), (x+1, y+1)
};

// The pair a tile of the flattened queue belongs to; p is the pair of an
// earlier tile of the same workgroup (tiles only move forward).
__device__ __forceinline__ int
advance_pair (const PairDev* __restrict__ pairs, int npairs, int p, int tile)
{
    while (p + 1 < npairs && pairs[p + 1].tile0 <= tile) p++;
    return p;
}

// BATCH false: a single pair, whose geometry the compiler may then keep in
// registers for the whole launch.  vf is the dense displacement field of
// pair 0 (KIND 4 writes it, KIND 5 reads it), float3 interleaved per fixed
// voxel.
// KINDX & TILE_PLAIN: float storage with no z-pair image, the corners read
// from the moving image itself (four 8-byte loads a voxel).
constexpr int TILE_PLAIN = 16;

// The tile kernel's static LDS at most: per-pair score sums and overlap
// counts, the per-wave tile scores, the next tile (rounded up).
constexpr size_t TILE_STATIC_LDS = MAX_BATCH * (sizeof (double) + sizeof (int))
    + (TILE_THREADS / 64) * sizeof (double) + 16;

template <int KINDX, typename TI, int V, int G, bool BATCH, bool DYN>
__global__ __launch_bounds__ (TILE_THREADS)
void k_tiles (const PairDev* __restrict__ pairs, TileQueue q, const float* __restrict__ coeff0,
    const float* __restrict__ blut, double* __restrict__ part_score,
    int* __restrict__ part_count, MiArgs mi, float* __restrict__ vf, float fill,
    const int* __restrict__ skip)
{
    if (skip && *skip) return;
    // what this kind computes
    constexpr int KIND = KINDX & 15;
    constexpr bool PLAIN = (KINDX & TILE_PLAIN) != 0;
    static_assert (!PLAIN || std::is_same<TI, float>::value, "plain corners are float storage only");
    constexpr bool MI = KIND == 2;
    constexpr bool VF_OUT = KIND == 4;          // the displacement field only
    constexpr bool VF_IN = KIND == 5;           // the displacement read from vf
    constexpr bool WARP = KIND == 6 || KIND == 7;   // warped moving image into vf
    constexpr bool NEAREST = KIND == 7;             // (nearest voxel instead of trilinear)
    constexpr bool DERIV = KIND == 1 || KIND == 5;  // analytic trilinear derivative
    constexpr bool GRAD = KIND <= 2 || KIND == 5;   // dc/dv and the condense
    constexpr bool SCORE = !MI && !VF_OUT && !WARP; // squared differences, overlap count
    constexpr int MODE = VF_OUT ? DISPLACEMENT : (VF_IN ? OFFSETS : PROJECTED);
    const int npairs = BATCH ? q.npairs : 1;

    extern __shared__ float4 sm4[];
    __shared__ double s_ps[BATCH ? MAX_BATCH : 1];  // per-pair sums of this workgroup
    __shared__ int s_pn[BATCH ? MAX_BATCH : 1];
    __shared__ double s_ws[TILE_THREADS / 64];      // counted tiles: the last tile's score, per wave
    __shared__ int s_next;                          // counted tiles: the workgroup's next tile
    const int t = threadIdx.x;
    if (SCORE && t < npairs) {
        s_ps[t] = 0.0;
        s_pn[t] = 0;
    }
    // Persistent workgroups: those on XCD x = blockIdx % 8 take the tiles
    // of that XCD's contiguous run of the (flattened) tile queue, so
    // neighbouring tiles, which share image lines and knots, meet in its
    // L2.  A workgroup's first tile is run start + its index on the XCD.
    // On a short run (a few tiles per workgroup) the later ones follow at a
    // fixed stride.  On a long run (a batch of large pairs, MI pass 2 at
    // 512 x 512 x 160: tens of tiles per workgroup) a fixed stride lets the
    // workgroups drift apart, and the L2 then no longer holds their
    // scattered footprints (at 16 pairs of 232 x 198 x 178, 1.5 x the HBM
    // fetch of single-pair launches): there (DYN, built for batched
    // launches and MI pass 2) the later tiles come from the XCD's counter in
    // queue order, taken one tile ahead (its low 32 bits
    // count the tiles taken past the first wpx, its high 32 bits the
    // workgroups that have finished; the last to finish clears it).  The
    // counter's device-scope atomic, whose return the tile's loads wait
    // behind, costs a short run more than the drift it saves.  Which
    // workgroup takes a counted tile matters to no result: the score goes
    // out per tile (at a fixed stride, per (pair, workgroup)) and the
    // overlap counts are integers.
    const int wpx = gridDim.x >> 3;
    const int tbeg = (blockIdx.x & 7) * q.tiles_per_xcd;
    const int tend = min (tbeg + q.tiles_per_xcd, q.total_tiles);
    unsigned long long* const ctr = q.ctr + TILE_CTR_STRIDE * (blockIdx.x & 7);
    int tile = tbeg + (blockIdx.x >> 3);
    // the next counted tile, clamped to the run (past it: tend)
    const unsigned counted = tend - tbeg > wpx ? (unsigned) (tend - tbeg - wpx) : 0u;
    auto take = [&] () -> int {
        const unsigned k = (unsigned) (uint32_t) atomicAdd (ctr, 1ull);
        return k < counted ? tbeg + wpx + (int) k : tend;
    };
    constexpr bool dyn = DYN;
    int fut = tend;     // (thread 0) the tile after the current one, in flight
    if (dyn && t == 0 && tile < tend) fut = take ();
    int ptile = -1;     // the last tile whose score partial is pending

    // every pair shares the tile edges, so the LDS layout and the LUTs
    // are those of pair 0
    const Shape<V> S (pairs[0].g);
    const int vx = S.vx, vy = S.vy, vz = S.vz;
    const int nvox = vx * vy * vz;
    const int slots = tile_slots (nvox, G);
    const TileLds L = carve (reinterpret_cast<float*> (sm4), vx, vy, vz, slots);
    if (tile < tend) {
    tile_luts (L, S, blut);
    if (MI) {
        const int nj = mi.h.fbins * mi.h.mbins;
        for (int i = t; i < nj; i += TILE_THREADS) L.tab[i] = mi.table[i];
    }
    constexpr int ES = MI ? 1 : Img<TI>::ES;    // bytes per voxel of what is loaded

    // This thread's first-batch voxel slots t + 256u (x fastest), the same
    // in every tile, packed by slot_pack (x, y, the column index and z).
    // With compile-time tile edges and at most four slots per thread, every
    // batch's slots are kept (NSK of them), not only the first batch's.
    constexpr int NS = V > 0 ? (V * V * V + TILE_THREADS * G - 1) / (TILE_THREADS * G) * G : G;
    constexpr int NSK = NS <= 4 ? NS : G;
    int sxyz[NSK];
#pragma unroll
    for (int u = 0; u < NSK; u++) {
        const int v = u * TILE_THREADS + t;
        const int z = S.div_xy (v);
        const int r = v - z * (vx * vy);
        const int y = S.div_x (r);
        sxyz[u] = slot_pack (r - y * vx, y, r, z);
    }

    // Fetched one tile ahead: the knot coefficients (threads < 64) and the
    // first batch's fixed voxels (fixed bins for MI, stored displacements
    // for VF_IN); voxels past the image read 0.
    int p = advance_pair (pairs, npairs, 0, tile);
    int np = p;                             // pair of the tile fetched ahead
    TilePos nq = tile_pos (pairs[p].g, tile - pairs[p].tile0);
    float cpre[3] = {0.f, 0.f, 0.f};
    // Float score kinds with compile-time tile edges stage each tile's fixed
    // voxels in LDS: the threads with stage index s < vy * vz load row s of
    // the tile (y = s mod vy, z = s / vy) with ceil(vx / 4) 16-byte loads, a
    // tile ahead, and write it into the third dc/dv plane at slot s * vx + x.
    // That plane holds nothing before the voxel pass, where each thread
    // reads its voxels' values there before writing their dc/dv over them;
    // a 16-byte load costs the vector memory pipeline no more than a 4-byte
    // one.
    constexpr bool STAGE = SCORE && !VF_IN && V > 0 && std::is_same<TI, float>::value;
    constexpr int NQ = (V + 3) / 4;
    float frow[STAGE ? 4 * NQ : 1];
    float fpre[G];
    int fbpre[G];
    float dpre[G][3];
    // the voxel data of the first batch of tile position q of pair pp
    auto fetch_first = [&] (int pp, const TilePos& qq) {
        const KGeom& gg = pairs[pp].g;
        const int d0 = gg.fdim[0], d01 = gg.fdim[0] * gg.fdim[1];
        const int fb = qq.bx0 + d0 * qq.by0 + d01 * qq.bz0;
        const uint32_t nf = (uint32_t) (d01 * gg.fdim[2]);
        const Rsrc fr = make_rsrc (MI ? (const void*) mi.fbin : pairs[pp].fimg, img_bytes (nf, ES));
        const Rsrc vr = make_rsrc (vf, VF_IN ? nf * 12u : 0u);
#pragma unroll
        for (int u = 0; u < G; u++) {
            const int x = slot_x (sxyz[u]), y = slot_y (sxyz[u]), z = slot_z (sxyz[u]);
            const int fo = fb + x + (int) __umul24 (d0, y) + (int) __umul24 (d01, z);
            if (SCORE && !STAGE) fpre[u] = Img<TI>::one (fr, fo * ES, 0);
            if (MI) fbpre[u] = load_u8 (fr, fo, 0);
            if (VF_IN) Img<float>::vec3 (dpre[u], vr, fo * 12, 0);
        }
    };
    // row s = z * vy + y of tile position qq of pair pp, into frow, for
    // the threads with stage index s < vy * vz
    auto fetch_rows = [&] (int pp, const TilePos& qq, int s) {
        if (!STAGE || s >= vy * vz) return;
        const KGeom& gg = pairs[pp].g;
        const int d0 = gg.fdim[0], d01 = gg.fdim[0] * gg.fdim[1];
        const Rsrc fr = make_rsrc (pairs[pp].fimg, (uint32_t) (d01 * gg.fdim[2]) * 4u);
        const int z = S.div_y (s), y = s - z * vy;
        const int off = 4 * (qq.bx0 + d0 * (qq.by0 + y) + d01 * (qq.bz0 + z));
#pragma unroll
        for (int k = 0; k < NQ; k++)
            Img<float>::quad (frow[4 * k], frow[4 * k + 1], frow[4 * k + 2], frow[4 * k + 3], fr,
                off + 16 * k, 0);
    };
    auto store_rows = [&] (int s) {
        if (!STAGE || s >= vy * vz) return;
        float* o = L.dv + 2 * L.dvp + s * vx;
#pragma unroll
        for (int x = 0; x < vx; x++) o[x] = frow[x];
    };
    if (MODE != OFFSETS && t < 64)
        tile_coef_fetch (cpre, pairs[p].g, nq.knot0, t, (coeff0 && p == 0) ? coeff0 : pairs[p].coeff);
    fetch_first (p, nq);
    fetch_rows (p, nq, t);

    // this thread's squared differences and its wave's overlap count over
    // the tiles of the current pair, folded into s_ps/s_pn when the pair
    // changes (counted tiles: the squared differences over the current
    // tile only, sent out per tile)
    double score = 0.0;
    int count = 0;
    int rot = 0;
    while (tile < tend) {
        const TilePos P = nq;
        if (SCORE && np != p) {
            // the workgroup moved on to the next pair: fold the last one
            int wc = (t & 63) == 0 ? count : 0;
            block_sum (L, score, wc);
            if (t == 0) {
                s_ps[p] += score;
                s_pn[p] += wc;
            }
            score = 0.0;
            count = 0;
        }
        p = np;
        const KGeom& g = pairs[p].g;
        const PairDev& pd = pairs[p];
        // the tile's moving-index frame (wave-uniform: scalar loads)
        const TileBase tb = pd.tbase[tile - pd.tile0];
        int next = tile + wpx;
        if (dyn) {
            if (t == 0) {
                s_next = fut;
                if (fut < tend) fut = take ();
            }
        } else if (next < tend) {
            np = advance_pair (pairs, npairs, p, next);
            nq = tile_pos (pairs[np].g, next - pairs[np].tile0);
        }
        // The thread index and the slot coordinates are re-read opaquely
        // each tile, so the compiler recomputes what derives from them
        // instead of keeping it all live across the loop (which costs
        // occupancy).
        int tt = t;
        asm volatile ("" : "+v" (tt));
#pragma unroll
        for (int u = 0; u < NSK; u++) asm volatile ("" : "+v" (sxyz[u]));
        // The per-tile stages keep fewer than 256 threads busy, so which
        // wave leads them rotates from tile to tile (stage index tr = t +
        // 64 * tile count): over a few tiles every wave, and so every SIMD,
        // takes the same share.  tn is the next tile's stage index, under
        // which its coefficients are fetched ahead.
        const int tr = (tt + rot) & (TILE_THREADS - 1);
        rot += 64;
        const int tn = (tt + rot) & (TILE_THREADS - 1);
        __syncthreads ();       // the previous tile is done with coef, t1, s2, red
        if (dyn) {
            if (SCORE && t == 0 && ptile >= 0)
                part_score[ptile] = ((s_ws[0] + s_ws[1]) + s_ws[2]) + s_ws[3];
            // (the same in every lane: kept scalar, as the stride's is)
            next = __builtin_amdgcn_readfirstlane (s_next);
            if (next < tend) {
                np = advance_pair (pairs, npairs, p, next);
                nq = tile_pos (pairs[np].g, next - pairs[np].tile0);
            }
        }
        const bool has_next = next < tend;
        tile_stage1<V, MODE> (L, S, g, cpre, tr);
        __syncthreads ();
        if (MODE != OFFSETS && tn < 64 && has_next)
            tile_coef_fetch (cpre, pairs[np].g, nq.knot0, tn,
                (coeff0 && np == 0) ? coeff0 : pairs[np].coeff);
        if (MODE != OFFSETS) {
            tile_t1 (L, S, tr);
            __syncthreads ();
        }
        // MSE and MI tiles whose every column stays in the moving image's
        // interior take the voxel pass without clamps or inside tests (the
        // same floors, fractions and decisions; the trilinear lerps take
        // their shared-difference form): one flag per wave, ANDed after
        // the barrier (red is free here).
        constexpr bool FASTABLE = MODE == PROJECTED && !VF_OUT;
        const bool col_ok = tile_columns<V, MODE> (L, S, g, tb, tr);
        store_rows (tr);
        int* wflag = reinterpret_cast<int*> (L.red + 12);
        if (FASTABLE) {
            const bool w = __all (col_ok);
            if ((t & 63) == 0) wflag[t >> 6] = w ? 1 : 0;
        }
        __syncthreads ();
        bool fast = false;
        if (FASTABLE) {
            const int4 f4 = *reinterpret_cast<const int4*> (wflag);
            fast = (f4.x & f4.y & f4.z & f4.w) != 0;
        }

        // Stage 4: the voxels.  Thread t takes voxels t, t + 256, ... of
        // the tile, G at a time, issuing all their loads before using any.
        // Voxels outside the ROI or the tile still run (their fixed load may
        // fall past the image and read 0) and are masked at the end; z < ez
        // also rules out the slots past the tile.  A tile whose every column
        // stays in the moving image's interior and that lies whole inside
        // the ROI takes the lean pass (FAST): no clamp, no inside test, no
        // per-voxel mask but that of the slots past the tile, and its
        // overlap count is its voxel count.
        const int ex = min (vx, g.roi_end[0] - P.bx0);
        const int ey = min (vy, g.roi_end[1] - P.by0);
        const int ez = min (vz, g.roi_end[2] - P.bz0);
        const bool full = ex == vx && ey == vy && ez == vz;
        fast = fast && full;
        constexpr bool ROIABLE = KIND <= 3;     // the score and gradient kinds
        const int fd0 = g.fdim[0], fd01 = g.fdim[0] * g.fdim[1];
        const int md0 = g.mdim[0], md01 = g.mdim[0] * g.mdim[1];
        const int fbase = P.bx0 + fd0 * P.by0 + fd01 * P.bz0;
        const uint32_t nfix = (uint32_t) (fd01 * g.fdim[2]);
        const Rsrc fr = make_rsrc (MI ? (const void*) mi.fbin : pd.fimg, img_bytes (nfix, ES));
        const Rsrc mr = MI ? make_rsrc (mi.mquad, (uint32_t) (md01 * g.mdim[2]) * 4u)
                           : make_rsrc (pd.mimg, img_bytes ((uint32_t) (md01 * g.mdim[2]), ES));
        const Rsrc gr = make_rsrc (pd.mgrad,
            KIND == 0 ? img_bytes ((uint32_t) (md01 * g.mdim[2]) * 3, ES) : 0u);
        const Rsrc vr = make_rsrc (vf, VF_IN ? nfix * 12u : 0u);
        const int SX = md0 * ES, SXY = md01 * ES;
        // Float trilinear kinds read the corners from the z-pair image: two
        // 16-byte loads a voxel instead of four 8-byte ones (the vector
        // memory pipeline's cost goes by instructions, not bytes).
        constexpr bool ZPAIR = !MI && !NEAREST;
        constexpr int EZ = 2 * Img<TI>::ES;     // bytes per z-pair record
        const Rsrc zr = make_rsrc (pd.mpair, ZPAIR && !PLAIN ? (uint32_t) (md01 * g.mdim[2]) * EZ : 0u);
        const int SXZ = md0 * EZ;
        // The moving-index bounds of the tile's local frame, read once per
        // tile into registers: left in memory, they are reloaded per voxel
        // (each reload a scalar-load round trip).  (The values are
        // wave-uniform; readfirstlane puts them in scalar registers.)
        float tlo[3], thi[3], clo[3], chi[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            tlo[k] = uniform (tb.thr_lo[k]);
            thi[k] = uniform (tb.thr_hi[k]);
            clo[k] = uniform (tb.cl_lo[k]);
            chi[k] = uniform (tb.cl_hi[k]);
        }
        const int boff = __builtin_amdgcn_readfirstlane (tb.boff);
        // Regions of interest (MSE flavor i, src/register/bspline_mse.cxx:
        // 172-204; MI flavor h pass 2, src/register/bspline_mi.cxx:
        // 1416-1430): the fixed ROI's byte at the voxel, and the moving ROI's
        // at the moving position, whose ROI index coordinates are
        // R (m + Bi) + r0 = R m + rt (rt formed per tile in double).  Tiles
        // of a pair with a ROI take the full voxel pass.
        const uint8_t* fmask = ROIABLE ? pd.fmask : nullptr;
        const uint8_t* mroi = ROIABLE ? pd.mroi.img : nullptr;
        const bool use_roi = ROIABLE && (fmask != nullptr || mroi != nullptr);
        fast = fast && !use_roi;
        float rR[9], rt[3], rtop[3];
        int rd0 = 0, rd01 = 0;
        const Rsrc fmr = make_rsrc (fmask, ROIABLE ? nfix : 0u);
        const Rsrc rmr = make_rsrc (mroi, ROIABLE && mroi
            ? (uint32_t) (pd.mroi.dim[0] * pd.mroi.dim[1] * pd.mroi.dim[2]) : 0u);
        if (ROIABLE && mroi) {
            const int b2 = boff / md01, rem = boff - b2 * md01;
            const int b1 = rem / md0, b0 = rem - b1 * md0;
            const RoiVol& rv = pd.mroi;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                rt[k] = uniform ((float) (rv.r0[k] + rv.Rd[3*k] * b0 + rv.Rd[3*k+1] * b1
                    + rv.Rd[3*k+2] * b2));
                rtop[k] = (float) (rv.dim[k] - 1);
            }
#pragma unroll
            for (int k = 0; k < 9; k++) rR[k] = uniform (rv.R[k]);
            rd0 = rv.dim[0];
            rd01 = rv.dim[0] * rv.dim[1];
        }
        if (fast && t < 64) count += nvox;      // every voxel of a lean tile is inside

        auto voxels = [&] (auto fast_c) {
        constexpr bool FAST = decltype (fast_c)::value;
#pragma unroll
        for (int base = 0; base < nvox; base += TILE_THREADS * G) {
            const bool first = base == 0;
            bool in[G];
            float fr_[G][3];
            float fval[G];
            int fbv[G];
            float a[G][8];
            float mg[G][3];
            int bn[G][8];
            float nv[G];            // WARP, NEAREST: the nearest moving voxel
            int wofs[G];            // WARP: the fixed voxel written, or -1
            int rk[G];              // ROIs: 0 where a ROI excludes the voxel
#pragma unroll
            for (int u = 0; u < G; u++) {
                float dd[3] = {0.f, 0.f, 0.f};
                const int v = base + u * TILE_THREADS + tt;
                if (STAGE) fval[u] = L.dv[2 * L.dvp + v];
                // this thread's slot (a constant of the unrolled loops for
                // compile-time edges; at runtime edges only the first
                // batch's slots are kept)
                const int su = V > 0 ? base / TILE_THREADS + u : u;
                int sw;
                if (V > 0 ? su < NSK : first) {
                    sw = sxyz[su];
                } else {
                    const int z = S.div_xy (v);
                    const int rr = v - z * (vx * vy);
                    const int y = S.div_x (rr);
                    sw = slot_pack (rr - y * vx, y, rr, z);
                }
                // slots that may lie past the tile (the last batch's)
                const bool may_pad = V > 0 ? base + u * TILE_THREADS + TILE_THREADS > V * V * V : true;
                if (first) {
                    if (SCORE && !STAGE) fval[u] = fpre[u];
                    if (MI) fbv[u] = fbpre[u];
                    if (VF_IN) { dd[0] = dpre[u][0]; dd[1] = dpre[u][1]; dd[2] = dpre[u][2]; }
                } else if (!STAGE) {
                    const int fofs = fbase + slot_x (sw) + (int) __umul24 (fd0, slot_y (sw))
                        + (int) __umul24 (fd01, slot_z (sw));
                    if (SCORE && !STAGE) fval[u] = Img<TI>::one (fr, fofs * ES, 0);
                    if (MI) fbv[u] = load_u8 (fr, fofs, 0);
                    if (VF_IN) Img<float>::vec3 (dd, vr, fofs * 12, 0);
                }
                bool act;
                if (FAST) {
                    act = may_pad ? v < nvox : true;
                } else {
                    act = slot_x (sw) < ex && slot_y (sw) < ey && slot_z (sw) < ez;
                }
                if (WARP) {
                    const int fofs = fbase + slot_x (sw) + (int) __umul24 (fd0, slot_y (sw))
                        + (int) __umul24 (fd01, slot_z (sw));
                    wofs[u] = act ? fofs : -1;
                }
                // the column's Horner coefficients, evaluated at the slot's z
                const float4* cp = reinterpret_cast<const float4*> (L.col) + slot_col (sw);
                const int ncol = vx * vy;
                const float4 Q0 = cp[0], Q1 = cp[ncol], Q2 = cp[2 * ncol];
                const float zf = slot_zf (sw);
                float m[3];
                m[0] = fmaf (fmaf (fmaf (Q0.w, zf, Q0.z), zf, Q0.y), zf, Q0.x);
                m[1] = fmaf (fmaf (fmaf (Q1.w, zf, Q1.z), zf, Q1.y), zf, Q1.x);
                m[2] = fmaf (fmaf (fmaf (Q2.w, zf, Q2.z), zf, Q2.y), zf, Q2.x);
                if (VF_IN) {
                    // P * (position + d - moving origin), regrouped as for
                    // the interpolated displacement
                    const float* Pm = g.mproj;
#pragma unroll
                    for (int k = 0; k < 3; k++)
                        m[k] = m[k] + fmaf (Pm[3 * k + 2], dd[2], fmaf (Pm[3 * k + 1], dd[1], Pm[3 * k] * dd[0]));
                }
                if (VF_OUT) {
                    // bspline_interpolate_vf (src/base/bspline_interpolate.cxx:112-142):
                    // the ROI voxels of the field get the displacement in mm
                    if (act) {
                        const int fofs = fbase + slot_x (sw) + (int) __umul24 (fd0, slot_y (sw))
                            + (int) __umul24 (fd01, slot_z (sw));
                        float* o = vf + 3 * (size_t) fofs;
                        o[0] = m[0];
                        o[1] = m[1];
                        o[2] = m[2];
                    }
                    continue;
                }

                // Volume::is_inside (src/base/volume.cxx:899-906) and
                // li_clamp (src/base/interpolate.cxx:17-41), branch free:
                // f = clamp(floor(m), 0, dim-2), frac = clamp(m - f, 0, 1)
                // and round = f + (frac >= 0.5) give exactly its three
                // branches (and the rounding of ROUND_PLM_LONG).  Any m, NaN
                // included, yields in-range indices, so the loads are issued
                // unconditionally.  A lean tile's m stay in the interior,
                // where the clamps are identities and floor and fraction are
                // one instruction each (the fraction m - floor(m) is exact
                // for m >= 0).
                bool inside = true;
                int f[3], c[3];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    float fa;
                    if (FAST) {
                        f[k] = floor_int (m[k]);
                        fa = __builtin_amdgcn_fractf (m[k]);
                    } else {
                        // bitwise, not short-circuit: no branch per component
                        inside = inside & (m[k] > tlo[k]) & (m[k] < thi[k]);
                        const float fl = __builtin_amdgcn_fmed3f (floorf (m[k]), clo[k], chi[k]);
                        fa = fminf (fmaxf (m[k] - fl, 0.f), 1.f);
                        f[k] = (int) fl;
                    }
                    c[k] = fa >= 0.5f ? 1 : 0;
                    fr_[u][k] = fa;
                }
                in[u] = inside & act;
                // element offsets from the tile's base voxel as signed 24-bit
                // products (the local floor is >= -Bi, and |f| < mdim: slices
                // hold < 2^23 voxels, checked at context creation)
                if (NEAREST) {
                    // bspline_warp_dcos's nearest voxel: the rounded index
                    const int ridx = boff + (f[0] + c[0]) + __mul24 (f[1] + c[1], md0)
                        + __mul24 (f[2] + c[2], md01);
                    nv[u] = Img<TI>::one (mr, ridx * ES, 0);
                } else if (ZPAIR) {
                    // the floor voxel's index; the rounded voxel's (the
                    // gradient sample of flavor g) is one per-axis step on
                    const int e = boff + f[0] + __mul24 (f[1], md0) + __mul24 (f[2], md01);
                    if constexpr (PLAIN) {
                        // the four x-pairs of rows (y, z), (y + 1, z),
                        // (y, z + 1), (y + 1, z + 1): one address register,
                        // the row steps as scalar offsets
                        const int po = e * 4;
                        const auto v0 = __builtin_amdgcn_raw_buffer_load_b64 (mr, po, 0, 0);
                        const auto v1 = __builtin_amdgcn_raw_buffer_load_b64 (mr, po, SX, 0);
                        const auto v2 = __builtin_amdgcn_raw_buffer_load_b64 (mr, po, SXY, 0);
                        const auto v3 = __builtin_amdgcn_raw_buffer_load_b64 (mr, po, SX + SXY, 0);
                        a[u][0] = __uint_as_float (v0[0]); a[u][1] = __uint_as_float (v0[1]);
                        a[u][2] = __uint_as_float (v1[0]); a[u][3] = __uint_as_float (v1[1]);
                        a[u][4] = __uint_as_float (v2[0]); a[u][5] = __uint_as_float (v2[1]);
                        a[u][6] = __uint_as_float (v3[0]); a[u][7] = __uint_as_float (v3[1]);
                    } else {
                        const int off = e * EZ;
                        Img<TI>::quad (a[u][0], a[u][4], a[u][1], a[u][5], zr, off, 0);
                        Img<TI>::quad (a[u][2], a[u][6], a[u][3], a[u][7], zr, off, SXZ);
                    }
                    if (KIND == 0) {
                        // nearest voxel: the gradient sample of flavor g
                        // (src/register/bspline_mse.cxx:679-681)
                        // (e can pass 2^24, so a full multiply by the
                        // constant, not a 24-bit one)
                        const int goff = e * (3 * ES) + (c[0] ? 3 * ES : 0) + (c[1] ? 3 * SX : 0)
                            + (c[2] ? 3 * SXY : 0);
                        Img<TI>::vec3 (mg[u], gr, goff, 0);
                    }
                } else {
                    const int idx = boff + f[0] + __mul24 (md0, f[1]) + __mul24 (md01, f[2]);
                    // the bin quads of slices z and z + 1
                    bin_quads (bn[u], mr, 4 * idx, 4 * md01);
                }
                if constexpr (ROIABLE && !FAST) {
                    rk[u] = 1;
                    if (use_roi) {
                        if (fmask) {
                            const int fofs = fbase + slot_x (sw) + (int) __umul24 (fd0, slot_y (sw))
                                + (int) __umul24 (fd01, slot_z (sw));
                            rk[u] = load_u8 (fmr, fofs, 0);
                        }
                        if (mroi) {
                            int ro = 0;
#pragma unroll
                            for (int k = 0; k < 3; k++) {
                                const float pf = fmaf (rR[3*k], m[0], fmaf (rR[3*k+1], m[1],
                                    fmaf (rR[3*k+2], m[2], rt[k])));
                                const int ik = (int) __builtin_amdgcn_fmed3f (floorf (pf), 0.f, rtop[k]);
                                ro += k == 0 ? ik : (k == 1 ? ik * rd0 : ik * rd01);
                            }
                            rk[u] = rk[u] & load_u8 (rmr, ro, 0);
                        }
                    }
                }
            }
            if (!VF_OUT) {
            float bsum = 0.f;       // this batch's squared differences (<= G terms)
#pragma unroll
            for (int u = 0; u < G; u++) {
                const int v = base + u * TILE_THREADS + tt;
                if constexpr (ROIABLE && !FAST) in[u] = in[u] & (rk[u] != 0);
                if (WARP) {
                    // bspline_warp_dcos (src/base/bspline_warp.cxx:27-167):
                    // the moving value where the voxel lands, the fill
                    // value where it leaves the moving image
                    float val;
                    if (NEAREST) {
                        val = nv[u];
                    } else {
                        Sample s;
                        s.l2[0] = fr_[u][0]; s.l2[1] = fr_[u][1]; s.l2[2] = fr_[u][2];
                        float md[3];
                        val = trilinear<false, FAST> (s, a[u], md);
                    }
                    asm volatile ("" : "+v" (val));
                    if (wofs[u] >= 0) vf[wofs[u]] = in[u] ? val : fill;
                    continue;
                }
                float dc0 = 0.f, dc1 = 0.f, dc2 = 0.f;
                if (!MI) {
                    Sample s;
                    s.l2[0] = fr_[u][0]; s.l2[1] = fr_[u][1]; s.l2[2] = fr_[u][2];
                    float md[3];
                    float m_val = trilinear<DERIV, FAST> (s, a[u], md);
                    // Opaque here, so the select below cannot become a
                    // branch that sinks this voxel's loads (and a wait for
                    // them) into it.
                    float fv = fval[u];
                    asm volatile ("" : "+v" (m_val), "+v" (fv));
                    // (a lean tile's in[] is true but for the slots past
                    // the tile, and constant where none can be)
                    const float diff = in[u] ? m_val - fv : 0.f;
                    bsum += diff * diff;
                    // the wave's overlap count, on the scalar unit
                    if (!FAST) count += (int) __popcll (__ballot (in[u]));
                    if (KIND == 0) {
                        dc0 = diff * mg[u][0];
                        dc1 = diff * mg[u][1];
                        dc2 = diff * mg[u][2];
                    } else if (DERIV) {
                        dc0 = diff * md[0];
                        dc1 = diff * md[1];
                        dc2 = diff * md[2];
                    }
                } else {
                    float l1[3], l2[3];
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        l2[k] = fr_[u][k];
                        l1[k] = 1.0f - l2[k];
                    }
                    dc0 = 0.f; dc1 = 0.f; dc2 = 0.f;
                    float dw[24];
                    dw[0] = -l1[1] * l1[2];  dw[1] = -l1[0] * l1[2];  dw[2] = -(l1[0] * l1[1]);
                    dw[3] = l1[1] * l1[2];   dw[4] = -l2[0] * l1[2];  dw[5] = -(l2[0] * l1[1]);
                    dw[6] = -l2[1] * l1[2];  dw[7] = l1[0] * l1[2];   dw[8] = -(l1[0] * l2[1]);
                    dw[9] = l2[1] * l1[2];   dw[10] = l2[0] * l1[2];  dw[11] = -(l2[0] * l2[1]);
                    dw[12] = -l1[1] * l2[2]; dw[13] = -l1[0] * l2[2]; dw[14] = l1[0] * l1[1];
                    dw[15] = l1[1] * l2[2];  dw[16] = -l2[0] * l2[2]; dw[17] = l2[0] * l1[1];
                    dw[18] = -l2[1] * l2[2]; dw[19] = l1[0] * l2[2];  dw[20] = l1[0] * l2[1];
                    dw[21] = l2[1] * l2[2];  dw[22] = l2[0] * l2[2];  dw[23] = l2[0] * l2[1];
                    const float* row = &L.tab[fbv[u] * mi.h.mbins];
#pragma unroll
                    for (int n = 0; n < 8; n++) {
                        const float ds = row[bn[u][n]];
                        dc0 -= dw[3 * n + 0] * ds;
                        dc1 -= dw[3 * n + 1] * ds;
                        dc2 -= dw[3 * n + 2] * ds;
                    }
                    // projected in place, component by component, as the
                    // reference does (src/register/bspline_mi.txx:107-109);
                    // its division by N before the projection (:104-106), a
                    // linear map's common factor, is the reduce's FP64 1/N
                    // instead of three float divisions a voxel
                    const float* Pm = g.mproj;
                    dc0 = dc0 * Pm[0] + dc1 * Pm[1] + dc2 * Pm[2];
                    dc1 = dc0 * Pm[3] + dc1 * Pm[4] + dc2 * Pm[5];
                    dc2 = dc0 * Pm[6] + dc1 * Pm[7] + dc2 * Pm[8];
                    if (!in[u]) { dc0 = 0.f; dc1 = 0.f; dc2 = 0.f; }
                }
                if (GRAD) {
                    // every slot below tile_slots is in LDS; slots past the
                    // tile hold zeros and are never read
                    L.dv[v] = dc0;
                    L.dv[L.dvp + v] = dc1;
                    L.dv[2 * L.dvp + v] = dc2;
                }
            }
            score += (double) bsum;
            }
            if (first && has_next) fetch_first (np, nq);
        }
        };
        if constexpr (FASTABLE) {
            if (fast) voxels (std::true_type {});
            else voxels (std::false_type {});
        } else {
            voxels (std::false_type {});
        }
        if (has_next) fetch_rows (np, nq, tn);
        if (SCORE && dyn) {
            // a counted tile's score: each thread's voxels in their fixed
            // order, the lanes by the butterfly, the waves in order (thread
            // 0 writes it after the next barrier)
            const double ws = wave_sum (score);
            if ((t & 63) == 0) s_ws[t >> 6] = ws;
            score = 0.0;
            ptile = tile;
        }
        if (GRAD) {
            __syncthreads ();
            tile_condense<V, DERIV> (L, S, g, P.knot0, pd.cond, tr);
        }
        tile = next;
    }
    if (SCORE) {
        // fold the last pair: each thread's tiles in their fixed order, then
        // the threads in a fixed order
        int wc = (t & 63) == 0 ? count : 0;
        block_sum (L, score, wc);
        if (t == 0) {
            s_ps[p] += score;
            s_pn[p] += wc;
        }
    }
    }
    // one partial per (pair, workgroup), zeros for the pairs this workgroup
    // did not touch (counted tiles: the last tile's score partial, and the
    // counts only)
    if (SCORE) {
        __syncthreads ();
        if (dyn) {
            if (t == 0 && ptile >= 0) part_score[ptile] = ((s_ws[0] + s_ws[1]) + s_ws[2]) + s_ws[3];
        } else if (t < npairs) {
            part_score[t * gridDim.x + blockIdx.x] = s_ps[t];
        }
        if (t < npairs) part_count[t * gridDim.x + blockIdx.x] = s_pn[t];
    }
    // done with the counter: the XCD's last workgroup to finish clears it
    // (every workgroup took its last tile before adding its finish, on the
    // same word)
    if (dyn && t == 0) {
        const unsigned long long old = atomicAdd (ctr, 1ull << 32);
        if ((int) (old >> 32) == wpx - 1) atomicExch (ctr, 0ull);
    }
}

// ---------------------------------------------------------------------------
// Reduce: REDUCE_BLOCK (64) gradient components per workgroup of 256
// threads, four per component: thread (q, c) sums slots 16q .. 16q + 15 of
// component c in FP64, and the four partial sums are added in q order --
// a fixed order, so the result is bitwise reproducible (the slot sum of
// bspline_condense_smetric_grad, src/register/bspline.cxx:400-413, in
// double) -- then scaled by 2/N for MSE (src/register/bspline_mse.cxx:43-73)
// or 1/N for MI (the per-voxel division of src/register/bspline_mi.txx:
// 104-106, moved here: the condense and the projection are linear).  Workgroups are dealt to the pairs of the launch in order.
// For MSE every workgroup also sums its pair's per-workgroup overlap counts
// (integers, so every workgroup gets the same N with no cross-workgroup
// traffic), fetched with the slot loads; the pair's first workgroup also
// sums the score partials in a fixed order -- per (pair, workgroup), or
// with tile_scores (the tile kernel took counted tiles) per tile in tile
// order -- and writes score = raw/N (FLT_MAX if N < 1) and N.

constexpr int REDUCE_THREADS = 4 * REDUCE_BLOCK;

template <bool MSE>
__global__ __launch_bounds__ (REDUCE_THREADS)
void k_reduce (const PairDev* __restrict__ pairs, int npairs, const double* __restrict__ part_score,
    const int* __restrict__ part_count, int nparts, int tile_scores, EvalOut out0, int with_grad)
{
    if (out0.skip && *out0.skip) return;
    __shared__ long long s_n[REDUCE_THREADS / 64];
    __shared__ double s_s[REDUCE_THREADS / 64];
    __shared__ double s_q[4][REDUCE_BLOCK];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = REDUCE_THREADS / 64;
    // the pair this workgroup works for (workgroups are dealt to pairs in
    // order, pairs[p].blk0 being the first of pair p)
    int p = 0;
    while (p + 1 < npairs && pairs[p + 1].blk0 <= (int) blockIdx.x) p++;
    const PairDev& pd = pairs[p];
    const int blk = blockIdx.x - pd.blk0;
    const bool lead = blk == 0;
    const int nc = 3 * pd.g.num_knots;
    const int i = blk * REDUCE_BLOCK + lane;        // this thread's component
    const int q = wave;                              // and its quarter of the slots
    const float* c = pd.cond + (size_t) (16 * q) * nc + (i < nc ? i : 0);
    double* grad = (out0.grad && p == 0) ? out0.grad : pd.grad;
    // all 16 slot loads are issued before any is consumed, then summed in
    // slot order; the pair's partials are fetched meanwhile.
    // The loads are unconditional (c stays in range) so nothing serializes
    // them.
    float v[16];
    if (with_grad) {
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = c[(size_t) m * nc];
    } else {
#pragma unroll
        for (int m = 0; m < 16; m++) v[m] = 0.f;
    }
    long long n = 0;
    double sc = 0.0;
    if (MSE) {
        // the overlap counts (integers: any order gives the same N), 16
        // bytes a lane where the pair's counts are 16-byte aligned (the
        // tile grid is a multiple of eight workgroups), two loads a lane in
        // flight together for up to 2048 partials
        const int* pc = part_count + (size_t) p * nparts;
        const double* ps = part_score + (size_t) p * nparts;
        int t4 = 0;
        if ((reinterpret_cast<uintptr_t> (pc) & 15) == 0) {
            const int4* pc4 = reinterpret_cast<const int4*> (pc);
            const int n4 = nparts >> 2;
            for (int t0 = 0; t0 < n4; t0 += 2 * REDUCE_THREADS) {
                const int ta = t0 + threadIdx.x, tb = ta + REDUCE_THREADS;
                const int4 a = ta < n4 ? pc4[ta] : make_int4 (0, 0, 0, 0);
                const int4 b = tb < n4 ? pc4[tb] : make_int4 (0, 0, 0, 0);
                n += ((long long) a.x + a.y) + ((long long) a.z + a.w)
                    + ((long long) b.x + b.y) + ((long long) b.z + b.w);
            }
            t4 = 4 * n4;
        }
        for (int t = t4 + threadIdx.x; t < nparts; t += REDUCE_THREADS) n += pc[t];
        // the lead's per-workgroup score partials, thread t adding t,
        // t + 256, ... in turn
        if (lead && !tile_scores)
            for (int t = threadIdx.x; t < nparts; t += REDUCE_THREADS) sc += ps[t];
        if (lead && tile_scores) {
            // per-tile score partials (counted tiles), in tile order
            constexpr int CT = 32;
            const double* pt = part_score + pd.tile0;
            const int nt = pd.g.num_tiles;
            for (int t0 = 0; t0 < nt; t0 += CT * REDUCE_THREADS) {
                double s[CT];
#pragma unroll
                for (int k = 0; k < CT; k++) {
                    const int t = t0 + k * REDUCE_THREADS + threadIdx.x;
                    s[k] = t < nt ? pt[t] : 0.0;
                }
#pragma unroll
                for (int k = 0; k < CT; k++) sc += s[k];
            }
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < 16; m++) acc += (double) v[m];
    s_q[q][lane] = acc;
    double scale = 1.0;
    if (MSE) {
        n = wave_sum_ll (n);
        sc = wave_sum (sc);
        if (lane == 0) { s_n[wave] = n; s_s[wave] = sc; }
    }
    if (!MSE) {
        // MI: dc/dv's factor 1/N (N written by k_mi_score), in FP64 here
        const int64_t* nv = (out0.nvox && p == 0) ? out0.nvox : pd.nvox;
        const long long Nv = *nv;
        scale = Nv >= 1 ? 1.0 / (double) Nv : 0.0;
    }
    __syncthreads ();
    if (MSE) {
        long long Nv = 0;
        double Sv = 0.0;
        for (int w = 0; w < nwaves; w++) { Nv += s_n[w]; Sv += s_s[w]; }
        scale = Nv >= 1 ? 2.0 / (double) Nv : 0.0;
        if (lead && threadIdx.x == 0) {
            double* score = (out0.score && p == 0) ? out0.score : pd.score;
            int64_t* nvox = (out0.nvox && p == 0) ? out0.nvox : pd.nvox;
            *score = Nv >= 1 ? Sv / (double) Nv : (double) FLT_MAX;
            *nvox = Nv;
        }
    }
    if (q == 0 && with_grad && i < nc)
        grad[i] = (((s_q[0][lane] + s_q[1][lane]) + s_q[2][lane]) + s_q[3][lane]) * scale;
}

// ---------------------------------------------------------------------------
// Mattes MI pass 1 and score

// Displacement in the reference's k-j-i order, bx[i]*(by[j]*bz[k]) weights,
// no FMA: bit-identical to bspline_interp_pix_c
// (src/base/bspline_interpolate.cxx:78-110).  The tile's coefficients are
// read straight from global memory at wave-uniform addresses, so they come
// through the scalar cache into SGPRs and every multiply takes its
// coefficient as a scalar operand.
__device__ __forceinline__ void
exact_displacement (float out[3], const float* bx, const float* by, const float* bz,
    const float* __restrict__ coeff, int knot0, int cd0, int cd01)
{
    float o0 = 0.f, o1 = 0.f, o2 = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float C = bz[k];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float B = by[j] * C;
            const float* cf = &coeff[3 * (knot0 + k * cd01 + j * cd0)];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float A = bx[i] * B;
                o0 += A * cf[3 * i + 0];
                o1 += A * cf[3 * i + 1];
                o2 += A * cf[3 * i + 2];
            }
        }
    }
    out[0] = o0; out[1] = o1; out[2] = o2;
}

// Fixed-point PV weights (units of 2^-32): q = floor(w*2^32 + 0.5), the
// remainder to the first largest weight, so each voxel adds exactly 2^32.
// For a weight 0 <= w < 1, x = w * 2^32 is exact in float and below 2^32,
// its truncation is floor(x) and its fraction x - floor(x) (v_fract) is
// exact, so q = floor(x) + (frac >= 1/2) is the double expression's value
// and fits 32 bits.  The eight q then sum to
// 2^32 - r with |r| <= 4, so the 32-bit sum, negated, is r, and the first
// largest q plus r still fits 32 bits.  A weight of 1 (a corner on a voxel
// centre; the other seven are then 0) is 2^32: returned as true, with that
// corner in amax and the q left meaningless.
__device__ __forceinline__ bool
pv_quantize (const float w[8], uint32_t q[8], int& amax)
{
    uint32_t sum = 0u;
    amax = 0;
    float wmax = w[0];
#pragma unroll
    for (int n = 0; n < 8; n++) {
        const float x = w[n] * 4294967296.0f;
        // (the clamp keeps w = 1's conversion in range)
        q[n] = (uint32_t) fminf (x, 4294967040.0f) + (__builtin_amdgcn_fractf (x) >= 0.5f ? 1u : 0u);
        sum += q[n];
        if (n > 0 && w[n] > wmax) {
            wmax = w[n];
            amax = n;
        }
    }
    const uint32_t r = 0u - sum;
#pragma unroll
    for (int n = 0; n < 8; n++)
        if (n == amax) q[n] += r;
    return wmax == 1.0f;
}

// The corners that fall in corner 0's bin merged into q[0] (their own q
// set to 0), in 32 bits: the merged value M lies in [0, 2^32], and the
// 32-bit sum m is M mod 2^32, so M = 2^32 exactly where m is 0 and some
// merged q is not (the eight q sum to 2^32).  Returns M.
__device__ __forceinline__ unsigned long long
pv_merge (uint32_t q[8], const int b[8])
{
    uint32_t m = q[0], any = q[0];
#pragma unroll
    for (int n = 1; n < 8; n++) {
        const uint32_t v = b[n] == b[0] ? q[n] : 0u;
        m += v;
        any |= v;
        q[n] -= v;
    }
    return (m == 0u && any != 0u) ? (1ull << 32) : (unsigned long long) m;
}

// Pass 1: PV-8 joint histogram.  Follows the pass-1 loop of
// bspline_score_g_mi (src/register/bspline_mi.cxx:1567-1615) with
// Joint_histogram::add_pvi_8 (src/register/joint_histogram.cxx:423-487).
// Persistent: each workgroup walks tiles with a grid stride and keeps a
// private int64 histogram in LDS, flushed once with integer atomics.  Bins
// come from the stage's bin images.  Because integer addition is
// associative, contributions to one bin may be merged before the atomic:
// the corners of a voxel that share corner 0's bin are merged in
// registers, so a voxel in a uniform region costs one LDS atomic.  (The
// lanes of a wavefront are not merged: the conflicts of same-address LDS
// atomics cost the LDS pipe less than the shuffles of a wave reduction
// cost the VALU, 0.73 against 0.82 ms at 512x512x160.)

template <bool ALIGNED>
__global__ __launch_bounds__ (TILE_THREADS)
void k_mi_hist (KGeom g, HGeom h, const uint8_t* __restrict__ fbin,
    const uint32_t* __restrict__ mquad, const float* __restrict__ coeff,
    const float* __restrict__ blut, unsigned long long* __restrict__ jhist,
    unsigned long long* __restrict__ fhist, const float* __restrict__ ptab,
    const uint8_t* __restrict__ fmask, RoiVol mroi, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    // the histograms first (the dynamic LDS base is 16-byte aligned), then
    // the LUTs; every pointer stays an LDS pointer, so the atomics are LDS
    // atomics
    extern __shared__ unsigned long long sm_u64[];
    const KGeom& G_ = g;
    const int vx = G_.vpr[0], vy = G_.vpr[1], vz = G_.vpr[2];
    unsigned long long* s_j = sm_u64;
    unsigned long long* s_f = s_j + h.fbins * h.mbins;
    float* s_bx = reinterpret_cast<float*> (s_f + h.fbins);
    float* s_by = s_bx + 4 * vx;
    float* s_bz = s_by + 4 * vy;
    const Rsrc qr = make_rsrc (mquad, (uint32_t) (G_.mdim[0] * G_.mdim[1] * G_.mdim[2]) * 4u);
    const int qdz = 4 * G_.mdim[0] * G_.mdim[1];

    const int nbins = h.fbins * h.mbins + h.fbins;
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) s_j[i] = 0ull;
    const int nl = 4 * (vx + vy + vz);
    for (int i = threadIdx.x; i < nl; i += blockDim.x) s_bx[i] = blut[i];

    const int nvox = vx * vy * vz;
    const int rounds = (nvox + TILE_THREADS - 1) / TILE_THREADS;
    for (int tile = blockIdx.x; tile < G_.num_tiles; tile += gridDim.x) {
        int p[3];
        tile_coords (G_, tile, p);
        const int knot0 = p[0] + G_.cdims[0] * (p[1] + G_.cdims[1] * p[2]);
        __syncthreads ();
        // every lane runs the same number of rounds (the wave-level merges
        // need the whole wavefront), inactive lanes contribute nothing
        for (int r = 0; r < rounds; r++) {
            const int v = r * TILE_THREADS + threadIdx.x;
            const int qz = fdiv (v, G_.mag_vxy);
            const int rr = v - qz * vx * vy;
            const int qy = fdiv (rr, G_.mag_vx);
            const int qx = rr - qy * vx;
            const int fi = vx * p[0] + qx, fj = vy * p[1] + qy, fk = vz * p[2] + qz;
            bool act = v < nvox && fi < G_.fdim[0] && fj < G_.fdim[1] && fk < G_.fdim[2];
            // flavor h: the fixed ROI at the walk position
            // (src/register/bspline_mi.cxx:1314-1318)
            if (fmask && act) act = fmask[fi + G_.fdim[0] * (fj + G_.fdim[1] * fk)] != 0;
            Sample s;
            float d[3] = {0.f, 0.f, 0.f};
            float fxyz[3];
            if (ptab) {
                // the fixed positions of the reference's LOOP_Z/Y/X walk,
                // where they differ from the direct formula (see
                // k_pass1_positions)
                const int fidx0 = act ? fi + G_.fdim[0] * (fj + G_.fdim[1] * fk) : 0;
                fxyz[0] = ptab[3 * fidx0 + 0];
                fxyz[1] = ptab[3 * fidx0 + 1];
                fxyz[2] = ptab[3 * fidx0 + 2];
            } else {
                position<ALIGNED> (fxyz, G_, fi, fj, fk);
            }
            if (act)
                exact_displacement (d, &s_bx[4 * qx], &s_by[4 * qy], &s_bz[4 * qz], coeff, knot0,
                    G_.cdims[0], G_.cdims[0] * G_.cdims[1]);
            act = correspond<ALIGNED> (G_, fxyz, d, s) && act;
            // and the moving ROI at mxyz = fxyz + d
            // (bspline_find_correspondence_dcos_roi, src/util/bspline_correspond.cxx:113-145)
            if (mroi.img && act) {
                const float mxyz[3] = {fxyz[0] + d[0], fxyz[1] + d[1], fxyz[2] + d[2]};
                act = mroi.img[roi_offset (mroi, mxyz)] != 0;
            }
            int b[8];
            bin_quads (b, qr, 4 * s.mvf, qdz);
            const int fidx = act ? fi + G_.fdim[0] * (fj + G_.fdim[1] * fk) : 0;
            const int fb = fbin[fidx];
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
            uint32_t q[8];
            int amax;
            const bool one = pv_quantize (w, q, amax);
            // (the fixed histogram is not counted here: a voxel adds exactly
            // 2^32 to its fixed bin's row, so k_mi_score reads the counts off
            // the row sums)
            const int row = fb * h.mbins;
            if (one) {
                // a corner on a voxel centre takes the whole voxel
                int bo = b[0];
#pragma unroll
                for (int n = 1; n < 8; n++)
                    if (n == amax) bo = b[n];
                if (act) atomicAdd (&s_j[row + bo], 1ull << 32);
            }
            act = act && !one;
            const unsigned long long q0 = pv_merge (q, b);
            if (act) atomicAdd (&s_j[row + b[0]], q0);
#pragma unroll
            for (int n = 1; n < 8; n++)
                if (act && q[n] != 0u) atomicAdd (&s_j[row + b[n]], (unsigned long long) q[n]);
        }
    }
    __syncthreads ();
    for (int i = threadIdx.x; i < h.fbins * h.mbins; i += blockDim.x)
        if (s_j[i]) atomicAdd (&jhist[i], s_j[i]);
}

// Pass 1 of the reference walks the fixed image with the LOOP_Z / LOOP_Y /
// LOOP_X macros (src/base/volume_macros.h:38-75, used at
// src/register/bspline_mi.cxx:1567-1572): each row (j, k) starts at
// (origin + k step[:,2]) + j step[:,1] and every x step adds step[:,0] to the
// float position, so voxel i of a row carries i rounded additions; the tile
// kernels form origin + step ijk directly (POSITION_FROM_COORDS, as pass 2
// does).  The two agree wherever the products and sums are exact (dyadic
// spacings on axis-aligned images); elsewhere the PV weights, and so the
// histogram, would differ.  One thread per row replays the walk: counts the
// voxels whose walk position differs from the direct one (mismatch), and
// with out writes the walk positions (float3 per voxel) for pass 1.
template <bool ALIGNED>
__global__ __launch_bounds__ (256)
void k_pass1_positions (KGeom g, float* __restrict__ out, unsigned int* __restrict__ mismatch)
{
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= g.fdim[1] * g.fdim[2]) return;
    const int j = row % g.fdim[1], k = row / g.fdim[1];
    const float fj = (float) j, fk = (float) k;
    const float* s = g.fstep;
    float w[3];
#pragma unroll
    for (int a = 0; a < 3; a++) w[a] = (g.img_origin[a] + fk * s[3 * a + 2]) + fj * s[3 * a + 1];
    unsigned int bad = 0;
    const size_t base = (size_t) g.fdim[0] * row;
    for (int i = 0; i < g.fdim[0]; i++) {
        float d[3];
        position<ALIGNED> (d, g, i, j, k);
        bad += (d[0] != w[0] || d[1] != w[1] || d[2] != w[2]) ? 1u : 0u;
        if (out) {
            out[3 * (base + i) + 0] = w[0];
            out[3 * (base + i) + 1] = w[1];
            out[3 * (base + i) + 2] = w[2];
        }
        w[0] += s[0];
        w[1] += s[3];
        w[2] += s[6];
    }
    if (bad) atomicAdd (mismatch, bad);
}

template <bool ALIGNED>
__global__ __launch_bounds__ (256)
void k_roi_fixed (KGeom g, RoiVol r, const float* __restrict__ ptab, uint8_t* __restrict__ out)
{
    const int64_t n = (int64_t) g.fdim[0] * g.fdim[1] * g.fdim[2];
    for (int64_t v = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; v < n;
         v += (int64_t) gridDim.x * blockDim.x) {
        const int i = (int) (v % g.fdim[0]);
        const int64_t jk = v / g.fdim[0];
        const int j = (int) (jk % g.fdim[1]), k = (int) (jk / g.fdim[1]);
        float xyz[3];
        if (ptab) {
            xyz[0] = ptab[3 * v];
            xyz[1] = ptab[3 * v + 1];
            xyz[2] = ptab[3 * v + 2];
        } else {
            position<ALIGNED> (xyz, g, i, j, k);
        }
        out[v] = r.img[roi_offset (r, xyz)] != 0 ? 1 : 0;
    }
}

hipError_t
launch_roi_fixed (const KGeom& g, const RoiVol& r, const float* ptab, uint8_t* out, hipStream_t st)
{
    if (g.aligned)
        hipLaunchKernelGGL ((k_roi_fixed<true>), dim3 (2048), dim3 (256), 0, st, g, r, ptab, out);
    else
        hipLaunchKernelGGL ((k_roi_fixed<false>), dim3 (2048), dim3 (256), 0, st, g, r, ptab, out);
    return hipGetLastError ();
}

hipError_t
launch_pass1_positions (const KGeom& g, float* out, unsigned int* mismatch, hipStream_t st)
{
    const int rows = g.fdim[1] * g.fdim[2];
    if (g.aligned)
        hipLaunchKernelGGL ((k_pass1_positions<true>), dim3 ((rows + 255) / 256), dim3 (256), 0, st, g,
            out, mismatch);
    else
        hipLaunchKernelGGL ((k_pass1_positions<false>), dim3 ((rows + 255) / 256), dim3 (256), 0, st, g,
            out, mismatch);
    return hipGetLastError ();
}

// Score from the integer histograms, and the per-bin dS/dP table for
// pass 2.  Follows mi_hist_score_omp (src/register/bspline_mi.cxx:285-310)
// and the dS_dP expression of bspline_mi_pvi_8_dc_dv_dcos
// (src/register/bspline_mi.txx:96-101).  One workgroup of 1024 threads.
__global__ __launch_bounds__ (1024)
void k_mi_score (HGeom h, const unsigned long long* __restrict__ jhist,
    unsigned long long* __restrict__ fhist, float* __restrict__ table,
    double* __restrict__ out_score, int64_t* __restrict__ out_nvox,
    float* __restrict__ out_nvox_f, int64_t* __restrict__ hist_out, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double s_m[MAX_BINS_1D];
    __shared__ double s_f[MAX_BINS_1D];
    __shared__ double s_red[1024];
    __shared__ long long s_n;
    const int fb = h.fbins, mb = h.mbins, nj = fb * mb;
    const double unit = 1.0 / 4294967296.0;
    // the fixed histogram from the joint one: every voxel adds exactly 2^32
    // to the row of its fixed bin, so a row's sum is its count times 2^32
    for (int a = threadIdx.x; a < fb; a += blockDim.x) {
        unsigned long long row = 0;
        for (int b = 0; b < mb; b++) row += jhist[a * mb + b];
        fhist[a] = row >> 32;
        s_f[a] = (double) (row >> 32);
    }
    for (int b = threadIdx.x; b < mb; b += blockDim.x) {
        unsigned long long m = 0;
        for (int a = 0; a < fb; a++) m += jhist[a * mb + b];
        s_m[b] = (double) m * unit;
    }
    __syncthreads ();
    if (threadIdx.x == 0) {
        long long n = 0;
        for (int i = 0; i < fb; i++) n += (long long) s_f[i];
        s_n = n;
    }
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
        for (int a = threadIdx.x; a < fb; a += blockDim.x) hist_out[nj + a] = (int64_t) s_f[a];
    if (threadIdx.x == 0) {
        *out_score = score;
        *out_nvox = N;
        *out_nvox_f = nvf;
    }
}

// ---------------------------------------------------------------------------
// Analytic regularization (flavor c of Bspline_regularize,
// src/register/bspline_regularize_analytic.cxx:943-1007).
//
// Per region, with p_a the 64 knot coefficients of component a indexed
// 16 kz + 4 ky + kx (find_knots_3, :129-165), the reference sums p'Vp over
// 32 matrices V = w X (x) Y (x) Z (get_Vmatrix, :283-338: X acts on the
// z digit, Z on the x digit) and the elastic couplings.  Grouping equal
// Kronecker monomials, that is
//   S = sum_a p_a' O_a p_a + p_y' Exy p_x + p_z' Exz p_x + p_z' Eyz p_y,
//   g_x = 2 O_x p_x + Exy p_y + Exz p_z   (and cyclically, :501-601),
// and each operator is applied separably: a 4x4 contraction along the x
// digit, then y, then z -- 276 FMAs per knot for the O_a instead of the
// 64-term rows of 21 dense matrices.  One 64-thread workgroup per region,
// thread t = knot slot t; all in double, as the reference.  The slot values
// go slot-major to cond (as reg_sort_sets, :87-105) and a reduce sums each
// knot's 64 slots in slot order (reg_update_grad, :110-127).

// out(t) = sum_c M[digit(t)][c] in(t with that digit = c); digit 0 (x)
// has stride 1, 1 (y) stride 4, 2 (z) stride 16.
template <int D>
__device__ __forceinline__ double
reg_contract (const double* M, const double* in, int t)
{
    constexpr int s = D == 0 ? 1 : (D == 1 ? 4 : 16);
    const int r = (t / s) & 3;
    const double* b = in + (t - r * s);
    return M[4 * r + 0] * b[0] + M[4 * r + 1] * b[s] + M[4 * r + 2] * b[2 * s] + M[4 * r + 3] * b[3 * s];
}

__global__ __launch_bounds__ (64)
void k_regularize (KGeom g, RegOp op, const float* __restrict__ coeff, double* __restrict__ part,
    double* __restrict__ cond, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    // per component a: p, the x-digit contractions Z0..Z3, Z01, Z10, and
    // the staged sums of the y step
    __shared__ double s_p[3][64];
    __shared__ double s_z[3][6][64];
    __shared__ double s_w[3][4][64];    // W_i: what X_i contracts
    __shared__ double s_e[3][4][64];    // elastic: Y01 T0, Y10 T0, Y0 Z01p, Y0 Z10p (by component)
    __shared__ double s_f[3][2][64];    // elastic: Y10 Z01p, Y01 Z10p
    __shared__ double s_red[64];
    const int t = threadIdx.x;
    const int r = blockIdx.x;
    const int r0 = r % g.rdims[0], r1 = (r / g.rdims[0]) % g.rdims[1], r2 = r / (g.rdims[0] * g.rdims[1]);
    const int kx = t & 3, ky = (t >> 2) & 3, kz = t >> 4;
    const int knot = (r0 + kx) + g.cdims[0] * ((r1 + ky) + g.cdims[1] * (r2 + kz));
    double p[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        p[a] = (double) coeff[3 * knot + a];
        s_p[a][t] = p[a];
    }
    const bool elastic = op.e1 != 0.0 || op.e2 != 0.0;
    __syncthreads ();
    // x digit: Z_l, l = 0..3, and the elastic Z01, Z10
#pragma unroll
    for (int a = 0; a < 3; a++) {
#pragma unroll
        for (int l = 0; l < 4; l++) s_z[a][l][t] = reg_contract<0> (op.Z[l], s_p[a], t);
        if (elastic) {
            s_z[a][4][t] = reg_contract<0> (op.Z[5], s_p[a], t);     // Z01: orders (0, 1)
            s_z[a][5][t] = reg_contract<0> (op.Z[4], s_p[a], t);     // Z10: orders (1, 0)
        }
    }
    __syncthreads ();
    // y digit: U_jl = Y_j Z_l p, weighted into W_i (the monomials grouped by
    // their X order i; component a's diagonal elastic weights folded in)
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const double dX = a == 0 ? op.ed : op.eh;
        const double dY = a == 1 ? op.ed : op.eh;
        const double dZ = a == 2 ? op.ed : op.eh;
        const double* T0 = s_z[a][0];
        const double* T1 = s_z[a][1];
        const double* T2 = s_z[a][2];
        const double* T3 = s_z[a][3];
        const double u00 = reg_contract<1> (op.Y[0], T0, t);
        const double u10 = reg_contract<1> (op.Y[1], T0, t);
        const double u01 = reg_contract<1> (op.Y[0], T1, t);
        const double u20 = reg_contract<1> (op.Y[2], T0, t);
        const double u02 = reg_contract<1> (op.Y[0], T2, t);
        const double u11 = reg_contract<1> (op.Y[1], T1, t);
        const double u30 = reg_contract<1> (op.Y[3], T0, t);
        const double u03 = reg_contract<1> (op.Y[0], T3, t);
        const double u21 = reg_contract<1> (op.Y[2], T1, t);
        const double u12 = reg_contract<1> (op.Y[1], T2, t);
        s_w[a][0][t] = op.cp * (u20 + u02) + op.cpm * u11 + (op.dp + dY) * u10 + (op.dp + dZ) * u01
            + op.tdp * u00 + op.top * (u30 + u03 + u21 + u12);
        s_w[a][1][t] = op.cpm * (u10 + u01) + (op.dp + dX) * u00 + op.top * (u20 + u02 + u11);
        s_w[a][2][t] = op.cp * u00 + op.top * (u10 + u01);
        s_w[a][3][t] = op.top * u00;
        if (elastic) {
            s_e[a][0][t] = reg_contract<1> (op.Y[5], T0, t);          // Y01 Z0 p
            s_e[a][1][t] = reg_contract<1> (op.Y[4], T0, t);          // Y10 Z0 p
            s_e[a][2][t] = reg_contract<1> (op.Y[0], s_z[a][4], t);   // Y0 Z01 p
            s_e[a][3][t] = reg_contract<1> (op.Y[0], s_z[a][5], t);   // Y0 Z10 p
            s_f[a][0][t] = reg_contract<1> (op.Y[4], s_z[a][4], t);   // Y10 Z01 p
            s_f[a][1][t] = reg_contract<1> (op.Y[5], s_z[a][5], t);   // Y01 Z10 p
        }
    }
    __syncthreads ();
    // z digit: O_a p_a = sum_i X_i W_i, and the couplings
    //   Exy = e1 X10 Y01 Z0 + e2 X01 Y10 Z0,  Exz = e1 X10 Y0 Z01 + e2 X01 Y0 Z10,
    //   Eyz = e1 X0 Y10 Z01 + e2 X0 Y01 Z10
    double R[3];
#pragma unroll
    for (int a = 0; a < 3; a++)
        R[a] = reg_contract<2> (op.X[0], s_w[a][0], t) + reg_contract<2> (op.X[1], s_w[a][1], t)
            + reg_contract<2> (op.X[2], s_w[a][2], t) + reg_contract<2> (op.X[3], s_w[a][3], t);
    double Exy_x = 0, Exy_y = 0, Exz_x = 0, Exz_z = 0, Eyz_y = 0, Eyz_z = 0;
    if (elastic) {
        auto exy = [&] (int a) {
            return op.e1 * reg_contract<2> (op.X[4], s_e[a][0], t) + op.e2 * reg_contract<2> (op.X[5], s_e[a][1], t);
        };
        auto exz = [&] (int a) {
            return op.e1 * reg_contract<2> (op.X[4], s_e[a][2], t) + op.e2 * reg_contract<2> (op.X[5], s_e[a][3], t);
        };
        auto eyz = [&] (int a) {
            return op.e1 * reg_contract<2> (op.X[0], s_f[a][0], t) + op.e2 * reg_contract<2> (op.X[0], s_f[a][1], t);
        };
        Exy_x = exy (0); Exy_y = exy (1);
        Exz_x = exz (0); Exz_z = exz (2);
        Eyz_y = eyz (1); Eyz_z = eyz (2);
    }
    const double gx = 2.0 * R[0] + Exy_y + Exz_z;
    const double gy = 2.0 * R[1] + Exy_x + Eyz_z;
    const double gz = 2.0 * R[2] + Exz_x + Eyz_y;
    // slot t of this knot, slot-major: cond[t][3 knot + c]
    const size_t plane = (size_t) 3 * g.num_knots;
    cond[(size_t) t * plane + 3 * knot + 0] = gx;
    cond[(size_t) t * plane + 3 * knot + 1] = gy;
    cond[(size_t) t * plane + 3 * knot + 2] = gz;
    // the region's score, summed over the wave in a fixed order
    double sr = p[0] * R[0] + p[1] * R[1] + p[2] * R[2] + p[1] * Exy_x + p[2] * Exz_x + p[2] * Eyz_y;
    s_red[t] = sr;
    __syncthreads ();
    if (t == 0) {
        double a = 0.0;
        for (int i = 0; i < 64; i++) a += s_red[i];
        part[r] = a;
    }
}

// grad[i] = (accumulate ? grad[i] : 0) + sum over the 64 slots in slot
// order; workgroup 0 also sums the region scores in region order.
__global__ __launch_bounds__ (256)
void k_reg_reduce (const double* __restrict__ cond, const double* __restrict__ part, int nregions,
    int nc, double* __restrict__ score, double* __restrict__ grad, int accumulate,
    const int* __restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double s_q[4][64];
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    const double* c = cond + (size_t) (16 * q) * nc + (i < nc ? i : 0);
    double v[16];
#pragma unroll
    for (int m = 0; m < 16; m++) v[m] = c[(size_t) m * nc];
    double acc = 0.0;
#pragma unroll
    for (int m = 0; m < 16; m++) acc += v[m];
    s_q[q][lane] = acc;
    __syncthreads ();
    if (q == 0 && i < nc) {
        const double sum = ((s_q[0][lane] + s_q[1][lane]) + s_q[2][lane]) + s_q[3][lane];
        grad[i] = accumulate ? grad[i] + sum : sum;
    }
    if (blockIdx.x == 0) {
        __syncthreads ();
        // 256 strided partial sums, then those in thread order
        double a = 0.0;
        for (int r = threadIdx.x; r < nregions; r += 256) a += part[r];
        s_q[q][lane] = a;
        __syncthreads ();
        if (threadIdx.x == 0) {
            double sum = 0.0;
            for (int k = 0; k < 256; k++) sum += s_q[k >> 6][k & 63];
            *score = sum;
        }
    }
}

// ---------------------------------------------------------------------------
// small kernels

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

// Bin image: floor((v - offset) / delta) with the IEEE division, as
// Joint_histogram::add_pvi_8 (src/register/joint_histogram.cxx:458-482)
// computes it per use; with a V-optimal key table (keys entries), the bin
// of that key; done once per stage.
template <typename TI>
__global__ void
k_bin_image (const TI* __restrict__ img, uint8_t* __restrict__ out, int64_t n, float offset,
    float delta, int bins, const int* __restrict__ lut, int keys)
{
    for (int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        const float v = ld (img, (int) i);
        out[i] = (uint8_t) (lut ? lut[hist_bin (v, offset, delta, keys)] : hist_bin (v, offset, delta, bins));
    }
}

// The moving image as z-pairs (see launch_zpair); once per stage.  T2 is
// the record: float2 for float storage, uint32_t (two BF16) for BF16.
template <typename TI, typename T2>
__global__ void
k_zpair (const TI* __restrict__ m, T2* __restrict__ out, int64_t n, int64_t d01)
{
    for (int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x) {
        const TI lo = m[i], hi = i + d01 < n ? m[i + d01] : (TI) 0;
        if constexpr (std::is_same<TI, float>::value) out[i] = make_float2 (lo, hi);
        else out[i] = (uint32_t) lo | ((uint32_t) hi << 16);
    }
}

// The moving bin-quad image: record v holds the bins of voxels v, v + 1,
// v + d0, v + d0 + 1 (low byte first), 0 for a neighbour past the row or
// the slice (never read: li_clamp keeps the corner below dim - 1).
__global__ void
k_bin_quads (const uint8_t* __restrict__ b, uint32_t* __restrict__ out, int64_t n, int d0, int d1)
{
    for (int64_t v = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; v < n;
         v += (int64_t) gridDim.x * blockDim.x) {
        const int64_t x = v % d0, y = (v / d0) % d1;
        const bool xr = x + 1 < d0, yr = y + 1 < d1;
        uint32_t q = b[v];
        if (xr) q |= (uint32_t) b[v + 1] << 8;
        if (yr) q |= (uint32_t) b[v + d0] << 16;
        if (xr && yr) q |= (uint32_t) b[v + d0 + 1] << 24;
        out[v] = q;
    }
}

__global__ void
k_zero_u64 (unsigned long long* __restrict__ p, int n, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 0ull;
}

__global__ void
k_fill_f32 (float* __restrict__ p, int64_t n, float v)
{
    for (int64_t i = blockIdx.x * (int64_t) blockDim.x + threadIdx.x; i < n;
         i += (int64_t) gridDim.x * blockDim.x)
        p[i] = v;
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
tile_lds_bytes (const KGeom& g, int vox_per_thread)
{
    // the carve-up of the tile kernel
    const int vx = g.vpr[0], vy = g.vpr[1], vz = g.vpr[2];
    const size_t f = 4 * vx + 4 * vy + 8 * vz + 192 + tile_col_floats (vx, vy, vz)
        + tile_dv_floats (vx, tile_slots (vx * vy * vz, vox_per_thread));
    // + the kernel's static arrays
    return f * sizeof (float) + 16 * sizeof (double) + TILE_STATIC_LDS;
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

// What a tile-kernel launch reads besides the plan.
struct TileArgs {
    const float* coeff0;    // pair 0's coefficients (single-pair calls), or null
    MiArgs mi;              // MI pass 2
    float* vf;              // pair 0's displacement field (KIND 4 out, KIND 5 in), or
                            // the warped image (KIND 6, 7 out)
    float fill = 0.f;       // KIND 6, 7: the value where a voxel leaves the moving image
    const int* skip = nullptr;  // EvalOut::skip
};

// Persistent grid: as many workgroups as fit on the device at once
// (occupancy from the kernel's registers and this launch's LDS; at most
// tune.wg_per_cu per CU when set), a multiple of 8 so every XCD gets the
// same count, and no more than an XCD has tiles.
template <int KIND, typename TI, int V, int G, bool BATCH>
static int
launch_tiles_vgb (const Plan& plan, const Tune& tune, size_t lds, hipStream_t st,
    const TileArgs& ta)
{
    static std::mutex mu;
    static size_t cached_lds = 0;
    static int cached_occ = 0;
    int occ;
    {
        std::lock_guard<std::mutex> lock (mu);
        if (cached_lds != lds) {
            int n = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor (&n, k_tiles<KIND, TI, V, G, BATCH, false>,
                    TILE_THREADS, lds) != hipSuccess || n < 1)
                n = 1;
            cached_lds = lds;
            cached_occ = n;
        }
        occ = cached_occ;
    }
    if (tune.wg_per_cu > 0) occ = std::min (occ, tune.wg_per_cu);
    TileQueue q;
    q.total_tiles = plan.total_tiles;
    q.tiles_per_xcd = (plan.total_tiles + 7) / 8;
    q.npairs = plan.npairs;
    q.pad = 0;
    q.ctr = plan.tile_ctr;
    const int cus = tune.num_cus > 0 ? tune.num_cus : 8;
    // as many workgroups as fit (evening out the rounds instead -- 184
    // workgroups an XCD for 1099 tiles at 256x256x128 rather than 192, of
    // which 54 sit out the last of six rounds -- was slower: 51.3 against
    // 48.4 us; the concurrency of the full rounds is worth more)
    const int per_xcd = std::max (1, std::min (q.tiles_per_xcd, occ * cus / 8));
    // counted tiles on long runs, for the launches built with them (a batch,
    // MI pass 2); the occupancy above is the fixed-stride build's
    constexpr bool DYN_BUILT = BATCH || (KIND & 15) == 2;
    const bool dyn = DYN_BUILT && plan.tile_ctr && q.tiles_per_xcd >= TILE_DYNAMIC_ROUNDS * per_xcd;
    if constexpr (DYN_BUILT) {
        if (dyn) {
            hipLaunchKernelGGL ((k_tiles<KIND, TI, V, G, BATCH, true>), dim3 (8 * per_xcd), dim3 (TILE_THREADS),
                lds, st, plan.d_pairs, q, ta.coeff0, plan.blut, plan.part_score, plan.part_count, ta.mi,
                ta.vf, ta.fill, ta.skip);
            return -8 * per_xcd;
        }
    }
    hipLaunchKernelGGL ((k_tiles<KIND, TI, V, G, BATCH, false>), dim3 (8 * per_xcd), dim3 (TILE_THREADS),
        lds, st, plan.d_pairs, q, ta.coeff0, plan.blut, plan.part_score, plan.part_count, ta.mi,
        ta.vf, ta.fill, ta.skip);
    return 8 * per_xcd;
}

// Batched launches exist for the two MSE gradient kinds; everything else
// runs on single-pair plans.
template <int KIND, typename TI, int V, int G>
static int
launch_tiles_vg (const Plan& plan, const Tune& tune, size_t lds, hipStream_t st,
    const TileArgs& ta)
{
    if (KIND < 2 && plan.npairs > 1)
        return launch_tiles_vgb<KIND, TI, V, G, (KIND < 2)> (plan, tune, lds, st, ta);
    return launch_tiles_vgb<KIND, TI, V, G, false> (plan, tune, lds, st, ta);
}

template <int KIND, typename TI, int V>
static int
launch_tiles_v (const Plan& plan, const Tune& tune, int vpt, size_t extra, hipStream_t st,
    const TileArgs& ta)
{
    const KGeom& g = plan.g0;
    // dynamic LDS: the carve-up (+ the MI table), without the static arrays
    constexpr size_t stat = TILE_STATIC_LDS;
    if (vpt == 1)
        return launch_tiles_vg<KIND, TI, V, 1> (plan, tune, tile_lds_bytes (g, 1) - stat + extra,
            st, ta);
    if (vpt == 2)
        return launch_tiles_vg<KIND, TI, V, 2> (plan, tune, tile_lds_bytes (g, 2) - stat + extra,
            st, ta);
    return launch_tiles_vg<KIND, TI, V, 4> (plan, tune, tile_lds_bytes (g, 4) - stat + extra,
        st, ta);
}

// The tile kernel is specialized at compile time for the isotropic tile
// edges 8, 10, 12 and 15 voxels; other vox_per_rgn run the generic build.
// vox_per_thread falls back to the largest value whose dc/dv slots fit in
// LDS (1 always fits: the context checked it at creation).  Returns the
// grid size, the number of per-workgroup partials per pair, negated when
// the launch took counted tiles (its score partials are then per tile).
// (The dynamic LDS passed at launch excludes the kernel's static arrays,
// which tile_lds_bytes counts for the fit.)
template <int KIND, typename TI>
static int
launch_tiles (const Plan& plan, const Tune& tune, size_t extra, hipStream_t st,
    const TileArgs& ta)
{
    const KGeom& g = plan.g0;
    int vpt = tune.vox_per_thread;
    while (vpt > 1 && tile_lds_bytes (g, vpt) + extra > tune.lds_limit) vpt >>= 1;
    const int v = g.vpr[0];
    const bool iso = g.vpr[1] == v && g.vpr[2] == v;
#define PLM_TILES(VV) return launch_tiles_v<KIND, TI, VV> (plan, tune, vpt, extra, st, ta)
    if (iso && v == 8) PLM_TILES (8);
    if (iso && v == 10) PLM_TILES (10);
    if (iso && v == 12) PLM_TILES (12);
    if (iso && v == 15) PLM_TILES (15);
    PLM_TILES (0);
#undef PLM_TILES
}

template <int KIND>
static int
launch_tiles_storage (const Plan& plan, const Tune& tune, int storage, hipStream_t st,
    const TileArgs& ta)
{
    if (storage == 0) return launch_tiles<KIND, float> (plan, tune, 0, st, ta);
    return launch_tiles<KIND, uint16_t> (plan, tune, 0, st, ta);
}

hipError_t
launch_mse (const Plan& plan, const Tune& tune, int mode, int storage, const float* coeff0,
    const EvalOut& out0, hipStream_t st, hipEvent_t* ev, const ReduceHook* hook)
{
    TileArgs ta {coeff0, MiArgs {}, nullptr};
    ta.skip = out0.skip;
    hipError_t e = mark (ev, 0, st);
    if (e != hipSuccess) return e;
    int nparts;
    if (mode == MSE_REF_NN) nparts = launch_tiles_storage<0> (plan, tune, storage, st, ta);
    else if (mode == MSE_EXACT) nparts = launch_tiles_storage<1> (plan, tune, storage, st, ta);
    else nparts = launch_tiles_storage<3> (plan, tune, storage, st, ta);
    LAUNCH_CHECK ();
    if ((e = mark (ev, 1, st)) != hipSuccess) return e;
    // the score alone needs one reduce workgroup per pair (single-pair plans)
    const bool grad = mode != MSE_SCORE_ONLY;
    if (hook && grad && plan.npairs == 1) {
        if ((e = hook->launch (hook->user, plan, nparts < 0 ? -nparts : nparts, nparts < 0 ? 1 : 0, out0, st))
            != hipSuccess)
            return e;
    } else {
        hipLaunchKernelGGL (k_reduce<true>, dim3 (grad ? plan.reduce_blocks : 1), dim3 (REDUCE_THREADS),
            0, st, plan.d_pairs, plan.npairs, plan.part_score, plan.part_count, nparts < 0 ? -nparts : nparts,
            nparts < 0 ? 1 : 0, out0, grad ? 1 : 0);
        LAUNCH_CHECK ();
    }
    return mark (ev, 2, st);
}

// The MSE evaluation of a single-pair float plan whose pair has no z-pair
// image (the corners come from the moving image), at vox_per_thread 2.
template <int KIND>
static int
launch_tiles_plain (const Plan& plan, const Tune& tune, hipStream_t st, const TileArgs& ta)
{
    const KGeom& g = plan.g0;
    constexpr size_t stat = TILE_STATIC_LDS;
    const size_t lds = tile_lds_bytes (g, 2) - stat;
    const int v = g.vpr[0];
    const bool iso = g.vpr[1] == v && g.vpr[2] == v;
    constexpr int K = KIND | TILE_PLAIN;
    if (iso && v == 8) return launch_tiles_vgb<K, float, 8, 2, false> (plan, tune, lds, st, ta);
    if (iso && v == 10) return launch_tiles_vgb<K, float, 10, 2, false> (plan, tune, lds, st, ta);
    if (iso && v == 12) return launch_tiles_vgb<K, float, 12, 2, false> (plan, tune, lds, st, ta);
    if (iso && v == 15) return launch_tiles_vgb<K, float, 15, 2, false> (plan, tune, lds, st, ta);
    return launch_tiles_vgb<K, float, 0, 2, false> (plan, tune, lds, st, ta);
}

hipError_t
launch_mse_plain (const Plan& plan, const Tune& tune, int mode, const float* coeff0,
    const EvalOut& out0, hipStream_t st)
{
    const TileArgs ta {coeff0, MiArgs {}, nullptr};
    int nparts;
    if (mode == MSE_REF_NN) nparts = launch_tiles_plain<0> (plan, tune, st, ta);
    else if (mode == MSE_EXACT) nparts = launch_tiles_plain<1> (plan, tune, st, ta);
    else nparts = launch_tiles_plain<3> (plan, tune, st, ta);
    LAUNCH_CHECK ();
    const bool grad = mode != MSE_SCORE_ONLY;
    hipLaunchKernelGGL (k_reduce<true>, dim3 (grad ? plan.reduce_blocks : 1), dim3 (REDUCE_THREADS),
        0, st, plan.d_pairs, plan.npairs, plan.part_score, plan.part_count, nparts < 0 ? -nparts : nparts,
        nparts < 0 ? 1 : 0, out0, grad ? 1 : 0);
    return hipGetLastError ();
}

hipError_t
launch_vf (const Plan& plan, const Tune& tune, const float* coeff0, float* vf, hipStream_t st)
{
    const TileArgs ta {coeff0, MiArgs {}, vf};
    // the field covers the whole fixed image; voxels outside the ROI stay 0
    const KGeom& g = plan.g0;
    hipError_t e = launch_zero (vf, 3 * (int64_t) g.fdim[0] * g.fdim[1] * g.fdim[2], st);
    if (e != hipSuccess) return e;
    launch_tiles<4, float> (plan, tune, 0, st, ta);
    return hipGetLastError ();
}

hipError_t
launch_mse_from_vf (const Plan& plan, const Tune& tune, int storage, const float* vf,
    const EvalOut& out0, hipStream_t st, hipEvent_t* ev)
{
    const TileArgs ta {nullptr, MiArgs {}, const_cast<float*> (vf)};
    hipError_t e = mark (ev, 0, st);
    if (e != hipSuccess) return e;
    const int nparts = launch_tiles_storage<5> (plan, tune, storage, st, ta);
    LAUNCH_CHECK ();
    if ((e = mark (ev, 1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL (k_reduce<true>, dim3 (plan.reduce_blocks), dim3 (REDUCE_THREADS), 0, st,
        plan.d_pairs, plan.npairs, plan.part_score, plan.part_count, nparts < 0 ? -nparts : nparts,
        nparts < 0 ? 1 : 0, out0, 1);
    LAUNCH_CHECK ();
    return mark (ev, 2, st);
}

hipError_t
launch_mi (const Plan& plan, const Tune& tune, const HGeom& h, int hist_blocks,
    const uint8_t* fbin, const uint32_t* mquad, const float* coeff, unsigned long long* jhist,
    unsigned long long* fhist, float* table, float* nvox_f, const EvalOut& out0,
    int64_t* hist_out, const float* ptab, const uint8_t* fmask1, const RoiVol& mroi,
    hipStream_t st, hipEvent_t* ev)
{
    const KGeom& g = plan.g0;
    const int nbins = h.fbins * h.mbins + h.fbins;
    hipError_t e = mark (ev, 0, st);
    if (e != hipSuccess) return e;
    // jhist and fhist are one allocation: joint bins then fixed bins.
    hipLaunchKernelGGL (k_zero_u64, dim3 ((nbins + 255) / 256), dim3 (256), 0, st, jhist, nbins,
        out0.skip);
    LAUNCH_CHECK ();
    const size_t lds1 = sizeof (unsigned long long) * nbins
        + sizeof (float) * 4 * (g.vpr[0] + g.vpr[1] + g.vpr[2]);
    const size_t table_bytes = sizeof (float) * h.fbins * h.mbins;
    if ((e = mark (ev, 1, st)) != hipSuccess) return e;
    if (g.aligned)
        hipLaunchKernelGGL ((k_mi_hist<true>), dim3 (hist_blocks), dim3 (TILE_THREADS), lds1, st,
            g, h, fbin, mquad, coeff, plan.blut, jhist, fhist, ptab, fmask1, mroi, out0.skip);
    else
        hipLaunchKernelGGL ((k_mi_hist<false>), dim3 (hist_blocks), dim3 (TILE_THREADS), lds1, st,
            g, h, fbin, mquad, coeff, plan.blut, jhist, fhist, ptab, fmask1, mroi, out0.skip);
    LAUNCH_CHECK ();
    if ((e = mark (ev, 2, st)) != hipSuccess) return e;
    hipLaunchKernelGGL (k_mi_score, dim3 (1), dim3 (1024), 0, st, h, jhist, fhist, table,
        out0.score, out0.nvox, nvox_f, hist_out, out0.skip);
    LAUNCH_CHECK ();
    if ((e = mark (ev, 3, st)) != hipSuccess) return e;
    TileArgs ta {coeff, MiArgs {h, table, nvox_f, fbin, mquad}, nullptr};
    ta.skip = out0.skip;
    // pass 2 reads only bin images; the image type parameter is unused
    launch_tiles<2, float> (plan, tune, table_bytes, st, ta);
    LAUNCH_CHECK ();
    if ((e = mark (ev, 4, st)) != hipSuccess) return e;
    hipLaunchKernelGGL (k_reduce<false>, dim3 (plan.reduce_blocks), dim3 (REDUCE_THREADS), 0, st,
        plan.d_pairs, plan.npairs, nullptr, nullptr, 0, 0, out0, 1);
    LAUNCH_CHECK ();
    return mark (ev, 5, st);
}

hipError_t
launch_bin_image (const void* img, int storage, uint8_t* out, int64_t n, float offset, float delta,
    int bins, hipStream_t st, const int* lut, int keys)
{
    if (storage == 0)
        hipLaunchKernelGGL ((k_bin_image<float>), dim3 (1024), dim3 (256), 0, st,
            (const float*) img, out, n, offset, delta, bins, lut, keys);
    else
        hipLaunchKernelGGL ((k_bin_image<uint16_t>), dim3 (1024), dim3 (256), 0, st,
            (const uint16_t*) img, out, n, offset, delta, bins, lut, keys);
    return hipGetLastError ();
}

hipError_t
launch_zpair (const void* m, int storage, void* out, int64_t n, int64_t d01, hipStream_t st)
{
    if (storage == 0)
        hipLaunchKernelGGL ((k_zpair<float, float2>), dim3 (1024), dim3 (256), 0, st,
            (const float*) m, (float2*) out, n, d01);
    else
        hipLaunchKernelGGL ((k_zpair<uint16_t, uint32_t>), dim3 (1024), dim3 (256), 0, st,
            (const uint16_t*) m, (uint32_t*) out, n, d01);
    return hipGetLastError ();
}

hipError_t
launch_bin_quads (const uint8_t* bins, uint32_t* out, const int dim[3], hipStream_t st)
{
    const int64_t n = (int64_t) dim[0] * dim[1] * dim[2];
    hipLaunchKernelGGL (k_bin_quads, dim3 (1024), dim3 (256), 0, st, bins, out, n, dim[0], dim[1]);
    return hipGetLastError ();
}

hipError_t
launch_zero (float* p, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL (k_fill_f32, dim3 (1024), dim3 (256), 0, st, p, n, 0.f);
    return hipGetLastError ();
}

hipError_t
launch_regularize (const KGeom& g, const RegOp& op, const float* coeff, double* part, double* cond,
    double* score, double* grad, int accumulate, hipStream_t st, const int* skip)
{
    hipLaunchKernelGGL (k_regularize, dim3 (g.num_tiles), dim3 (64), 0, st, g, op, coeff, part, cond,
        skip);
    LAUNCH_CHECK ();
    const int nc = 3 * g.num_knots;
    hipLaunchKernelGGL (k_reg_reduce, dim3 ((nc + 63) / 64), dim3 (256), 0, st, cond, part,
        g.num_tiles, nc, score, grad, accumulate, skip);
    return hipGetLastError ();
}

hipError_t
launch_warp (const Plan& plan, const Tune& tune, int storage, const float* coeff0, float* out,
    float fill, int linear, hipStream_t st)
{
    const KGeom& g = plan.g0;
    // voxels outside the ROI keep the fill value; the tile kernel writes
    // every ROI voxel
    bool whole = true;
    for (int a = 0; a < 3; a++)
        whole = whole && g.roi_off[a] == 0 && g.roi_end[a] == g.fdim[a];
    if (!whole) {
        hipLaunchKernelGGL (k_fill_f32, dim3 (1024), dim3 (256), 0, st, out,
            (int64_t) g.fdim[0] * g.fdim[1] * g.fdim[2], fill);
        LAUNCH_CHECK ();
    }
    const TileArgs ta {coeff0, MiArgs {}, out, fill};
    if (linear) launch_tiles_storage<6> (plan, tune, storage, st, ta);
    else launch_tiles_storage<7> (plan, tune, storage, st, ta);
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
