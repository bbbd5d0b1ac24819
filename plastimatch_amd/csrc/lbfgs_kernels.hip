// Vector kernels of the device-resident L-BFGS: the vector work of
// liblbfgs's lbfgs() (libs/liblbfgs-1.9/lib/lbfgs.c:449-609) and of its
// More-Thuente trials (:903-906), in double, the MSE gradient reduce with
// that work folded in, and the pass that runs the scalar logic on the
// device.
//
// Built with -ffp-contract=off: each element update rounds its product and
// its sum separately, as liblbfgs's vecadd / vecdiff do.

#include <float.h>
#include <stdint.h>

#include "bspline_kernels.h"
#include "lbfgs_kernels.h"
#include "lbfgs_logic.h"

namespace plmhip {

namespace {

// Workgroup sum of one per-thread value: a butterfly within each wave,
// then the four wave sums as (w0 + w1) + (w2 + w3); valid in thread 0.
__device__ __forceinline__ double
block_sum1 (double v)
{
    __shared__ double red[LB_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor (v, o, 64);
    if (lane == 0) red[w] = v;
    __syncthreads ();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// Workgroup sums of NC per-thread values into out[c * nb + block]:
// components go through LDS sixteen at a time -- thread (c, l) adds
// threads l, l + 16, ..., l + 240 of component c in turn, then thread c
// adds its sixteen partial sums in turn.  Two LDS round trips per chunk
// instead of a shuffle butterfly per component.
template <int NC>
__device__ __forceinline__ void
block_partials (const double (&v)[NC], double* __restrict__ out, int nb)
{
    __shared__ double buf[16][LB_THREADS];
    __shared__ double mid[16][16];
    const int t = threadIdx.x;
#pragma unroll
    for (int c0 = 0; c0 < NC; c0 += 16) {
#pragma unroll
        for (int c = 0; c < 16; ++c)
            if (c0 + c < NC) buf[c][t] = v[c0 + c];
        __syncthreads ();
        {
            const int c = t >> 4, l = t & 15;
            double s = 0.0;
#pragma unroll
            for (int q = 0; q < LB_THREADS / 16; ++q) s += buf[c][l + 16 * q];
            mid[c][l] = s;
        }
        __syncthreads ();
        if (t < 16 && c0 + t < NC) {
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < 16; ++l) s += mid[t][l];
            out[(c0 + t) * nb + blockIdx.x] = s;
        }
        __syncthreads ();
    }
}

// The point of the next evaluation (LbAct::kind): 0 the start point x0; 1
// a new direction d = -H g from the coefficients the scalar logic solved
// for, its first trial xp + stp d and the partials of g.d; 2 a trial
// xp + stp d along the current d.  x and the float coefficients are
// written.
template <int MMAX>
__global__ void __launch_bounds__ (LB_THREADS)
k_lb_next (const LbAct* __restrict__ ap, LbAct av, LbBufs b)
{
    const LbAct* act = ap ? ap : &av;
    if (act->done) return;
    const int kind = act->kind;
    const double stp = act->stp;
    double* __restrict__ x = b.X[act->xi];
    const double* __restrict__ xp = b.X[act->xpi];
    const double* __restrict__ g = b.G[act->gpi];
    const int nd = act->dir_b;
    const double gamma = act->gamma;
    const int64_t n = b.n;
    double v = 0.0;
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        double xe;
        if (kind == 0) {
            xe = b.x0[e];
        } else if (kind == 1) {
            const double ge = g[e];
            double de;
            if (nd == 0) {
                de = -ge;
            } else {
                // nd <= MMAX (the run's m): the bound is a constant, so the
                // action's slots and coefficients and every history load
                // are issued before the sums start
                double t1 = 0.0, t2 = 0.0;
#pragma unroll
                for (int i = 0; i < MMAX; ++i) {
                    if (i < nd) {
                        const int64_t off = (int64_t) act->slot[i] * n + e;
                        t1 = t1 + act->ca[i] * b.S[off];
                        t2 = t2 + act->cu[i] * b.Y[off];
                    }
                }
                de = -gamma * ge - t1 + gamma * t2;
            }
            b.d[e] = de;
            xe = xp[e] + stp * de;
            v += ge * de;
        } else {
            xe = xp[e] + stp * b.d[e];
        }
        x[e] = xe;
        b.coeff[e] = (float) xe;
    }
    if (kind == 1) {
        const double s = block_sum1 (v);
        if (threadIdx.x == 0) b.slope[blockIdx.x] = s;
    }
}

// The post pass's work on element e of gradient value ge: ge kept in the
// trial's g buffer, s = x - xp and y = g - gp written into the would-be
// pair's slot pe, and the element's terms added to every post component
// (LbAct::post_bp <= BPMAX earlier pairs; the start point has no pair and
// takes g for d).
template <int BPMAX>
struct PostElem {
    static constexpr int NC = 7 + 5 * BPMAX;
    const double* __restrict__ x;
    const double* __restrict__ xp;
    double* __restrict__ gk;
    const double* __restrict__ gp;
    const double* __restrict__ d;
    int pe, bp;
    bool pair;
    int slot[BPMAX > 0 ? BPMAX : 1];

    __device__ __forceinline__
    PostElem (const LbAct* act, const LbBufs& b)
    {
        pe = act->post_e;
        bp = act->post_bp;
        pair = pe >= 0;
        x = b.X[act->xi];
        xp = b.X[act->xpi];
        gk = b.G[act->gi];
        gp = b.G[act->gpi];
        d = act->kind == 0 ? nullptr : b.d;
        const int m = b.m;
#pragma unroll
        for (int i = 1; i <= BPMAX; ++i) slot[i - 1] = (pe - i + m) % m;
    }

    // The element's inputs, none of which the pass writes (the earlier
    // pairs' slots are never pe), so they may be fetched before its
    // gradient value is known.
    struct In {
        double xe, xpe, gpe, de;
        double sj[BPMAX > 0 ? BPMAX : 1], yj[BPMAX > 0 ? BPMAX : 1];
    };

    __device__ __forceinline__ void
    load (const LbBufs& b, int64_t e, In& w) const
    {
        const int64_t n = b.n;
        w.xe = x[e];
        w.de = d ? d[e] : 0.0;
        if (pair) {
            w.xpe = xp[e];
            w.gpe = gp[e];
#pragma unroll
            for (int i = 0; i < BPMAX; ++i) {
                if (i < bp) {
                    w.sj[i] = b.S[(int64_t) slot[i] * n + e];
                    w.yj[i] = b.Y[(int64_t) slot[i] * n + e];
                }
            }
        }
    }

    __device__ __forceinline__ void
    add (const LbBufs& b, int64_t e, double ge, const In& w, double (&v)[NC]) const
    {
        const int64_t n = b.n;
        const double xe = w.xe;
        gk[e] = ge;
        const double de = d ? w.de : ge;
        v[0] += ge * de;
        v[1] += xe * xe;
        v[2] += ge * ge;
        if (pair) {
            const double se = xe - w.xpe;
            const double ye = ge - w.gpe;
            b.S[(int64_t) pe * n + e] = se;
            b.Y[(int64_t) pe * n + e] = ye;
            v[3] += se * ye;
            v[4] += ye * ye;
            v[5] += se * ge;
            v[6] += ye * ge;
#pragma unroll
            for (int i = 0; i < BPMAX; ++i) {
                if (i < bp) {
                    const double sj = w.sj[i];
                    const double yj = w.yj[i];
                    v[7 + 5 * i + 0] += se * yj;
                    v[7 + 5 * i + 1] += sj * ye;
                    v[7 + 5 * i + 2] += ye * yj;
                    v[7 + 5 * i + 3] += sj * ge;
                    v[7 + 5 * i + 4] += yj * ge;
                }
            }
        }
    }
};

// After the evaluation: the gradient it wrote (geval) through PostElem,
// then the workgroup's partial of every post component.
template <int BPMAX>
__global__ void __launch_bounds__ (LB_THREADS)
k_lb_post (const LbAct* __restrict__ ap, LbAct av, LbBufs b)
{
    const LbAct* act = ap ? ap : &av;
    if (act->done) return;
    using PE = PostElem<BPMAX>;
    const PE pe (act, b);
    double v[PE::NC];
#pragma unroll
    for (int c = 0; c < PE::NC; ++c) v[c] = 0.0;
    const int64_t n = b.n;
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        typename PE::In w;
        pe.load (b, e, w);
        pe.add (b, e, b.geval[e], w, v);
    }
    block_partials<PE::NC> (v, b.part, gridDim.x);
}

__device__ __forceinline__ double
wave_sum_d (double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor (v, o, 64);
    return v;
}

__device__ __forceinline__ long long
wave_sum_i64 (long long v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor (v, o, 64);
    return v;
}

// The MSE gradient reduce and the post pass in one launch, for a single-pair
// MSE objective without regularization whose n elements take one thread
// each in k_lb_post's grid (n <= lb_blocks (n) * 256).  Thread e forms
// gradient component e as k_reduce<true> does -- its 64 condense slots as
// four runs of 16 summed in FP64 in slot order, the runs added in order,
// times 2/N -- and hands it to PostElem as k_lb_post would from geval; N
// and the score are k_reduce's too (every workgroup sums the overlap
// counts, the first one the score partials: per tile in tile order when
// the tile kernel took counted tiles, else per workgroup, thread t adding
// t, t + 256, ... and the four wave sums in order).  The gradient, score,
// N and every post partial are the bits of k_reduce followed by k_lb_post.
template <int BPMAX>
__global__ void __launch_bounds__ (LB_THREADS)
k_lb_reduce_post (const PairDev* __restrict__ pairs, const double* __restrict__ part_score,
    const int* __restrict__ part_count, int nparts, int tile_scores, EvalOut out0,
    const LbAct* __restrict__ ap, LbAct av, LbBufs b)
{
    const LbAct* act = ap ? ap : &av;
    if (act->done) return;
    __shared__ long long s_n[LB_THREADS / 64];
    __shared__ double s_s[LB_THREADS / 64];
    using PE = PostElem<BPMAX>;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const PairDev& pd = pairs[0];
    const int64_t n = b.n;
    const int64_t e = (int64_t) blockIdx.x * LB_THREADS + t;
    const bool live = e < n;
    // the 64 slot loads first, unconditional (the index stays in range)
    const float* c = pd.cond + (live ? e : 0);
    float sv[64];
#pragma unroll
    for (int m = 0; m < 64; m++) sv[m] = c[(int64_t) m * n];
    // and the post pass's inputs of the element, in the same round
    const PE pe (act, b);
    typename PE::In w;
    pe.load (b, live ? e : 0, w);
    // N: the overlap counts, 16 bytes a lane where aligned
    long long cnt = 0;
    int t4 = 0;
    if ((reinterpret_cast<uintptr_t> (part_count) & 15) == 0) {
        const int4* pc4 = reinterpret_cast<const int4*> (part_count);
        const int n4 = nparts >> 2;
        for (int t0 = 0; t0 < n4; t0 += 2 * LB_THREADS) {
            const int ta = t0 + t, tb = ta + LB_THREADS;
            const int4 a = ta < n4 ? pc4[ta] : make_int4 (0, 0, 0, 0);
            const int4 q = tb < n4 ? pc4[tb] : make_int4 (0, 0, 0, 0);
            cnt += ((long long) a.x + a.y) + ((long long) a.z + a.w)
                + ((long long) q.x + q.y) + ((long long) q.z + q.w);
        }
        t4 = 4 * n4;
    }
    for (int k = t4 + t; k < nparts; k += LB_THREADS) cnt += part_count[k];
    double sc = 0.0;
    if (blockIdx.x == 0) {
        const double* ps = tile_scores ? part_score + pd.tile0 : part_score;
        const int np = tile_scores ? pd.g.num_tiles : nparts;
        for (int k = t; k < np; k += LB_THREADS) sc += ps[k];
    }
    cnt = wave_sum_i64 (cnt);
    sc = wave_sum_d (sc);
    if (lane == 0) { s_n[wave] = cnt; s_s[wave] = sc; }
    __syncthreads ();
    long long Nv = 0;
    double Sv = 0.0;
    for (int w = 0; w < LB_THREADS / 64; w++) { Nv += s_n[w]; Sv += s_s[w]; }
    const double scale = Nv >= 1 ? 2.0 / (double) Nv : 0.0;
    if (blockIdx.x == 0 && t == 0) {
        *out0.score = Nv >= 1 ? Sv / (double) Nv : (double) FLT_MAX;
        *out0.nvox = Nv;
    }
    double a[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < 16; m++) acc += (double) sv[16 * q + m];
        a[q] = acc;
    }
    const double ge = (((a[0] + a[1]) + a[2]) + a[3]) * scale;
    double v[PE::NC];
#pragma unroll
    for (int k = 0; k < PE::NC; ++k) v[k] = 0.0;
    if (live) pe.add (b, e, ge, w, v);
    block_partials<PE::NC> (v, b.part, gridDim.x);
}

// One workgroup: the totals -- the score slots, the slope (sum of the nb
// partials of k_lb_next) and every component LbAct::ncomp (sums of the nb
// partials of k_lb_post), thread (k, l) of a group of sixteen totals adding
// partials l, l + 16, ... in turn, then the sixteen in turn.  Without a
// machine the totals go to host; with one, its state is staged in LDS,
// thread 0 advances it over the totals, and the next action (or the run's
// end) goes back to device memory, with a note to the polling host.
__global__ void __launch_bounds__ (LB_THREADS)
k_lb_finish (const LbAct* __restrict__ ap, LbAct av, LbMachine* __restrict__ machine, LbBufs b,
    double* __restrict__ host, LbFlag* __restrict__ flag)
{
    const LbAct* act = ap ? ap : &av;
    if (act->done) return;
    __shared__ double mid[48][16];
    __shared__ double tot[4 + 7 + 5 * (LB_MAX_M - 1)];
    __shared__ LbMachine ms;
    const int t = threadIdx.x;
    const int nb = b.nb;
    const int total = act->ncomp + 1;
    // the machine's words are fetched first, so their latency hides under
    // the sums (it is 8-byte aligned)
    constexpr int W = (int) (sizeof (LbMachine) / 8);
    constexpr int WT = (W + LB_THREADS - 1) / LB_THREADS;
    static_assert (sizeof (LbMachine) % 8 == 0, "LbMachine size");
    uint64_t* ws = reinterpret_cast<uint64_t*> (&ms);
    uint64_t* wg = reinterpret_cast<uint64_t*> (machine);
    uint64_t mw[WT];
    if (machine) {
#pragma unroll
        for (int i = 0; i < WT; i++)
            if (t + i * LB_THREADS < W) mw[i] = wg[t + i * LB_THREADS];
    }
    if (t < 3) tot[t] = b.score[t];
    // RG groups of sixteen totals a round (48: every total of a history of
    // up to 9 pairs in one round), all of a round's loads in flight together
    constexpr int RG = 3;
    for (int k0 = 0; k0 < total; k0 += 16 * RG) {
        const int l = t & 15;
        // every partial of this thread loaded before any is added (nb <=
        // LB_MAX_BLOCKS), then added in q order; the absent ones add +0.0,
        // which leaves a sum that started at +0.0 unchanged
        double pv[RG][LB_MAX_BLOCKS / 16];
#pragma unroll
        for (int r = 0; r < RG; r++) {
            const int k = k0 + 16 * r + (t >> 4);
            const double* p = k == 0 ? b.slope : b.part + (int64_t) (k - 1) * nb;
#pragma unroll
            for (int i = 0; i < LB_MAX_BLOCKS / 16; i++) {
                const int q = l + 16 * i;
                pv[r][i] = (k < total && q < nb) ? p[q] : 0.0;
            }
        }
#pragma unroll
        for (int r = 0; r < RG; r++) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < LB_MAX_BLOCKS / 16; i++) s += pv[r][i];
            mid[16 * r + (t >> 4)][l] = s;
        }
        __syncthreads ();
        if (t < 16 * RG && k0 + t < total) {
            double u = 0.0;
#pragma unroll
            for (int j = 0; j < 16; ++j) u += mid[t][j];
            tot[3 + k0 + t] = u;
        }
        __syncthreads ();
    }
    if (!machine) {
        for (int i = t; i < 3 + total; i += LB_THREADS) host[i] = tot[i];
        return;
    }
    // the machine through LDS
#pragma unroll
    for (int i = 0; i < WT; i++)
        if (t + i * LB_THREADS < W) ws[t + i * LB_THREADS] = mw[i];
    __syncthreads ();
    if (t == 0) ms.advance (tot);
    __syncthreads ();
    for (int i = t; i < W; i += LB_THREADS) wg[i] = ws[i];
    if (t == 0) {
        // evaluations and the end in one 8-byte store, so the host never
        // sees one without the other
        LbFlag f {ms.evals, ms.a.done, ms.nrep, 0};
        *reinterpret_cast<volatile unsigned long long*> (flag) =
            *reinterpret_cast<const unsigned long long*> (&f);
    }
}

__global__ void __launch_bounds__ (LB_THREADS)
k_lb_convert (const double* __restrict__ x, float* __restrict__ coeff, int64_t n)
{
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride)
        coeff[e] = (float) x[e];
}

template <int BPMAX>
hipError_t
launch_post_m (const LbAct* act, const LbAct& av, const LbBufs& b, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_post<BPMAX>, dim3 (b.nb), dim3 (LB_THREADS), 0, st, act, av, b);
    return hipGetLastError ();
}

} // namespace

template <int MMAX>
hipError_t
launch_next_m (const LbAct* act, const LbAct& av, const LbBufs& b, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_next<MMAX>, dim3 (b.nb), dim3 (LB_THREADS), 0, st, act, av, b);
    return hipGetLastError ();
}

hipError_t
lb_launch_next (const LbAct* act, const LbAct& av, const LbBufs& b, hipStream_t st)
{
    switch (b.m) {
    case 1: return launch_next_m<1> (act, av, b, st);
    case 2: return launch_next_m<2> (act, av, b, st);
    case 3: return launch_next_m<3> (act, av, b, st);
    case 4: return launch_next_m<4> (act, av, b, st);
    case 5: return launch_next_m<5> (act, av, b, st);
    case 6: return launch_next_m<6> (act, av, b, st);
    case 7: return launch_next_m<7> (act, av, b, st);
    case 8: return launch_next_m<8> (act, av, b, st);
    case 9: return launch_next_m<9> (act, av, b, st);
    case 10: return launch_next_m<10> (act, av, b, st);
    case 11: return launch_next_m<11> (act, av, b, st);
    case 12: return launch_next_m<12> (act, av, b, st);
    case 13: return launch_next_m<13> (act, av, b, st);
    case 14: return launch_next_m<14> (act, av, b, st);
    case 15: return launch_next_m<15> (act, av, b, st);
    case 16: return launch_next_m<16> (act, av, b, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t
lb_launch_post (const LbAct* act, const LbAct& av, const LbBufs& b, hipStream_t st)
{
    switch (b.m - 1) {
    case 0: return launch_post_m<0> (act, av, b, st);
    case 1: return launch_post_m<1> (act, av, b, st);
    case 2: return launch_post_m<2> (act, av, b, st);
    case 3: return launch_post_m<3> (act, av, b, st);
    case 4: return launch_post_m<4> (act, av, b, st);
    case 5: return launch_post_m<5> (act, av, b, st);
    case 6: return launch_post_m<6> (act, av, b, st);
    case 7: return launch_post_m<7> (act, av, b, st);
    case 8: return launch_post_m<8> (act, av, b, st);
    case 9: return launch_post_m<9> (act, av, b, st);
    case 10: return launch_post_m<10> (act, av, b, st);
    case 11: return launch_post_m<11> (act, av, b, st);
    case 12: return launch_post_m<12> (act, av, b, st);
    case 13: return launch_post_m<13> (act, av, b, st);
    case 14: return launch_post_m<14> (act, av, b, st);
    case 15: return launch_post_m<15> (act, av, b, st);
    default: return hipErrorInvalidValue;
    }
}

template <int BPMAX>
hipError_t
launch_reduce_post_m (const PairDev* pairs, const double* part_score, const int* part_count, int nparts,
    int tile_scores, const EvalOut& out0, const LbAct* act, const LbAct& av, const LbBufs& b, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_reduce_post<BPMAX>, dim3 (b.nb), dim3 (LB_THREADS), 0, st, pairs, part_score,
        part_count, nparts, tile_scores, out0, act, av, b);
    return hipGetLastError ();
}

hipError_t
lb_launch_reduce_post (const PairDev* pairs, const double* part_score, const int* part_count, int nparts,
    int tile_scores, const EvalOut& out0, const LbAct* act, const LbAct& av, const LbBufs& b, hipStream_t st)
{
    if ((int64_t) b.nb * LB_THREADS < b.n || !out0.score || !out0.nvox) return hipErrorInvalidValue;
    switch (b.m - 1) {
    case 0: return launch_reduce_post_m<0> (pairs, part_score, part_count, nparts, tile_scores, out0, act, av, b, st);
    case 1: return launch_reduce_post_m<1> (pairs, part_score, part_count, nparts, tile_scores, out0, act, av, b, st);
    case 2: return launch_reduce_post_m<2> (pairs, part_score, part_count, nparts, tile_scores, out0, act, av, b, st);
    case 3: return launch_reduce_post_m<3> (pairs, part_score, part_count, nparts, tile_scores, out0, act, av, b, st);
    case 4: return launch_reduce_post_m<4> (pairs, part_score, part_count, nparts, tile_scores, out0, act, av, b, st);
    case 5: return launch_reduce_post_m<5> (pairs, part_score, part_count, nparts, tile_scores, out0, act, av, b, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t
lb_launch_finish (const LbAct* act, const LbAct& av, LbMachine* machine, const LbBufs& b, double* host,
    LbFlag* flag, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_finish, dim3 (1), dim3 (LB_THREADS), 0, st, act, av, machine, b, host, flag);
    return hipGetLastError ();
}

hipError_t
lb_launch_convert (const double* x, float* coeff, int64_t n, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_convert, dim3 (lb_blocks (n)), dim3 (LB_THREADS), 0, st, x, coeff, n);
    return hipGetLastError ();
}

} // namespace plmhip
