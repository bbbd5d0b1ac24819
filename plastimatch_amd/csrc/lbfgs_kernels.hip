// Vector kernels of the device-resident L-BFGS: the vector work of
// liblbfgs's lbfgs() (libs/liblbfgs-1.9/lib/lbfgs.c:449-609) and of its
// More-Thuente trials (:903-906), in double, and the pass that runs the
// scalar logic on the device.
//
// Built with -ffp-contract=off: each element update rounds its product and
// its sum separately, as liblbfgs's vecadd / vecdiff do.

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
                double t1 = 0.0, t2 = 0.0;
                for (int i = 0; i < nd; ++i) {
                    const int64_t off = (int64_t) act->slot[i] * n + e;
                    t1 = t1 + act->ca[i] * b.S[off];
                    t2 = t2 + act->cu[i] * b.Y[off];
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

// After the evaluation: the gradient it wrote kept in the trial's g buffer,
// and the partials of every post component (LbAct::post_bp <= BPMAX earlier
// pairs; the start point has no pair and takes g for d).
template <int BPMAX>
__global__ void __launch_bounds__ (LB_THREADS)
k_lb_post (const LbAct* __restrict__ ap, LbAct av, LbBufs b)
{
    const LbAct* act = ap ? ap : &av;
    if (act->done) return;
    constexpr int NC = 7 + 5 * BPMAX;
    const int pe = act->post_e, bp = act->post_bp, m = b.m;
    const bool pair = pe >= 0;
    const double* __restrict__ x = b.X[act->xi];
    const double* __restrict__ xp = b.X[act->xpi];
    const double* __restrict__ g = b.geval;
    double* __restrict__ gk = b.G[act->gi];
    const double* __restrict__ gp = b.G[act->gpi];
    const double* __restrict__ d = act->kind == 0 ? nullptr : b.d;
    const int64_t n = b.n;
    double v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = 0.0;
    int slot[BPMAX > 0 ? BPMAX : 1];
#pragma unroll
    for (int i = 1; i <= BPMAX; ++i) slot[i - 1] = (pe - i + m) % m;
    const int64_t stride = (int64_t) gridDim.x * LB_THREADS;
    for (int64_t e = (int64_t) blockIdx.x * LB_THREADS + threadIdx.x; e < n; e += stride) {
        const double xe = x[e], ge = g[e];
        gk[e] = ge;
        const double de = d ? d[e] : ge;
        v[0] += ge * de;
        v[1] += xe * xe;
        v[2] += ge * ge;
        if (pair) {
            const double se = xe - xp[e];
            const double ye = ge - gp[e];
            b.S[(int64_t) pe * n + e] = se;
            b.Y[(int64_t) pe * n + e] = ye;
            v[3] += se * ye;
            v[4] += ye * ye;
            v[5] += se * ge;
            v[6] += ye * ge;
#pragma unroll
            for (int i = 0; i < BPMAX; ++i) {
                if (i < bp) {
                    const double sj = b.S[(int64_t) slot[i] * n + e];
                    const double yj = b.Y[(int64_t) slot[i] * n + e];
                    v[7 + 5 * i + 0] += se * yj;
                    v[7 + 5 * i + 1] += sj * ye;
                    v[7 + 5 * i + 2] += ye * yj;
                    v[7 + 5 * i + 3] += sj * ge;
                    v[7 + 5 * i + 4] += yj * ge;
                }
            }
        }
    }
    block_partials<NC> (v, b.part, gridDim.x);
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

hipError_t
lb_launch_next (const LbAct* act, const LbAct& av, const LbBufs& b, hipStream_t st)
{
    hipLaunchKernelGGL (k_lb_next, dim3 (b.nb), dim3 (LB_THREADS), 0, st, act, av, b);
    return hipGetLastError ();
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
