// Host check of the L-BFGS direction solve (LbMachine::solve_direction,
// plastimatch_amd/csrc/lbfgs_logic.h), built by tests/test_lbfgs_direction.py.
//
// For every history length m = 1 .. 16 it runs 3m + 4 iterations on random
// vectors: each iteration forms a new pair (s, y) with s.y > 0 in slot e,
// hands the solve the products the post pass would (s_e.y_e, y_e.y_e,
// s_e.g, y_e.g and, per earlier pair j, s_e.y_j, s_j.y_e, y_e.y_j, s_j.g,
// y_j.g), forms d = -gamma g - sum ca_i s_i + gamma sum cu_i y_i as the
// trial-point kernel does, and compares d with liblbfgs's two-loop
// recursion (libs/liblbfgs-1.9/lib/lbfgs.c:555-609) over the same pairs.
// Where both forms of the solve apply (b <= 6) it also runs the
// runtime-length form on a copy of the machine and requires the same bits.
// Prints one line per m: "m <m> maxrel <max relative L2 error of d>
// bitdiff <count of solves whose two forms differ>".
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "lbfgs_kernels.h"
#include "lbfgs_logic.h"

using namespace plmhip;

namespace {

double
dot (const std::vector<double>& a, const std::vector<double>& b)
{
    double s = 0.0;
    for (size_t i = 0; i < a.size (); i++) s += a[i] * b[i];
    return s;
}

bool
same_solve (const LbMachine& a, const LbMachine& b, int nb)
{
    if (std::memcmp (&a.a.gamma, &b.a.gamma, sizeof (double)) != 0) return false;
    if (std::memcmp (a.a.ca, b.a.ca, sizeof (double) * nb) != 0) return false;
    if (std::memcmp (a.a.cu, b.a.cu, sizeof (double) * nb) != 0) return false;
    if (std::memcmp (a.a.slot, b.a.slot, sizeof (int) * nb) != 0) return false;
    if (std::memcmp (a.SY, b.SY, sizeof a.SY) != 0) return false;
    return std::memcmp (a.YY, b.YY, sizeof a.YY) == 0;
}

} // namespace

int
main ()
{
    const int n = 24;
    std::mt19937_64 rng (11);
    std::normal_distribution<double> N01 (0.0, 1.0);
    std::uniform_real_distribution<double> U (0.5, 2.0);
    static LbMachine M, C;
    for (int m = 1; m <= LB_MAX_M; ++m) {
        std::memset (&M, 0, sizeof M);
        plm_lbfgs_parms p {};
        p.m = m;
        M.init (p, 0, n);
        // y = A s with A symmetric positive definite, so s.y > 0
        std::vector<double> diag (n);
        for (double& v : diag) v = U (rng);
        std::vector<std::vector<double>> S (m, std::vector<double> (n)), Y (m, std::vector<double> (n));
        int end = 0, bitdiff = 0;
        double maxrel = 0.0;
        for (int k = 1; k <= 3 * m + 4; ++k) {
            const int b = m <= k ? m : k;
            const int e = end;
            for (int i = 0; i < n; i++) {
                S[e][i] = N01 (rng);
                Y[e][i] = diag[i] * S[e][i] + 0.05 * N01 (rng);
            }
            std::vector<double> g (n);
            for (double& v : g) v = N01 (rng);
            double red[7 + 5 * LB_MAX_M] = {0};
            red[3] = dot (S[e], Y[e]);
            red[4] = dot (Y[e], Y[e]);
            red[5] = dot (S[e], g);
            red[6] = dot (Y[e], g);
            for (int kk = 1; kk < b; ++kk) {
                const int j = (e - kk + m) % m;
                const int c = 7 + 5 * (kk - 1);
                red[c] = dot (S[e], Y[j]);
                red[c + 1] = dot (S[j], Y[e]);
                red[c + 2] = dot (Y[e], Y[j]);
                red[c + 3] = dot (S[j], g);
                red[c + 4] = dot (Y[j], g);
            }
            C = M;
            M.solve_direction (red, b, e);
            if (b <= 6) {
                C.a.dir_b = b;
                C.solve_any (red, b, e);
                bitdiff += same_solve (M, C, b) ? 0 : 1;
            }
            // d as the trial-point kernel forms it
            std::vector<double> d (n);
            for (int i = 0; i < n; i++) {
                double t1 = 0.0, t2 = 0.0;
                for (int q = 0; q < M.a.dir_b; ++q) {
                    t1 = t1 + M.a.ca[q] * S[M.a.slot[q]][i];
                    t2 = t2 + M.a.cu[q] * Y[M.a.slot[q]][i];
                }
                d[i] = -M.a.gamma * g[i] - t1 + M.a.gamma * t2;
            }
            // the two-loop recursion over the b newest pairs
            std::vector<double> q = g, alpha (b);
            for (int kk = 0; kk < b; ++kk) {
                const int j = (e - kk + m) % m;
                alpha[kk] = dot (S[j], q) / dot (Y[j], S[j]);
                for (int i = 0; i < n; i++) q[i] -= alpha[kk] * Y[j][i];
            }
            const double gamma = dot (S[e], Y[e]) / dot (Y[e], Y[e]);
            for (double& v : q) v *= gamma;
            for (int kk = b - 1; kk >= 0; --kk) {
                const int j = (e - kk + m) % m;
                const double beta = dot (Y[j], q) / dot (Y[j], S[j]);
                for (int i = 0; i < n; i++) q[i] += (alpha[kk] - beta) * S[j][i];
            }
            double num = 0.0, den = 0.0;
            for (int i = 0; i < n; i++) {
                num += (d[i] + q[i]) * (d[i] + q[i]);
                den += q[i] * q[i];
            }
            maxrel = std::fmax (maxrel, std::sqrt (num / den));
            end = (end + 1) % m;
        }
        std::printf ("m %d maxrel %.3e bitdiff %d\n", m, maxrel, bitdiff);
    }
    return 0;
}
