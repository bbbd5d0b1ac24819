"""Mattes MI of the HIP flavor against the oracle (flavor g of
src/register/bspline_mi.cxx): identical fixed-point histograms, score
within 1e-5 relative, gradient within 1e-4 relative L2."""
import numpy as np
import pytest

import oracle
from plastimatch_amd import synth
from plastimatch_amd.bspline import stage_for_pair
from tests.helpers import assert_parity_mi, oracle_case, rel, rel_scalar, small_pair

pytestmark = pytest.mark.gpu

CASES = [
    ((40, 36, 28), (1.0, 1.0, 2.5), (5, 5, 5), 32),
    ((48, 32, 30), (0.9765625, 0.9765625, 2.5), (12, 12, 12), 32),
    ((31, 33, 29), (1.0, 1.0, 1.0), (4, 6, 7), 20),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_mi_matches_oracle(hip_lib, case):
    dim, sp, vpr, bins = CASES[case]
    pair = small_pair(dim=dim, spacing=sp, vpr=vpr, seed=20 + case)
    pair.moving = synth.ct_remap(pair.moving.astype(np.float64)).astype(np.float32)
    oc = oracle_case(pair)
    assert oracle.pass1_position_mismatch(pair.fixed, sp, oc.g) == 0
    hp = oc.hist(bins)
    with stage_for_pair(pair, with_grad=False, bins=bins) as st:
        for k, coeff in enumerate([np.full(oc.num_coeff, 0.01, np.float32), pair.true_coeff,
                                   synth.random_coeff(pair.dim, vpr, 2.0, seed=case + 7)]):
            r = st.score(coeff, "mi", want_hist=True)
            o = oc.mi(coeff, hp)
            s64, g64, n64 = oc.mi_f64(coeff, hp)
            sh, gh, _ = oc.mi_f64(coeff, hp, hist=o)
            nj = bins * bins
            assert r.num_vox == o["num_vox"]
            assert np.array_equal(r.hist[:nj], o["fx_j_hist"])
            assert np.array_equal(r.hist[nj:], o["fx_f_hist"])
            assert r.hist[:nj].sum() == r.num_vox << 32
            assert_parity_mi(r.score, r.grad, (o["score"], o["grad"]), (s64, g64), (sh, gh),
                             label=f"case {case} coeff {k}:")


def test_full_size_mi_512x512x160(hip_lib):
    """Mattes MI at 512x512x160 (CT-like intensities, 12-voxel grid, N(0, 4
    mm) field, remapped moving intensities, 32x32 bins): the fixed-point
    histogram and N identical to the oracle's float flavor, the score within
    1e-5 and the gradient within 1e-4 (relative L2) of the full FP64
    oracle."""
    pair = synth.make_config("mi_512x512x160_vpr12", seed=0)
    oc = oracle_case(pair)
    hp = oc.hist(32)
    coeff = synth.random_coeff(pair.dim, pair.vox_per_rgn, 4.0, seed=401)
    with stage_for_pair(pair, with_grad=False, bins=32) as st:
        r = st.score(coeff, "mi", want_hist=True)
    o = oc.mi(coeff, hp)
    s64, g64, _ = oc.mi_f64(coeff, hp)
    sh, gh, _ = oc.mi_f64(coeff, hp, hist=o)
    nj = 32 * 32
    assert r.num_vox == o["num_vox"]
    assert np.array_equal(r.hist[:nj], o["fx_j_hist"])
    assert np.array_equal(r.hist[nj:], o["fx_f_hist"])
    d = assert_parity_mi(r.score, r.grad, (o["score"], o["grad"]), (s64, g64), (sh, gh),
                         label="512x512x160:")
    # the headline bars against the full FP64 oracle, asserted unconditionally
    assert d["score_full64"] < 1e-5 and d["grad_full64"] < 1e-4, d


def test_mi_deterministic(hip_lib):
    pair = small_pair(dim=(64, 48, 40), vpr=(8, 8, 8))
    with stage_for_pair(pair, with_grad=False, bins=32) as st:
        a = st.score(pair.true_coeff, "mi", want_hist=True)
        b = st.score(pair.true_coeff, "mi", want_hist=True)
        assert np.array_equal(a.hist, b.hist)
        assert a.score == b.score
        assert np.array_equal(a.grad, b.grad)
