"""V-optimal MI histograms on the device (histogram_type=vopt:
bspline_initialize_mi_hist_vopt, src/register/joint_histogram.cxx:195-324,
and the key tables of add_pvi_8 / the pass-2 derivative) against the
oracle: the same key grid and key -> bin table bit for bit, and Mattes MI on
those bins with identical fixed-point histograms and the score and gradient
within the usual bars."""
import ctypes

import numpy as np
import pytest

import oracle
from plastimatch_amd import _abi, synth
from plastimatch_amd.bspline import Joint_histogram, Volume, hist_vopt, stage_for_pair
from tests.helpers import assert_parity_mi, oracle_case, small_pair

pytestmark = pytest.mark.gpu


def images():
    rng = np.random.default_rng(4)
    pair = small_pair(dim=(40, 36, 28), seed=31)
    yield "ct-like", pair.fixed
    yield "ct-remapped", synth.ct_remap(pair.moving.astype(np.float64)).astype(np.float32)
    # integer intensities: most keys empty
    yield "integers", np.round(rng.normal(100.0, 60.0, size=(20, 22, 24))).astype(np.float32)
    yield "two-level", np.where(rng.uniform(size=(16, 18, 20)) < 0.3, -1000.0, 40.0).astype(np.float32)
    yield "gamma", (rng.gamma(2.0, 200.0, size=(24, 20, 16)) - 1000.0).astype(np.float32)


@pytest.mark.parametrize("bins", [8, 32])
@pytest.mark.parametrize("where", ["host", "device"])
def test_hist_vopt_matches_oracle(hip_lib, bins, where):
    import torch
    for name, img in images():
        ref = oracle.hist_vopt(img, bins)
        data = torch.from_numpy(img).to("cuda:0") if where == "device" else img
        off, dl, lut = hist_vopt(Volume(data, (1.0, 1.0, 1.0)), bins)
        assert (off, dl) == ref[:2], name
        assert np.array_equal(lut, ref[2]), (name, np.flatnonzero(lut != ref[2])[:10])


def test_hist_vopt_full_size_ct(hip_lib):
    """The 256x256x128 CT-like image (8.4 M voxels): key table bit for bit."""
    import torch
    pair = synth.make_config("mse_256x256x128_vpr10", seed=0)
    for img in (pair.fixed, pair.moving):
        ref = oracle.hist_vopt(img, 32)
        off, dl, lut = hist_vopt(Volume(torch.from_numpy(img).to("cuda:0"), pair.spacing), 32)
        assert (off, dl) == ref[:2] and np.array_equal(lut, ref[2])


def test_hist_vopt_refusals(hip_lib):
    img = np.full((8, 8, 8), 5.0, np.float32)
    with pytest.raises(_abi.PlmHipError):
        hist_vopt(Volume(img, (1.0, 1.0, 1.0)), 16)
    ramp = np.arange(512, dtype=np.float32).reshape(8, 8, 8)
    with pytest.raises(_abi.PlmHipError):
        hist_vopt(Volume(ramp, (1.0, 1.0, 1.0)), _abi.VOPT_KEYS + 1)
    # more bins than keys: equal spacing, as the reference falls back
    jh = Joint_histogram(_abi.VOPT_KEYS + 1, 16, "vopt").initialize(Volume(ramp, (1, 1, 1)), Volume(ramp, (1, 1, 1)))
    assert jh.fixed_type == "eqsp" and jh.fixed_keys is None and jh.moving_type == "vopt"


@pytest.mark.parametrize("case", [0, 1])
def test_mi_on_vopt_bins_matches_oracle(hip_lib, case):
    dims = [((40, 36, 28), (1.0, 1.0, 2.5), (5, 5, 5), 32), ((31, 33, 29), (1.0, 1.0, 1.0), (4, 6, 7), 16)]
    dim, sp, vpr, bins = dims[case]
    pair = small_pair(dim=dim, spacing=sp, vpr=vpr, seed=40 + case)
    pair.moving = synth.ct_remap(pair.moving.astype(np.float64)).astype(np.float32)
    oc = oracle_case(pair)
    hp = oc.hist(bins, vopt=True)
    with stage_for_pair(pair, with_grad=False, bins=bins, hist_type="vopt") as st:
        assert np.array_equal(st.hist.fixed_keys, hp._luts[0])
        assert np.array_equal(st.hist.moving_keys, hp._luts[1])
        for k, coeff in enumerate([np.full(oc.num_coeff, 0.01, np.float32), pair.true_coeff]):
            r = st.score(coeff, "mi", want_hist=True)
            o = oc.mi(coeff, hp)
            s64, g64, _ = oc.mi_f64(coeff, hp)
            sh, gh, _ = oc.mi_f64(coeff, hp, hist=o)
            nj = bins * bins
            assert r.num_vox == o["num_vox"]
            assert np.array_equal(r.hist[:nj], o["fx_j_hist"])
            assert np.array_equal(r.hist[nj:], o["fx_f_hist"])
            assert_parity_mi(r.score, r.grad, (o["score"], o["grad"]), (s64, g64), (sh, gh),
                             label=f"vopt case {case} coeff {k}:")


def test_set_hist_keys_argument_errors(hip_lib):
    pair = small_pair(dim=(24, 20, 16), seed=3)
    with stage_for_pair(pair, with_grad=False, bins=8) as st:
        bad = np.zeros(_abi.VOPT_KEYS, np.int32)
        bad[-1] = 8
        with pytest.raises(_abi.PlmHipError):
            _abi.check(st.lib.plm_hip_bspline_set_hist_keys(st.ctx, bad.ctypes.data_as(ctypes.c_void_p), None))
    with stage_for_pair(pair, with_grad=False) as st:
        ok = np.zeros(_abi.VOPT_KEYS, np.int32)
        with pytest.raises(_abi.PlmHipError):
            _abi.check(st.lib.plm_hip_bspline_set_hist_keys(st.ctx, ok.ctypes.data_as(ctypes.c_void_p), None))
