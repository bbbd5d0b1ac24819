"""MSE score+gradient of the HIP flavor against the oracle (flavors g and p
of src/register/bspline_mse.cxx), through the C ABI.

Bars: score within 1e-5 relative, gradient within 1e-4 relative L2, the
overlap count identical; errors against the FP64 oracle are checked too.
"""
import numpy as np
import pytest

from plastimatch_amd import synth
from plastimatch_amd.bspline import (BsplineHip, Bspline_xform, Volume, stage_for_pair,
                                     volume_gradient)
from tests.helpers import assert_count, assert_parity, bf16_round, oracle_case, rel, rel_scalar, small_pair

pytestmark = pytest.mark.gpu

CASES = [
    # dim, spacing, vpr, origin
    ((40, 36, 28), (1.0, 1.0, 2.5), (5, 5, 5), (0.0, 0.0, 0.0)),
    ((33, 29, 31), (1.0, 1.0, 1.0), (4, 6, 7), (-12.5, 3.0, 100.0)),
    ((64, 48, 20), (0.9765625, 0.9765625, 2.5), (10, 10, 10), (0.0, 0.0, 0.0)),
    ((30, 30, 30), (1.5, 1.25, 2.0), (15, 15, 15), (0.0, 0.0, 0.0)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("grad_mode", ["ref_nn", "exact"])
def test_mse_matches_oracle(hip_lib, case, grad_mode):
    dim, sp, vpr, org = CASES[case]
    pair = small_pair(dim=dim, spacing=sp, vpr=vpr, origin=org, seed=10 + case)
    oc = oracle_case(pair)
    with stage_for_pair(pair, moving_grad=Volume(oc.m_grad, sp, org)) as st:
        for k, coeff in enumerate([np.zeros(oc.num_coeff, np.float32), pair.true_coeff,
                                   synth.random_coeff(pair.dim, vpr, 2.0, seed=case + 7)]):
            r = st.score(coeff, "mse", grad_mode)
            flavor = "g" if grad_mode == "ref_nn" else "p"
            s_ref, g_ref, n_ref, _ = oc.mse(coeff, flavor)
            s64, g64, n64 = oc.mse_f64(coeff, 0 if grad_mode == "ref_nn" else 1)
            assert_count(r.num_vox, n_ref, n64, f"case {case} {grad_mode} coeff {k}:")
            assert_parity(r.score, r.grad, (s_ref, g_ref), (s64, g64),
                          label=f"case {case} {grad_mode} coeff {k}:")


def test_volume_gradient_bit_exact(hip_lib):
    pair = small_pair(dim=(37, 22, 19), spacing=(0.8, 1.3, 2.5))
    oc = oracle_case(pair)
    g = volume_gradient(Volume(pair.moving, pair.spacing, pair.origin))
    assert np.array_equal(g.img, oc.m_grad)


def test_identical_images_zero_score(hip_lib):
    pair = small_pair()
    fixed = Volume(pair.fixed, pair.spacing)
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    with BsplineHip(bxf, fixed, fixed, volume_gradient(fixed)) as st:
        r = st.score(np.zeros(bxf.num_coeff, np.float32), "mse", "ref_nn")
        assert r.score == 0.0
        assert np.all(r.grad == 0.0)
        assert r.num_vox == pair.fixed.size


def test_constant_offset_score(hip_lib):
    pair = small_pair()
    fixed = Volume(pair.fixed, pair.spacing)
    moving = Volume(pair.fixed + np.float32(3.0), pair.spacing)
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    with BsplineHip(bxf, fixed, moving) as st:
        r = st.score(np.zeros(bxf.num_coeff, np.float32), "mse", "exact")
        assert abs(r.score - 9.0) < 1e-4


def test_no_overlap_gives_flt_max(hip_lib):
    pair = small_pair()
    oc = oracle_case(pair)
    with stage_for_pair(pair) as st:
        coeff = np.full(oc.num_coeff, 1e4, np.float32)
        r = st.score(coeff, "mse", "ref_nn")
        s_ref, g_ref, n_ref, _ = oc.mse(coeff, "g")
        assert n_ref == 0 and r.num_vox == 0
        assert r.status == 1
        assert np.float32(r.score) == np.float32(s_ref)
        assert np.all(r.grad == 0)


def test_deterministic(hip_lib):
    pair = small_pair(dim=(64, 48, 40))
    with stage_for_pair(pair) as st:
        a = st.score(pair.true_coeff, "mse", "ref_nn")
        b = st.score(pair.true_coeff, "mse", "ref_nn")
        assert a.score == b.score
        assert np.array_equal(a.grad, b.grad)


def test_bf16_storage_matches_oracle_on_rounded_images(hip_lib):
    pair = small_pair(dim=(48, 40, 24), vpr=(8, 8, 8))
    pair.fixed = bf16_round(pair.fixed)
    pair.moving = bf16_round(pair.moving)
    oc = oracle_case(pair)
    mg = Volume(oc.m_grad, pair.spacing)
    st_exact = stage_for_pair(pair, storage="bf16", with_grad=False)
    r = st_exact.score(pair.true_coeff, "mse", "exact")
    s_ref, g_ref, n_ref, _ = oc.mse(pair.true_coeff, "p")
    s64, g64, n64 = oc.mse_f64(pair.true_coeff, 1)
    assert_count(r.num_vox, n_ref, n64, "bf16:")
    assert_parity(r.score, r.grad, (s_ref, g_ref), (s64, g64), label="bf16:")
    st_exact.close()


def test_full_size_mse_256x256x128(hip_lib):
    pair = synth.make_config("mse_256x256x128_vpr10", seed=0)
    oc = oracle_case(pair)
    with stage_for_pair(pair, moving_grad=Volume(oc.m_grad, pair.spacing)) as st:
        coeff = synth.random_coeff(pair.dim, pair.vox_per_rgn, 3.0, seed=5)
        r = st.score(coeff, "mse", "ref_nn")
        s_ref, g_ref, n_ref, _ = oc.mse(coeff, "g")
        s64, g64, n64 = oc.mse_f64(coeff, 0)
        assert_count(r.num_vox, n_ref, n64, "256x256x128:")
        assert_parity(r.score, r.grad, (s_ref, g_ref), (s64, g64), label="256x256x128:")
        r2 = st.score(coeff, "mse", "ref_nn")
        assert np.array_equal(r.grad, r2.grad)


@pytest.mark.parametrize("grad_mode", ["ref_nn", "exact"])
def test_full_size_mse_128cube_vpr15(hip_lib, grad_mode):
    """The 128^3 workload at a 15-voxel grid (40 Gaussians plus noise, moving
    warped by a N(0, 2 mm) field) at zero and at the true coefficients,
    against the oracle's flavors g and p and FP64."""
    pair = synth.make_config("mse_128cube_vpr15", seed=0)
    oc = oracle_case(pair)
    flavor, mode = ("g", 0) if grad_mode == "ref_nn" else ("p", 1)
    with stage_for_pair(pair, moving_grad=Volume(oc.m_grad, pair.spacing)) as st:
        for name, coeff in (("zero", np.zeros(oc.num_coeff, np.float32)), ("true", pair.true_coeff)):
            r = st.score(coeff, "mse", grad_mode)
            s_ref, g_ref, n_ref, _ = oc.mse(coeff, flavor)
            s64, g64, n64 = oc.mse_f64(coeff, mode)
            label = f"128^3 vpr 15 {grad_mode} {name}:"
            assert_count(r.num_vox, n_ref, n64, label)
            assert_parity(r.score, r.grad, (s_ref, g_ref), (s64, g64), label=label)


@pytest.mark.parametrize("vpr", [(7, 6, 5), (10, 10, 10)])
def test_launch_shapes_agree_and_autotune(hip_lib, vpr):
    """Every tile-kernel launch shape (voxels per thread, workgroups per CU)
    gives the same gradient bits (the condense order does not depend on it)
    and the same score to rounding; the autotuner picks one of them."""
    pair = small_pair(dim=(64, 48, 40), vpr=vpr, seed=41)
    with stage_for_pair(pair) as st:
        assert st.get_tuning()[1] == 0
        res = {}
        for v in (1, 2, 4):
            for w in (0, 4, 1):
                st.set_tuning(v)
                st.set_occupancy(w)
                assert st.get_tuning() == (v, w)
                res[v, w] = st.score(pair.true_coeff, "mse", "ref_nn")
        ref = res[1, 0]
        for key, r in res.items():
            assert np.array_equal(r.grad, ref.grad), key
            # the squared differences of a thread's batch of G voxels are
            # summed in FP32 before the FP64 total, so G moves the score by
            # rounding only (a few FP32 roundings of a batch sum, ~1e-9 of
            # the total here)
            assert rel_scalar(r.score, ref.score) < 1e-8, key
            assert r.num_vox == ref.num_vox, key
        choice, ms = st.autotune("mse", "ref_nn", reps=3)
        assert choice in (1, 2, 4) and ms > 0
        v, w = st.get_tuning()
        assert v == choice and w in (0, 4, 5, 6)
        r = st.score(pair.true_coeff, "mse", "ref_nn")
        assert np.array_equal(r.grad, ref.grad)


def test_tile_counters_match_fixed_stride(hip_lib):
    """At one workgroup per CU a 128x128x96 pair with 8-voxel tiles gives
    each workgroup at least 12 tiles, so the tile kernel takes its tiles
    from the XCD counters; with as many workgroups as fit each walks one or
    two at a fixed stride.  Both give the same score, count and gradient
    bits, call after call (each launch leaves its counters zero), for MSE
    with either derivative, the score alone and MI pass 2."""
    pair = small_pair(dim=(128, 128, 96), vpr=(8, 8, 8), seed=43, n_blobs=20)
    with stage_for_pair(pair, bins=24) as st:
        st.set_tuning(1)
        for metric, mode in (("mse", "ref_nn"), ("mse", "exact"), ("mse", "none"), ("mi", "ref_nn")):
            res = []
            for w in (0, 1, 1, 0, 1):
                st.set_occupancy(w)
                res.append((w, st.score(pair.true_coeff, metric, mode)))
            ref = res[0][1]
            for w, r in res:
                assert r.score == ref.score, (metric, mode, w)
                assert r.num_vox == ref.num_vox, (metric, mode, w)
                assert np.array_equal(r.grad, ref.grad), (metric, mode, w)
