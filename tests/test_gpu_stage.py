"""Stage preprocessing on the GPU against the oracle: the subsampling of
registration_resample_volume (plm_hip_stage_resample /
plm_hip_volume_resample) bit for bit, and the EQSP histogram range scan
(plm_hip_hist_eqsp) exactly.  The oracle's restatement is pinned by closed
forms in tests/test_oracle_resample.py."""
import numpy as np
import pytest

import oracle
from plastimatch_amd import synth
from plastimatch_amd.bspline import (BsplineHip, Bspline_xform, Joint_histogram, Volume, hist_eqsp,
                                     resample_geometry, stage_resample, volume_gradient, volume_resample)
from tests.helpers import assert_count, assert_parity

pytestmark = pytest.mark.gpu

SP, ORG = (1.0, 1.5, 2.5), (-3.0, 2.0, 10.0)


def image(shape=(19, 23, 30), seed=0):
    return np.random.default_rng(seed).normal(100, 40, shape).astype(np.float32)


def on(img, where):
    if where == "device":
        import torch
        return torch.from_numpy(img).cuda()
    return img


def host(a):
    return a.cpu().numpy() if hasattr(a, "cpu") else a


@pytest.mark.parametrize("where", ["host", "device"])
@pytest.mark.parametrize("nearest", [False, True])
@pytest.mark.parametrize("kind,rate", [("voxel_rate", (2, 3, 2)), ("voxel_rate", (4, 4, 1)),
                                       ("voxel_legacy", (2, 2, 1)), ("voxel_legacy", (3, 5, 2)),
                                       ("mm", (2.2, 1.7, 3.1)), ("pct", (0.45, 0.7, 0.6)),
                                       ("voxel_rate", (1, 1, 1))])
def test_stage_resample_matches_oracle(hip_lib, kind, rate, nearest, where):
    img = image()
    got = stage_resample(Volume(on(img, where), SP, ORG), kind, rate, nearest)
    ref, o, s = oracle.stage_resample(img, SP, ORG, kind, rate, nearest)
    assert got.origin == o and got.spacing == s
    assert np.array_equal(host(got.img), ref)


def test_volume_resample_beyond_the_input(hip_lib):
    img = image(seed=1)
    dim, o, s = (25, 20, 15), (-4.3, 0.7, 5.0), (1.3, 2.1, 3.3)
    got = volume_resample(Volume(img, SP, ORG), dim, o, s)
    ref = oracle.volume_resample(img, SP, ORG, dim, o, s)
    assert np.array_equal(host(got.img), ref)
    assert np.count_nonzero(ref == 0) > 0


@pytest.mark.parametrize("kind,rate", [("voxel_legacy", (2, 2, 1)), ("voxel_rate", (4, 4, 1))])
def test_stage_resample_full_size(hip_lib, kind, rate):
    pair = synth.make_config("mse_256x256x128_vpr10", seed=0)
    got = stage_resample(Volume(on(pair.moving, "device"), pair.spacing, pair.origin), kind, rate)
    ref, o, s = oracle.stage_resample(pair.moving, pair.spacing, pair.origin, kind, rate)
    print(f"{kind} {rate}: {pair.moving.shape} -> {ref.shape}")
    assert np.array_equal(host(got.img), ref)


@pytest.mark.parametrize("where", ["host", "device"])
def test_hist_eqsp_matches_oracle(hip_lib, where):
    rng = np.random.default_rng(0)
    f = rng.normal(0, 300, (10, 11, 12)).astype(np.float32)
    m = rng.normal(40, 100, (33, 17, 70)).astype(np.float32)
    jh = Joint_histogram(32, 20).initialize(Volume(on(f, where), (1, 1, 1)), Volume(on(m, where), (1, 1, 1)))
    assert (jh.fixed_offset, jh.fixed_delta) == oracle.hist_eqsp(f, 32)
    assert (jh.moving_offset, jh.moving_delta) == oracle.hist_eqsp(m, 20)
    assert hist_eqsp(Volume(on(m, where), (1, 1, 1)), 7) == oracle.hist_eqsp(m, 7)


def test_stage_chain_on_device(hip_lib):
    """A stage as registration_data.cxx:350-384 builds it, on the device:
    both images subsampled, the moving gradient of the subsampled moving
    image, then one score -- against the oracle on the same volumes."""
    pair = synth.make_pair("stage", (48, 44, 36), (1.0, 1.0, 2.0), (5, 5, 5), 2.0, 10, seed=7, ct=True)
    fx = stage_resample(Volume(on(pair.fixed, "device"), pair.spacing), "voxel_legacy", (2, 2, 1))
    mv = stage_resample(Volume(on(pair.moving, "device"), pair.spacing), "voxel_legacy", (2, 2, 1))
    bxf = Bspline_xform.for_volume(fx, (5, 5, 5))
    with BsplineHip(bxf, fx, mv, volume_gradient(mv)) as st:
        coeff = np.random.default_rng(3).normal(0, 1.0, bxf.num_coeff).astype(np.float32)
        r = st.score(coeff, "mse", "ref_nn")
    f_h, m_h = host(fx.img), host(mv.img)
    case = oracle.Case(f_h, m_h, fx.spacing, (5, 5, 5), origin=fx.origin)
    s, g, n, _ = case.mse(coeff, "g")
    s64, g64, n64 = case.mse_f64(coeff, 0)
    assert_count(r.num_vox, n, n64, "stage chain")
    assert_parity(r.score, r.grad, (s, g), (s64, g64), label="stage chain")
