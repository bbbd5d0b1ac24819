"""Rotated and non-dyadic geometry on the device: the general-direction
paths (direction cosines off the diagonal, so the tile kernel's full 3x3
moving-index map, MI pass 1 without the aligned shortcut, the projected
moving gradient), with a moving image of its own origin, spacing and
direction cosines, against the oracle.

Fixed spacings such as 1.1 mm are not dyadic, so even an axis-aligned image
walks pass 1's positions (the reference's LOOP_X recurrence) differently
from the direct formula; the device keeps the walk's positions for pass 1
there, and its histograms must still equal the float flavor's bin for bin.
"""
import numpy as np
import pytest

import oracle
from plastimatch_amd import synth
from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Joint_histogram, Volume, volume_gradient
from tests.helpers import assert_count, assert_parity, assert_parity_mi, rel

pytestmark = pytest.mark.gpu


def rot(ax, ay, az):
    """Row-major direction cosines Rz(az) Ry(ay) Rx(ax), as float32."""
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return tuple(float(v) for v in (rz @ ry @ rx).astype(np.float32).reshape(-1))


GEOMS = {
    "rotated": dict(f_sp=(1.0, 1.1, 2.3), f_org=(-3.0, 5.0, 1.5), f_dc=rot(0.05, -0.03, 0.12),
                    m_sp=(0.95, 1.05, 2.2), m_org=(-2.0, 4.0, 0.5), m_dc=rot(0.0, 0.02, 0.10)),
    "aligned_non_dyadic": dict(f_sp=(1.1, 0.9, 2.3), f_org=(-3.3, 5.1, 1.7), f_dc=rot(0, 0, 0),
                               m_sp=(1.05, 0.95, 2.2), m_org=(-2.0, 4.0, 0.5), m_dc=rot(0, 0, 0)),
}


def build(name, seed, bins=None):
    gm = GEOMS[name]
    pair = synth.make_pair(name, (44, 38, 30), gm["f_sp"], (6, 5, 4), 1.5, 12, seed=seed, ct=True)
    fixed = Volume(pair.fixed, gm["f_sp"], gm["f_org"], gm["f_dc"])
    moving = Volume(pair.moving, gm["m_sp"], gm["m_org"], gm["m_dc"])
    oc = oracle.Case(pair.fixed, pair.moving, gm["f_sp"], pair.vox_per_rgn, origin=gm["f_org"], dc=gm["f_dc"],
                     m_spacing=gm["m_sp"], m_origin=gm["m_org"], m_dc=gm["m_dc"])
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    return pair, fixed, moving, oc, bxf


@pytest.mark.parametrize("name", list(GEOMS))
def test_moving_gradient_rotated_bit_exact(hip_lib, name):
    pair, fixed, moving, oc, bxf = build(name, 3)
    g = volume_gradient(moving)
    assert np.array_equal(g.img, oc.m_grad)


@pytest.mark.parametrize("name", list(GEOMS))
@pytest.mark.parametrize("grad_mode", ["ref_nn", "exact"])
def test_mse_general_geometry(hip_lib, name, grad_mode):
    pair, fixed, moving, oc, bxf = build(name, 4)
    flavor, mode = ("g", 0) if grad_mode == "ref_nn" else ("p", 1)
    with BsplineHip(bxf, fixed, moving, Volume(oc.m_grad, moving.spacing, moving.origin,
                                                moving.direction_cosines)) as st:
        for k, coeff in enumerate([np.zeros(oc.num_coeff, np.float32),
                                   synth.random_coeff(pair.dim, pair.vox_per_rgn, 1.5, seed=11)]):
            r = st.score(coeff, "mse", grad_mode)
            s_ref, g_ref, n_ref, _ = oc.mse(coeff, flavor)
            s64, g64, n64 = oc.mse_f64(coeff, mode)
            label = f"{name} {grad_mode} coeff {k}:"
            assert n64 > 0.3 * pair.fixed.size, "the geometries must overlap"
            assert_count(r.num_vox, n_ref, n64, label)
            assert_parity(r.score, r.grad, (s_ref, g_ref), (s64, g64), label=label)


@pytest.mark.parametrize("name", list(GEOMS))
def test_mi_general_geometry(hip_lib, name):
    pair, fixed, moving, oc, bxf = build(name, 5)
    hp = oc.hist(24)
    hist = Joint_histogram(24, 24).initialize(fixed, moving)
    assert (hist.fixed_offset, hist.fixed_delta) == (hp.f_offset, hp.f_delta)
    mismatch = oracle.pass1_position_mismatch(pair.fixed, fixed.spacing, oc.g, fixed.origin,
                                              fixed.direction_cosines)
    print(f"{name}: pass-1 walk positions differing from the direct formula: {mismatch}")
    assert mismatch > 0
    with BsplineHip(bxf, fixed, moving, hist=hist) as st:
        for k, coeff in enumerate([np.full(oc.num_coeff, 0.01, np.float32),
                                   synth.random_coeff(pair.dim, pair.vox_per_rgn, 1.5, seed=12)]):
            r = st.score(coeff, "mi", want_hist=True)
            o = oc.mi(coeff, hp)
            s64, g64, _ = oc.mi_f64(coeff, hp)
            sh, gh, _ = oc.mi_f64(coeff, hp, hist=o)
            nj = 24 * 24
            assert r.num_vox == o["num_vox"]
            assert np.array_equal(r.hist[:nj], o["fx_j_hist"])
            assert np.array_equal(r.hist[nj:], o["fx_f_hist"])
            assert_parity_mi(r.score, r.grad, (o["score"], o["grad"]), (s64, g64), (sh, gh),
                             label=f"{name} coeff {k}:")


@pytest.mark.parametrize("linear", [True, False])
def test_warp_rotated(hip_lib, linear):
    import torch
    fill = -1234.5
    pair, fixed, moving, oc, bxf = build("rotated", 6)
    coeff = synth.random_coeff(pair.dim, pair.vox_per_rgn, 2.0, seed=13)
    with BsplineHip(bxf, fixed, moving) as st:
        out = torch.full((st.fixed_voxels,), 7.0, dtype=torch.float32, device="cuda")
        st.warp(torch.from_numpy(coeff).cuda(), out, fill, linear)
        torch.cuda.synchronize()
        got = out.cpu().numpy().reshape(pair.fixed.shape)
    ref = oracle.warp(oc.g, pair.moving, moving.spacing, coeff, linear, fill, m_origin=moving.origin,
                      m_dc=moving.direction_cosines)
    d_got, d_ref = got == fill, ref == fill
    flips = np.count_nonzero(d_got != d_ref)
    both = ~d_got & ~d_ref
    print(f"rotated warp linear={linear}: default voxels {d_ref.sum()}, inside-test flips {flips}")
    assert both.sum() > 0.3 * got.size
    assert flips <= max(2, got.size // 20000)
    span = float(pair.moving.max() - pair.moving.min())
    if linear:
        err = np.abs(got[both].astype(np.float64) - ref[both])
        print(f"  max |err| {err.max():.3e} of range {span:.1f}")
        assert err.max() < 1e-3 * span
        assert np.sqrt((err ** 2).mean()) < 1e-5 * span
    else:
        picks = np.count_nonzero(got[both] != ref[both])
        print(f"  nearest-voxel flips {picks} of {both.sum()}")
        assert picks <= max(2, got.size // 5000)
