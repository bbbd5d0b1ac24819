"""The differentiable MSE path: the score-only forward (PLM_GRAD_NONE), the
dense displacement field (bspline_interpolate_vf), the gradient from a
stored field, and the torch.autograd.Function built on them.

The adjoint is checked against the oracle's flavor p gradient (the
reference CPU flavor and FP64, 1e-5 / 1e-4 bars) and against central finite
differences (h = 1e-3 mm) of the FP64 score, at 1e-4 relative L2, on
coefficients whose support holds no voxel that li_clamp clamps at the
moving-image border: there li_value_dx keeps its slope while the clamped
value stays put (src/base/interpolate.cxx:17-41, 162-244), so flavor p's
gradient is not the derivative of its score, and those coefficients are
reported apart.  The full-size case is 192x192x96 at an 8-voxel grid with
200 random coefficients.
"""
import numpy as np
import pytest

import oracle
from plastimatch_amd import synth
from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume
from tests.helpers import assert_parity, rel, rel_scalar, small_pair

pytestmark = pytest.mark.gpu

ROI_OFF = (4, 4, 4)
ROI_DIM = (24, 22, 16)


def interior_stage(seed=5, roi=True):
    pair = small_pair(dim=(32, 30, 24), spacing=(1.0, 1.0, 1.0), vpr=(6, 6, 6), ct=False, seed=seed)
    fixed = Volume(pair.fixed, pair.spacing)
    moving = Volume(pair.moving, pair.spacing)
    kw = dict(roi_offset=ROI_OFF, roi_dim=ROI_DIM) if roi else {}
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn, **kw)
    st = BsplineHip(bxf, fixed, moving)
    oc = oracle.Case(pair.fixed, pair.moving, pair.spacing, pair.vox_per_rgn,
                     **({"roi_offset": ROI_OFF, "roi_dim": ROI_DIM} if roi else {}))
    return pair, bxf, st, oc


def dev_tensors(n):
    import torch
    dev = torch.device("cuda:0")
    return (torch.zeros(1, dtype=torch.float64, device=dev), torch.zeros(n, dtype=torch.float64, device=dev),
            torch.zeros(1, dtype=torch.int64, device=dev))


def test_displacement_field_matches_oracle(hip_lib):
    import torch
    pair, bxf, st, oc = interior_stage()
    coeff = synth.random_coeff(ROI_DIM, pair.vox_per_rgn, 2.0, seed=3)
    d_c = torch.from_numpy(coeff).cuda()
    vf = torch.full((st.fixed_voxels * 3,), 7.0, dtype=torch.float32, device="cuda")
    st.interpolate_vf(d_c, vf)
    torch.cuda.synchronize()
    ref = oracle.interp_vf(oc.g, coeff)
    got = vf.cpu().numpy().reshape(ref.shape)
    # the ROI carries the displacement, everything else is zero
    mask = np.zeros(ref.shape[:3], bool)
    (ox, oy, oz), (dx, dy, dz) = ROI_OFF, ROI_DIM
    mask[oz:oz + dz, oy:oy + dy, ox:ox + dx] = True
    assert np.all(got[~mask] == 0.0)
    err = np.abs(got[mask] - ref[mask]).max()
    print("field max abs err (mm)", err, "max |d|", np.abs(ref).max())
    assert err < 1e-5 * max(1.0, np.abs(ref).max())
    st.close()


def test_score_only_and_stored_field_agree_with_full_eval(hip_lib):
    import torch
    pair, bxf, st, oc = interior_stage(roi=False)
    coeff = torch.from_numpy(synth.random_coeff(pair.dim, pair.vox_per_rgn, 2.0, seed=4)).cuda()
    s1, g1, n1 = dev_tensors(bxf.num_coeff)
    st.score_device(coeff, s1, g1, n1, "mse", "exact")
    s0, _, n0 = dev_tensors(bxf.num_coeff)
    st.score_device(coeff, s0, None, n0, "mse", "none")
    vf = torch.empty(st.fixed_voxels * 3, dtype=torch.float32, device="cuda")
    st.interpolate_vf(coeff, vf)
    s2, g2, n2 = dev_tensors(bxf.num_coeff)
    st.score_vf_device(vf, s2, g2, n2)
    torch.cuda.synchronize()
    # the score-only pass is the full pass without its gradient: same bits
    assert s0.item() == s1.item() and n0.item() == n1.item()
    # the stored field regroups the moving-index arithmetic: float rounding
    assert abs(n2.item() - n1.item()) <= 2
    assert rel_scalar(s2.item(), s1.item()) < 1e-5
    assert rel(g2.cpu().numpy(), g1.cpu().numpy()) < 1e-4
    st.close()


@pytest.mark.parametrize("mode", ["recompute", "stored"])
def test_autograd_adjoint_matches_reference_flavor(hip_lib, mode):
    import torch
    from plastimatch_amd.autograd import mse
    pair, bxf, st, oc = interior_stage(roi=False, seed=11)
    c = synth.random_coeff(pair.dim, pair.vox_per_rgn, 2.0, seed=12)
    coeff = torch.from_numpy(c).cuda().requires_grad_(True)
    score = mse(coeff, st, mode)
    score.backward()
    torch.cuda.synchronize()
    s_ref, g_ref, n_ref, _ = oc.mse(c, "p")
    s64, g64, _ = oc.mse_f64(c, 1)
    assert_parity(score.item(), coeff.grad.cpu().numpy().astype(np.float64), (s_ref, g_ref), (s64, g64),
                  label=f"autograd {mode}:")
    st.close()


def unclamped_coefficients(oc, coeff, count, seed):
    """count random coefficient indices whose support holds no clamped
    voxel (and at most as many of those that do, for the report)."""
    rng = np.random.default_rng(seed)
    inner, border = [], []
    for i in rng.permutation(oc.num_coeff):
        _, _, nclamp = oc.mse_f64_support(coeff, int(i), 0.0)
        (inner if nclamp == 0 else border).append(int(i))
        if len(inner) == count:
            break
    return np.array(inner), np.array(border[:count])


def adjoint_vs_fd(st, oc, c, mode, count, seed, label, h=1e-3):
    import torch
    from plastimatch_amd.autograd import mse
    coeff = torch.from_numpy(c).cuda().requires_grad_(True)
    score = mse(coeff, st, mode)
    score.backward()
    torch.cuda.synchronize()
    g = coeff.grad.cpu().numpy().astype(np.float64)
    inner, border = unclamped_coefficients(oc, c, count, seed)
    assert len(inner) == count
    fd, b = oc.fd_mse_f64(c, inner, h)
    assert np.all(b == 0)
    err = rel(g[inner], fd)
    msg = f"{label} {mode}: adjoint vs FP64 central differences (h = {h} mm) over {count} unclamped coefficients {err:.2e}"
    if len(border):
        fdb, _ = oc.fd_mse_f64(c, border, h)
        _, g64, _ = oc.mse_f64(c, 1)
        msg += (f"; over {len(border)} coefficients with clamped voxels: adjoint vs FD {rel(g[border], fdb):.2e}, "
                f"adjoint vs the FP64 flavor p gradient {rel(g[border], g64[border]):.2e}")
    print(msg)
    assert err < 1e-4, err
    return g


@pytest.mark.parametrize("mode", ["recompute", "stored"])
def test_autograd_matches_central_differences(hip_lib, mode):
    """The adjoint against central differences of the FP64 score on 40
    coefficients of an interior ROI (no clamped voxel).  The step is 1e-4
    mm: the trilinear interpolant is only piecewise smooth, and with a few
    thousand voxels per coefficient a voxel whose position crosses a cell
    boundary within a 1e-3 mm step already moves a difference by ~1e-4."""
    pair, bxf, st, oc = interior_stage(seed=5)
    c = synth.random_coeff(ROI_DIM, pair.vox_per_rgn, 1.0, seed=9)
    adjoint_vs_fd(st, oc, c, mode, 40, 3, "32x30x24, ROI 24x22x16", h=1e-4)
    st.close()


def test_autograd_full_size_192x192x96(hip_lib):
    """The differentiable path at 192x192x96, 8-voxel grid: both modes give
    the reference flavor p's score and gradient (1e-5 / 1e-4 of FP64), and
    the adjoint matches central differences of the FP64 score on 200 random
    unclamped coefficients (1e-4)."""
    import torch
    from plastimatch_amd.autograd import mse
    pair = synth.make_config("adjoint_192x192x96_vpr8", seed=0)
    fixed = Volume(pair.fixed, pair.spacing)
    moving = Volume(pair.moving, pair.spacing)
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    oc = oracle.Case(pair.fixed, pair.moving, pair.spacing, pair.vox_per_rgn)
    c = synth.random_coeff(pair.dim, pair.vox_per_rgn, 2.0, seed=44)
    s_ref, g_ref, n_ref, _ = oc.mse(c, "p")
    s64, g64, _ = oc.mse_f64(c, 1)
    with BsplineHip(bxf, fixed, moving) as st:
        grads = {}
        for mode in ("recompute", "stored"):
            coeff = torch.from_numpy(c).cuda().requires_grad_(True)
            score = mse(coeff, st, mode)
            score.backward()
            torch.cuda.synchronize()
            grads[mode] = coeff.grad.cpu().numpy().astype(np.float64)
            assert_parity(score.item(), grads[mode], (s_ref, g_ref), (s64, g64),
                          label=f"192x192x96 {mode}:")
        assert rel(grads["stored"], grads["recompute"]) < 1e-5
        adjoint_vs_fd(st, oc, c, "recompute", 200, 7, "192x192x96")
