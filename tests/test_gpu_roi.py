"""Regions of interest on the device: MSE flavor i (flavor r for the exact
gradient) and MI flavor h, through plm_hip_bspline_set_roi, against the
oracle with the same ROIs (src/register/bspline_mse.cxx:82-290, 1628-1862;
src/register/bspline_mi.cxx:1232-1496; inside_roi,
src/util/bspline_correspond.cxx:52-74, 113-145).

The fixed ROI is tested at each fixed voxel's float position exactly as the
reference does, so its decisions equal the float oracle's voxel for voxel;
MI pass 1 also tests the moving ROI at the reference's float moving
position, so the histograms stay identical bin for bin.  The tile kernels
test the moving ROI at their own moving-index coordinates (mapped to ROI
index coordinates in a tile-local frame), which can differ from the
reference's float position by ~1e-6 voxel: a decision flips only for a
voxel that lands that close to a ROI boundary.  The ROIs here are
ellipsoids on grids of their own, offset from the image grids by a
fraction of a voxel (as a structure rasterised at another resolution is),
so no voxel centre sits on a ROI voxel boundary at zero displacement; the
bars are the same fixed 1e-5 / 1e-4 as without ROIs, and the counts must
agree with FP64.
"""
import numpy as np
import pytest

import oracle
from plastimatch_amd import synth
from plastimatch_amd.bspline import (BsplineBatch, BsplineHip, Bspline_xform, Joint_histogram, Volume,
                                     stage_for_pair, stage_rois)
from tests.helpers import assert_count, assert_parity, assert_parity_mi, oracle_case, rel, rel_scalar, small_pair

pytestmark = pytest.mark.gpu

IDENTITY = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


def ellipsoid(dim, spacing, origin, center, radii, dc=IDENTITY):
    """uint8 ROI (z, y, x) on the grid origin + dc diag(spacing) ijk, 1
    inside the ellipsoid of the given centre and semi-axes (mm)."""
    nx, ny, nz = dim
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    m = np.asarray(dc, np.float64).reshape(3, 3) * np.asarray(spacing, np.float64)[None, :]
    p = [origin[a] + m[a, 0] * i + m[a, 1] * j + m[a, 2] * k for a in range(3)]
    q = sum(((p[a] - center[a]) / radii[a]) ** 2 for a in range(3))
    return (q < 1.0).astype(np.uint8)


def rois_for(pair):
    """A fixed ROI on a 2x coarser grid offset by a quarter of its voxel
    and a moving ROI on a grid of 1.5x the spacing, offset by 0.3 voxel."""
    nx, ny, nz = pair.dim
    sp = pair.spacing
    ext = [sp[a] * pair.dim[a] for a in range(3)]
    c = [pair.origin[a] + 0.5 * ext[a] for a in range(3)]
    fsp = tuple(2.0 * s for s in sp)
    forg = tuple(pair.origin[a] - 0.75 * fsp[a] for a in range(3))
    fdim = tuple(pair.dim[a] // 2 + 2 for a in range(3))
    froi = ellipsoid(fdim, fsp, forg, c, [0.42 * e for e in ext])
    msp = tuple(1.5 * s for s in sp)
    morg = tuple(pair.origin[a] - 0.7 * msp[a] for a in range(3))
    mdim = tuple(int(pair.dim[a] / 1.5) + 2 for a in range(3))
    mroi = ellipsoid(mdim, msp, morg, [c[0] + 2.0, c[1] - 1.5, c[2] + 1.0], [0.38 * e for e in ext])
    return (froi, fsp, forg), (mroi, msp, morg)


def roi_volume(r):
    img, sp, org = r[:3]
    dc = r[3] if len(r) > 3 else IDENTITY
    return Volume(img, sp, org, dc)


WHICH = {"fixed": (True, False), "moving": (False, True), "both": (True, True)}


@pytest.mark.parametrize("which", list(WHICH))
@pytest.mark.parametrize("grad_mode", ["ref_nn", "exact"])
def test_mse_roi_matches_oracle(hip_lib, which, grad_mode):
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=31)
    fr, mr = rois_for(pair)
    use_f, use_m = WHICH[which]
    fr, mr = (fr if use_f else None), (mr if use_m else None)
    oc = oracle_case(pair, fixed_roi=fr, moving_roi=mr)
    plain = oracle_case(pair)
    flavor, mode = ("g", 0) if grad_mode == "ref_nn" else ("p", 1)
    with stage_for_pair(pair, moving_grad=Volume(oc.m_grad, pair.spacing, pair.origin)) as st:
        st.set_roi(roi_volume(fr) if fr else None, roi_volume(mr) if mr else None)
        for k, coeff in enumerate([np.zeros(oc.num_coeff, np.float32),
                                   synth.random_coeff(pair.dim, pair.vox_per_rgn, 2.0, seed=41)]):
            r = st.score(coeff, "mse", grad_mode)
            s_ref, g_ref, n_ref, _ = oc.mse(coeff, flavor)
            s64, g64, n64 = oc.mse_f64(coeff, mode)
            n_all = plain.mse_f64(coeff, mode)[2]
            label = f"{which} ROI {grad_mode} coeff {k}:"
            print(f"{label} overlap {n64} of {n_all} without the ROI")
            assert 0.1 * n_all < n64 < 0.9 * n_all, "the ROI must cut the overlap"
            assert_count(r.num_vox, n_ref, n64, label)
            assert_parity(r.score, r.grad, (s_ref, g_ref), (s64, g64), label=label)


def test_mse_all_ones_roi_is_no_roi(hip_lib):
    """An all-ones ROI (on grids of their own) keeps every voxel: the count
    is the ROI-free one and the score and gradient agree to the rounding of
    the interpolation form (a ROI sends interior tiles through the full
    voxel pass, whose trilinear lerps round within a few ulp of the lean
    pass's); clearing the ROIs restores the ROI-free evaluation bit for
    bit."""
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=32)
    coeff = synth.random_coeff(pair.dim, pair.vox_per_rgn, 2.0, seed=42)
    ones_f = Volume(np.ones((16, 20, 22), np.uint8), (2.0, 2.0, 5.0), (-3.0, -2.5, -4.0))
    ones_m = Volume(np.ones((30, 40, 44), np.uint8), (1.0, 1.0, 1.0), (-1.5, -1.5, -1.5))
    for grad_mode in ("ref_nn", "exact", "none"):
        with stage_for_pair(pair) as st:
            a = st.score(coeff, "mse", grad_mode)
            st.set_roi(ones_f, ones_m)
            b = st.score(coeff, "mse", grad_mode)
            st.set_roi()
            c = st.score(coeff, "mse", grad_mode)
        assert c.num_vox == a.num_vox and c.score == a.score, grad_mode
        assert np.array_equal(c.grad, a.grad), grad_mode
        print(f"{grad_mode}: all-ones ROI vs none: score {rel_scalar(b.score, a.score):.2e}, "
              f"grad {rel(b.grad, a.grad) if grad_mode != 'none' else 0.0:.2e}")
        assert b.num_vox == a.num_vox, grad_mode
        assert rel_scalar(b.score, a.score) < 1e-8, grad_mode
        if grad_mode != "none":
            assert rel(b.grad, a.grad) < 1e-6, grad_mode


def test_mse_empty_fixed_roi(hip_lib):
    pair = small_pair(seed=33)
    with stage_for_pair(pair) as st:
        st.set_roi(Volume(np.zeros((4, 4, 4), np.uint8), (50.0, 50.0, 50.0)))
        r = st.score(np.zeros(pair.true_coeff.size, np.float32), "mse", "ref_nn")
    assert r.num_vox == 0 and r.score == np.finfo(np.float32).max and not np.any(r.grad)


@pytest.mark.parametrize("which", list(WHICH))
def test_mi_roi_matches_oracle(hip_lib, which):
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=34)
    pair.moving = synth.ct_remap(pair.moving.astype(np.float64)).astype(np.float32)
    fr, mr = rois_for(pair)
    use_f, use_m = WHICH[which]
    fr, mr = (fr if use_f else None), (mr if use_m else None)
    oc = oracle_case(pair, fixed_roi=fr, moving_roi=mr)
    bins = 24
    hp = oc.hist(bins)
    nj = bins * bins
    with stage_for_pair(pair, with_grad=False, bins=bins) as st:
        st.set_roi(roi_volume(fr) if fr else None, roi_volume(mr) if mr else None)
        for k, coeff in enumerate([np.full(oc.num_coeff, 0.01, np.float32),
                                   synth.random_coeff(pair.dim, pair.vox_per_rgn, 2.0, seed=43)]):
            r = st.score(coeff, "mi", want_hist=True)
            o = oc.mi(coeff, hp)
            s64, g64, n64 = oc.mi_f64(coeff, hp)
            sh, gh, _ = oc.mi_f64(coeff, hp, hist=o)
            label = f"MI {which} ROI coeff {k}:"
            print(f"{label} overlap {o['num_vox']} (FP64 {n64})")
            assert r.num_vox == o["num_vox"]
            assert np.array_equal(r.hist[:nj], o["fx_j_hist"])
            assert np.array_equal(r.hist[nj:], o["fx_f_hist"])
            assert_parity_mi(r.score, r.grad, (o["score"], o["grad"]), (s64, g64), (sh, gh), label=label)


def test_mi_roi_on_vopt_bins(hip_lib):
    """Flavor h on V-optimal bins: both ROIs and the key tables together
    (the bins are set up over the whole images, as Joint_histogram::
    initialize does whatever the ROIs): the joint and fixed histograms bin
    for bin, N, and the score and gradient against the oracle."""
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=36)
    pair.moving = synth.ct_remap(pair.moving.astype(np.float64)).astype(np.float32)
    fr, mr = rois_for(pair)
    oc = oracle_case(pair, fixed_roi=fr, moving_roi=mr)
    bins = 20
    hp = oc.hist(bins, vopt=True)
    nj = bins * bins
    with stage_for_pair(pair, with_grad=False, bins=bins, hist_type="vopt") as st:
        assert np.array_equal(st.hist.fixed_keys, hp._luts[0])
        assert np.array_equal(st.hist.moving_keys, hp._luts[1])
        st.set_roi(roi_volume(fr), roi_volume(mr))
        for k, coeff in enumerate([np.full(oc.num_coeff, 0.01, np.float32),
                                   synth.random_coeff(pair.dim, pair.vox_per_rgn, 2.0, seed=44)]):
            r = st.score(coeff, "mi", want_hist=True)
            o = oc.mi(coeff, hp)
            s64, g64, _ = oc.mi_f64(coeff, hp)
            sh, gh, _ = oc.mi_f64(coeff, hp, hist=o)
            assert r.num_vox == o["num_vox"]
            assert np.array_equal(r.hist[:nj], o["fx_j_hist"])
            assert np.array_equal(r.hist[nj:], o["fx_f_hist"])
            assert_parity_mi(r.score, r.grad, (o["score"], o["grad"]), (s64, g64), (sh, gh),
                             label=f"MI ROI on V-optimal bins, coeff {k}:")


ROT = dict(f_sp=(1.0, 1.1, 2.3), f_org=(-3.0, 5.0, 1.5),
           f_dc=(0.9927, -0.1197, 0.0147, 0.1192, 0.9922, 0.0346, -0.0187, -0.0326, 0.9993),
           m_sp=(0.95, 1.05, 2.2), m_org=(-2.0, 4.0, 0.5))


def test_mi_fixed_roi_rotated_walk_positions(hip_lib):
    """A rotated, non-dyadic fixed image: pass 1 takes the reference's
    LOOP_Z/Y/X walk positions, so its fixed-ROI decisions are taken there
    (a mask of their own), and pass 2's at the direct positions.  The ROI
    lies on the fixed image's own grid, where voxel centres fall on ROI
    voxel boundaries up to float rounding: the reference decides those ties
    by its float positions, and the histograms must still equal the float
    oracle's bin for bin."""
    dc = tuple(float(v) for v in np.asarray(ROT["f_dc"], np.float32))
    pair = synth.make_pair("rot", (44, 38, 30), ROT["f_sp"], (6, 5, 4), 1.5, 12, seed=35, ct=True)
    fixed = Volume(pair.fixed, ROT["f_sp"], ROT["f_org"], dc)
    moving = Volume(pair.moving, ROT["m_sp"], ROT["m_org"], dc)
    ext = [ROT["f_sp"][a] * pair.dim[a] for a in range(3)]
    froi = ellipsoid(pair.dim, ROT["f_sp"], ROT["f_org"],
                     [ROT["f_org"][a] + 0.45 * ext[a] for a in range(3)], [0.4 * e for e in ext], dc)
    oc = oracle.Case(pair.fixed, pair.moving, ROT["f_sp"], pair.vox_per_rgn, origin=ROT["f_org"], dc=dc,
                     m_spacing=ROT["m_sp"], m_origin=ROT["m_org"], m_dc=dc,
                     fixed_roi=(froi, ROT["f_sp"], ROT["f_org"], dc))
    assert oracle.pass1_position_mismatch(pair.fixed, ROT["f_sp"], oc.g, ROT["f_org"], dc) > 0
    bins = 20
    hp = oc.hist(bins)
    hist = Joint_histogram(bins, bins).initialize(fixed, moving)
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    with BsplineHip(bxf, fixed, moving, hist=hist) as st:
        st.set_roi(Volume(froi, ROT["f_sp"], ROT["f_org"], dc))
        coeff = synth.random_coeff(pair.dim, pair.vox_per_rgn, 1.5, seed=44)
        r = st.score(coeff, "mi", want_hist=True)
    o = oc.mi(coeff, hp)
    nj = bins * bins
    assert r.num_vox == o["num_vox"] and o["num_vox"] < 0.8 * pair.fixed.size
    assert np.array_equal(r.hist[:nj], o["fx_j_hist"])
    assert np.array_equal(r.hist[nj:], o["fx_f_hist"])
    # Pass 2's fixed-ROI decisions are the float oracle's too (the mask is
    # taken at the same float positions), so the score and gradient are held
    # to the float flavor.  FP64 decides the ties by exact positions -- a few
    # percent of the ROI's voxels go the other way -- so it is no yardstick
    # here; its distance is printed.
    sh, gh, _ = oc.mi_f64(coeff, hp, hist=o)
    es, eg = rel_scalar(r.score, o["score"]), rel(r.grad, o["grad"])
    print(f"rotated fixed ROI MI: vs float flavor score {es:.2e} grad {eg:.2e}; float flavor vs FP64 "
          f"with its ties: score {rel_scalar(o['score'], sh):.2e} grad {rel(o['grad'], gh):.2e}")
    assert es < 1e-5 and eg < 1e-4


def test_mi_intensity_range_rois(hip_lib):
    """MI with the user's intensity ranges (flavor h without ROI files):
    stage_rois turns them into ROIs on each image's grid (update_roi), the
    histogram keeps spanning the whole images."""
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=36)
    pair.moving = synth.ct_remap(pair.moving.astype(np.float64)).astype(np.float32)
    fixed = Volume(pair.fixed, pair.spacing, pair.origin)
    moving = Volume(pair.moving, pair.spacing, pair.origin)
    flo, fhi = np.percentile(pair.fixed, [20, 95])
    mlo, mhi = np.percentile(pair.moving, [10, 90])
    fr, mr = stage_rois(fixed, moving, None, None, float(flo), float(fhi), float(mlo), float(mhi))
    assert np.array_equal(fr.img, ((pair.fixed >= np.float32(flo)) & (pair.fixed <= np.float32(fhi))))
    oc = oracle_case(pair, fixed_roi=(fr.img, fr.spacing, fr.origin), moving_roi=(mr.img, mr.spacing, mr.origin))
    bins = 24
    hp = oc.hist(bins)
    with stage_for_pair(pair, with_grad=False, bins=bins) as st:
        st.set_roi(fr, mr)
        coeff = synth.random_coeff(pair.dim, pair.vox_per_rgn, 1.5, seed=45)
        r = st.score(coeff, "mi", want_hist=True)
    o = oc.mi(coeff, hp)
    s64, g64, _ = oc.mi_f64(coeff, hp)
    sh, gh, _ = oc.mi_f64(coeff, hp, hist=o)
    nj = bins * bins
    assert r.num_vox == o["num_vox"] and r.num_vox < 0.8 * pair.fixed.size
    assert np.array_equal(r.hist[:nj], o["fx_j_hist"])
    assert_parity_mi(r.score, r.grad, (o["score"], o["grad"]), (s64, g64), (sh, gh), label="range ROIs")


def test_batch_with_roi_matches_single(hip_lib):
    """A batch holding a context with ROIs and one without gives each what
    its own single-context evaluation gives (gradient bits, count; score to
    summation order)."""
    import torch
    dev = torch.device("cuda:0")
    pairs = [small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 1.5), vpr=(6, 6, 6), seed=37 + i)
             for i in range(2)]
    stages = [stage_for_pair(p) for p in pairs]
    fr, mr = rois_for(pairs[0])
    stages[0].set_roi(roi_volume(fr), roi_volume(mr))
    coeffs = [torch.from_numpy(synth.random_coeff(p.dim, p.vox_per_rgn, 2.0, seed=60 + i)).to(dev)
              for i, p in enumerate(pairs)]
    single = []
    for st, c in zip(stages, coeffs):
        s = torch.zeros(1, dtype=torch.float64, device=dev)
        g = torch.zeros(c.numel(), dtype=torch.float64, device=dev)
        nv = torch.zeros(1, dtype=torch.int64, device=dev)
        st.score_device(c, s, g, nv, "mse", "ref_nn")
        torch.cuda.synchronize()
        single.append((float(s.item()), g.cpu().numpy(), int(nv.item())))
    grads = [torch.zeros(c.numel(), dtype=torch.float64, device=dev) for c in coeffs]
    score = torch.zeros(2, dtype=torch.float64, device=dev)
    nvox = torch.zeros(2, dtype=torch.int64, device=dev)
    with BsplineBatch(stages) as b:
        b.set_buffers(coeffs, grads, score, nvox)
        b.score_device("ref_nn")
        torch.cuda.synchronize()
        for i, (s1, g1, n1) in enumerate(single):
            assert int(nvox[i].item()) == n1, i
            assert np.array_equal(grads[i].cpu().numpy(), g1), i
            assert rel_scalar(float(score[i].item()), s1) < 1e-12, i
    assert single[0][2] < 0.9 * single[1][2]
    for st in stages:
        st.close()


def test_roi_argument_errors(hip_lib):
    import torch
    pair = small_pair(seed=38)
    with stage_for_pair(pair) as st:
        bad = Volume(np.zeros((4, 4, 4), np.float32), (1.0, 1.0, 1.0))
        with pytest.raises(ValueError):
            st.set_roi(bad)
        st.set_roi(Volume(np.ones((4, 4, 4), np.uint8), (20.0, 20.0, 20.0)))
        vf = torch.zeros(3 * st.fixed_voxels, dtype=torch.float32, device="cuda")
        s = torch.zeros(1, dtype=torch.float64, device="cuda")
        g = torch.zeros(pair.true_coeff.size, dtype=torch.float64, device="cuda")
        nv = torch.zeros(1, dtype=torch.int64, device="cuda")
        with pytest.raises(RuntimeError, match="region of interest"):
            st.score_vf_device(vf, s, g, nv)
