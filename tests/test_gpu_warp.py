"""The warped moving image (plm_hip_bspline_warp) against the oracle's
restatement of bspline_warp_dcos (src/base/bspline_warp.cxx:27-167).

The GPU maps a fixed voxel to moving index coordinates with the regrouped
P (position + displacement - origin) of the tile kernel, the oracle with
the reference's float positions (x accumulated step by step) and its q_lut
displacement: the two land within ~1e-5 voxel of each other.  Trilinear
values then agree to that times the local image slope; a nearest-voxel pick
or the inside test flips only for a voxel that sits within that distance of
a rounding or border boundary, so those are counted, not required equal.
A zero field on a shared grid is exact (every voxel on a centre), and
voxels outside the ROI hold the default value.
"""
import numpy as np
import pytest

import oracle
from plastimatch_amd import synth
from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume, bf16_round
from tests.helpers import small_pair

pytestmark = pytest.mark.gpu

FILL = -1234.5


def stage(dim=(40, 36, 28), spacing=(1.0, 1.25, 2.0), vpr=(5, 6, 4), roi=None, storage="f32", seed=2):
    pair = small_pair(dim=dim, spacing=spacing, vpr=vpr, seed=seed)
    fixed = Volume(pair.fixed, pair.spacing)
    moving = Volume(pair.moving, pair.spacing)
    kw = {"roi_offset": roi[0], "roi_dim": roi[1]} if roi else {}
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn, **kw)
    st = BsplineHip(bxf, fixed, moving, storage=storage)
    g = oracle.geom(pair.dim, pair.spacing, pair.vox_per_rgn, **kw)
    return pair, bxf, st, g


def gpu_warp(st, coeff, linear):
    import torch
    out = torch.full((st.fixed_voxels,), 7.0, dtype=torch.float32, device="cuda")
    st.warp(torch.from_numpy(np.ascontiguousarray(coeff, np.float32)).cuda(), out, FILL, linear)
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("linear", [True, False])
def test_warp_zero_field_is_a_copy(hip_lib, linear):
    pair, bxf, st, g = stage()
    got = gpu_warp(st, np.zeros(bxf.num_coeff, np.float32), linear)
    assert np.array_equal(got.reshape(pair.moving.shape), pair.moving)
    st.close()


@pytest.mark.parametrize("linear", [True, False])
@pytest.mark.parametrize("roi", [None, ((3, 4, 2), (31, 27, 22))])
def test_warp_matches_oracle(hip_lib, linear, roi):
    pair, bxf, st, g = stage(roi=roi)
    coeff = synth.random_coeff(roi[1] if roi else pair.dim, pair.vox_per_rgn, 3.0, seed=5)
    got = gpu_warp(st, coeff, linear).reshape(pair.moving.shape)
    ref = oracle.warp(g, pair.moving, pair.spacing, coeff, linear, FILL)
    d_got, d_ref = got == FILL, ref == FILL
    flips = np.count_nonzero(d_got != d_ref)
    both = ~d_got & ~d_ref
    n = got.size
    print(f"linear={linear} roi={roi}: default voxels {d_ref.sum()}, inside-test flips {flips}")
    assert flips <= max(2, n // 20000)
    if roi:
        (ox, oy, oz), (dx, dy, dz) = roi
        mask = np.ones(got.shape, bool)
        mask[oz:oz + dz, oy:oy + dy, ox:ox + dx] = False
        assert np.all(got[mask] == FILL)
    span = float(pair.moving.max() - pair.moving.min())
    if linear:
        err = np.abs(got[both].astype(np.float64) - ref[both])
        print(f"  max |err| {err.max():.3e} of range {span:.1f}, rms {np.sqrt((err ** 2).mean()):.3e}")
        assert err.max() < 1e-3 * span
        assert np.sqrt((err ** 2).mean()) < 1e-5 * span
    else:
        picks = np.count_nonzero(got[both] != ref[both])
        print(f"  nearest-voxel flips {picks} of {both.sum()}")
        assert picks <= max(2, n // 5000)
    st.close()


def test_warp_bf16_storage(hip_lib):
    """BF16 storage warps the BF16-rounded moving image."""
    pair, bxf, st, g = stage(storage="bf16")
    coeff = synth.random_coeff(pair.dim, pair.vox_per_rgn, 3.0, seed=6)
    got = gpu_warp(st, coeff, True).reshape(pair.moving.shape)
    ref = oracle.warp(g, bf16_round(pair.moving), pair.spacing, coeff, True, FILL)
    both = (got != FILL) & (ref != FILL)
    span = float(pair.moving.max() - pair.moving.min())
    assert np.sqrt(((got[both].astype(np.float64) - ref[both]) ** 2).mean()) < 1e-5 * span
    st.close()
