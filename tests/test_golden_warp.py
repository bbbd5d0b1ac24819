"""The reference's own warp tests, reproduced: plm-warp-b and plm-warp-d
(Testing/CMakeLists.txt:4889-4960) warp the synthetic gauss-1 and
sphere-1 images (Testing/CMakeLists.txt:369-371, 637-640) with the
reference's B-spline file Testing/Data/xf-bspline-2.txt (a copy in
tests/golden/) through the native warper, and check

  plm-warp-b: mean x displacement of the output field in [5.77, 5.78],
              mean of the warped gauss image (trilinear) in [-775.7, -775.6];
  plm-warp-d: mean of the warped sphere image (nearest voxel) in [9.2, 9.3].

Plm_warp moves the transform onto the image's own grid first
(plm_warp_native -> xform_to_gpuit_bsp, src/base/plm_warp.cxx:189-240):
same origin, grid spacing from the file, so vox_per_rgn = round(19.73685 /
(50/38)) = 15 and the same 6 x 6 x 6 knots; the file's coefficients are
used on that grid (the conversion goes through ITK, absent here, and
moves them by the file's spacing rounding, ~4e-7 relative).

The oracle must land in the reference's intervals (CPU tests); so must
the device warp and displacement field (GPU tests).
"""
import os

import numpy as np
import pytest

import oracle
from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume
from tests.helpers import reference_synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
XF = os.path.join(GOLDEN, "xf-bspline-2.txt")


def image_and_grid(pattern):
    img, sp, org = reference_synth(pattern)
    loaded = Bspline_xform.load(XF)
    vol = Volume(img, sp, org)
    bxf = Bspline_xform.for_volume(vol, (15, 15, 15))
    assert bxf.cdims == loaded.cdims == (6, 6, 6)
    bxf.coeff = loaded.coeff.copy()
    return img, sp, org, vol, bxf


def test_xform_file_round_trip(tmp_path):
    bxf = Bspline_xform.load(XF)
    assert bxf.img_dim == (38, 38, 38) and bxf.vox_per_rgn == (15, 15, 15) and bxf.num_coeff == 648
    out = tmp_path / "xf.txt"
    bxf.save(str(out))
    assert out.read_text() == open(XF).read()
    old = Bspline_xform.load(os.path.join(GOLDEN, "xf-bspline-1.txt"))   # no direction_cosines line
    assert old.img_dim == (4, 4, 4) and old.num_coeff == 192
    assert old.direction_cosines == (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


def test_oracle_reproduces_plm_warp_b():
    img, sp, org, vol, bxf = image_and_grid("gauss")
    g = oracle.geom((38, 38, 38), sp, (15, 15, 15), origin=org)
    warped = oracle.warp(g, img, sp, bxf.coeff, True, 0.0, m_origin=org)
    vf = oracle.interp_vf(g, bxf.coeff)
    ave = float(np.mean(warped.astype(np.float64)))
    mean_x = float(np.mean(vf.reshape(-1, 3)[:, 0].astype(np.float64)))
    print(f"plm-warp-b (oracle): AVE {ave:.4f}, vf mean x {mean_x:.4f}")
    assert -775.7 <= ave <= -775.6
    assert 5.77 <= mean_x <= 5.78


def test_oracle_reproduces_plm_warp_d():
    img, sp, org, vol, bxf = image_and_grid("sphere")
    g = oracle.geom((38, 38, 38), sp, (15, 15, 15), origin=org)
    warped = oracle.warp(g, img, sp, bxf.coeff, False, 0.0, m_origin=org)
    ave = float(np.mean(warped.astype(np.uint8).astype(np.float64)))
    print(f"plm-warp-d (oracle): AVE {ave:.4f}")
    assert 9.2 <= ave <= 9.3


@pytest.mark.gpu
def test_device_reproduces_plm_warp_b(hip_lib):
    import torch
    img, sp, org, vol, bxf = image_and_grid("gauss")
    with BsplineHip(bxf, vol, vol) as st:
        d_c = torch.from_numpy(bxf.coeff).cuda()
        out = torch.empty(st.fixed_voxels, dtype=torch.float32, device="cuda")
        vf = torch.empty(3 * st.fixed_voxels, dtype=torch.float32, device="cuda")
        st.warp(d_c, out, 0.0, True)
        st.interpolate_vf(d_c, vf)
        torch.cuda.synchronize()
    ave = float(out.double().mean())
    mean_x = float(vf.view(-1, 3)[:, 0].double().mean())
    print(f"plm-warp-b (device): AVE {ave:.4f}, vf mean x {mean_x:.4f}")
    assert -775.7 <= ave <= -775.6
    assert 5.77 <= mean_x <= 5.78


@pytest.mark.gpu
def test_device_reproduces_plm_warp_d(hip_lib):
    import torch
    img, sp, org, vol, bxf = image_and_grid("sphere")
    with BsplineHip(bxf, vol, vol) as st:
        out = torch.empty(st.fixed_voxels, dtype=torch.float32, device="cuda")
        st.warp(torch.from_numpy(bxf.coeff).cuda(), out, 0.0, False)
        torch.cuda.synchronize()
    ave = float(out.cpu().numpy().astype(np.uint8).astype(np.float64).mean())
    print(f"plm-warp-d (device): AVE {ave:.4f}")
    assert 9.2 <= ave <= 9.3
