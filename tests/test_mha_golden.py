"""MetaImage I/O (plastimatch_amd.mha and the C++ host's read_mha /
write_mha) and the reference's registration example data
(extra/lua_examples/data: img_01.mha, img_03.mha, the B-spline transform
xf.txt registering them and its dense field vf.mha, copied into
tests/golden/reg_example/).

* The reference's files, written by ITK, read back with their geometry.
* The writer's header is the reference's format string field for field
  (src/base/mha_io.cxx:33-47, %g geometry, TransformMatrix by columns).
* Round trips in every element type, big-endian input, vector fields,
  Python <-> C++.
* Golden: bspline_interpolate_vf of xf.txt over its own grid is vf.mha bit
  for bit (the oracle; the device within float rounding, in the GPU test).
"""
import os
import subprocess

import numpy as np
import pytest

import oracle
from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume
from plastimatch_amd.mha import read_mha, write_mha

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EX = os.path.join(ROOT, "tests", "golden", "reg_example")
TOOL = os.path.join(ROOT, "plastimatch_amd", "host", "bspline_hip_check")


def test_reads_the_reference_files():
    im = read_mha(os.path.join(EX, "img_01.mha"))
    assert im.dim == (44, 44, 21) and im.img.dtype == np.float32
    assert im.spacing == (7.8125, 7.8125, 7.8125)
    assert im.origin == tuple(float(np.float32(v)) for v in (-167.969, -167.969, -89.844))
    vf = read_mha(os.path.join(EX, "vf.mha"))
    assert vf.img.shape == (21, 44, 44, 3)


def test_header_is_the_reference_format(tmp_path):
    p = tmp_path / "h.mha"
    dc = (0.0, 1.0, 0.0, -1.0, 0.0, 0.0, 0.0, 0.0, 1.0)
    write_mha(str(p), np.zeros((3, 4, 5), np.int16), (0.5, 1.25, 2.0), (-1.5, 2.25, 100.125), dc)
    head = p.read_bytes()[:-120].decode()
    assert head == ("ObjectType = Image\nNDims = 3\nBinaryData = True\nBinaryDataByteOrderMSB = False\n"
                    "TransformMatrix = 0 -1 0 1 0 0 0 0 1\nOffset = -1.5 2.25 100.125\n"
                    "CenterOfRotation = 0 0 0\nElementSpacing = 0.5 1.25 2\nDimSize = 5 4 3\n"
                    "AnatomicalOrientation = RAI\nElementType = MET_SHORT\nElementDataFile = LOCAL\n")
    back = read_mha(str(p))
    assert back.direction_cosines == dc


@pytest.mark.parametrize("dtype", [np.float32, np.int16, np.uint8])
def test_round_trip_types(tmp_path, dtype):
    rng = np.random.default_rng(1)
    a = (rng.normal(0, 100, (4, 5, 6)) % 200).astype(dtype)
    p = str(tmp_path / "a.mha")
    write_mha(p, a, (1.0, 2.0, 3.0), (4.0, 5.0, 6.0))
    b = read_mha(p)
    assert b.img.dtype == a.dtype and np.array_equal(a, b.img)


def test_big_endian_and_vector_field(tmp_path):
    rng = np.random.default_rng(2)
    v = rng.normal(0, 5, (3, 4, 5, 3)).astype(np.float32)
    p = str(tmp_path / "v.mha")
    write_mha(p, v, (1, 1, 1))
    assert np.array_equal(read_mha(p).img, v)
    # the same field stored big-endian
    raw = open(p, "rb").read()
    k = raw.index(b"ElementDataFile = LOCAL\n") + len(b"ElementDataFile = LOCAL\n")
    hdr = raw[:k].replace(b"BinaryDataByteOrderMSB = False", b"BinaryDataByteOrderMSB = True")
    pb = str(tmp_path / "vb.mha")
    open(pb, "wb").write(hdr + v.astype(">f4").tobytes())
    assert np.array_equal(read_mha(pb).img, v)


def test_cpp_reader_and_writer(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.dirname(TOOL)], check=True)
    for name in ("img_01.mha", "vf.mha"):
        out = str(tmp_path / name)
        r = subprocess.run([TOOL, "--mha-roundtrip", os.path.join(EX, name), out], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 0, r.stderr
        a, b = read_mha(os.path.join(EX, name)), read_mha(out)
        assert np.array_equal(a.img, b.img) and a.origin == b.origin and a.spacing == b.spacing
    # uchar and short inputs widen to float in the C++ reader
    s = (np.arange(60) % 7).astype(np.uint8).reshape(3, 4, 5)
    ps, po = str(tmp_path / "s.mha"), str(tmp_path / "so.mha")
    write_mha(ps, s, (1, 1, 1))
    r = subprocess.run([TOOL, "--mha-roundtrip", ps, po], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and np.array_equal(read_mha(po).img, s.astype(np.float32))


def _example():
    bxf = Bspline_xform.load(os.path.join(EX, "xf.txt"))
    g = oracle.geom(bxf.img_dim, bxf.img_spacing, bxf.vox_per_rgn, origin=bxf.img_origin)
    return bxf, g


def test_oracle_vf_is_the_reference_vf():
    bxf, g = _example()
    vf = read_mha(os.path.join(EX, "vf.mha")).img
    mine = oracle.interp_vf(g, bxf.coeff).reshape(vf.shape)
    assert np.array_equal(mine, vf)


def test_reference_transform_aligns_the_pair():
    bxf, g = _example()
    i1 = read_mha(os.path.join(EX, "img_01.mha")).img
    i3 = read_mha(os.path.join(EX, "img_03.mha")).img
    w = oracle.warp(g, i1, bxf.img_spacing, bxf.coeff, True, -1000.0, m_origin=bxf.img_origin)
    before = np.mean((i3.astype(np.float64) - i1) ** 2)
    after = np.mean((i3.astype(np.float64) - w) ** 2)
    print(f"MSE img_03 vs img_01 {before:.1f}, vs warped img_01 {after:.1f}")
    assert after < 0.65 * before


@pytest.mark.gpu
def test_device_vf_and_warp_of_the_reference_example(hip_lib):
    import torch
    bxf, g = _example()
    i1 = read_mha(os.path.join(EX, "img_01.mha")).img
    vf_ref = read_mha(os.path.join(EX, "vf.mha")).img
    vol = Volume(i1, bxf.img_spacing, bxf.img_origin)
    grid = Bspline_xform.for_volume(vol, bxf.vox_per_rgn)
    with BsplineHip(grid, vol, vol) as st:
        d_c = torch.from_numpy(bxf.coeff).cuda()
        vf = torch.empty(3 * st.fixed_voxels, dtype=torch.float32, device="cuda")
        out = torch.empty(st.fixed_voxels, dtype=torch.float32, device="cuda")
        st.interpolate_vf(d_c, vf)
        st.warp(d_c, out, -1000.0, True)
        torch.cuda.synchronize()
    got = vf.cpu().numpy().reshape(vf_ref.shape)
    err = np.abs(got - vf_ref).max()
    print(f"device vf vs the reference's vf.mha: max |diff| {err:.2e} of max |vf| {np.abs(vf_ref).max():.2f}")
    assert err <= 4e-6 * np.abs(vf_ref).max()
    ref_w = oracle.warp(g, i1, bxf.img_spacing, bxf.coeff, True, -1000.0, m_origin=bxf.img_origin)
    w = out.cpu().numpy().reshape(i1.shape)
    assert np.sqrt(np.mean((w.astype(np.float64) - ref_w) ** 2)) < 1e-5 * 1000.0
