"""The C++ host side (plastimatch_amd/host): Bspline_state / bspline_score()
with threading=hip, driven through bspline_hip_check like the reference's
check_grad tool."""
import json
import os
import subprocess

import numpy as np
import pytest

from plastimatch_amd import synth
from tests.helpers import assert_count, oracle_case, rel, rel_scalar, small_pair

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "plastimatch_amd", "host")
TOOL = os.path.join(HOST, "bspline_hip_check")


def run_tool(args, tmp_path, pair):
    f = os.path.join(tmp_path, "fixed.mha")
    m = os.path.join(tmp_path, "moving.mha")
    synth.write_mha(f, pair.fixed, pair.spacing, pair.origin)
    synth.write_mha(m, pair.moving, pair.spacing, pair.origin)
    r = subprocess.run([TOOL] + args + [f, m], capture_output=True, text=True, timeout=240)
    return r.returncode, json.loads(r.stdout.strip().splitlines()[-1]), r.stderr


def test_host_builds_and_reports_missing_device(tmp_path):
    subprocess.run(["make", "-s", "-C", HOST], check=True)
    assert os.path.exists(TOOL)
    pytest.importorskip("torch")
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present; the GPU tests cover the tool")
    pair = small_pair(dim=(12, 10, 8), vpr=(4, 4, 4))
    rc, out, _ = run_tool(["-g", "4"], str(tmp_path), pair)
    assert rc == 1 and out["status"] < 0 and out["error"]


@pytest.mark.gpu
@pytest.mark.parametrize("flavor", ["g", "p"])
def test_host_mse_matches_oracle(tmp_path, flavor):
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=31)
    oc = oracle_case(pair)
    gfile = os.path.join(str(tmp_path), "grad.raw")
    xfile = os.path.join(str(tmp_path), "coeff.raw")
    coeff = synth.random_coeff(pair.dim, pair.vox_per_rgn, 1.5, seed=11)
    coeff.tofile(xfile)
    rc, out, err = run_tool(["-g", "5", "-f", flavor, "-X", xfile, "-O", gfile], str(tmp_path), pair)
    assert rc == 0, (out, err)
    grad = np.fromfile(gfile, np.float64)
    s, g, n, _ = oc.mse(coeff, flavor)
    assert_count(out["num_vox"], n, oc.mse_f64(coeff, 0 if flavor == "g" else 1)[2], "host")
    assert rel_scalar(out["score"], s) < 1e-5
    assert rel(grad, g) < 1e-4


@pytest.mark.gpu
def test_host_mse_zero_coeff_parity(tmp_path):
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=32)
    oc = oracle_case(pair)
    gfile = os.path.join(str(tmp_path), "grad.raw")
    rc, out, err = run_tool(["-g", "5", "-O", gfile], str(tmp_path), pair)
    assert rc == 0, (out, err)
    s, g, n, _ = oc.mse(np.zeros(oc.num_coeff, np.float32), "g")
    grad = np.fromfile(gfile, np.float64)
    assert_count(out["num_vox"], n, oc.mse_f64(np.zeros(oc.num_coeff, np.float32), 0)[2], "host zero")
    assert rel_scalar(out["score"], s) < 1e-5
    assert rel(grad, g) < 1e-4


@pytest.mark.gpu
def test_host_central_difference(tmp_path):
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 1.0), vpr=(5, 5, 5), seed=33, ct=False)
    rc, out, err = run_tool(["-g", "5", "-f", "p", "-p", "ctr", "-e", "1e-2", "-i", "700",
                             "-R", "-0.5", "0.5"], str(tmp_path), pair)
    assert rc == 0, (out, err)
    assert abs(out["analytic"] - out["finite_difference"]) <= 0.05 * abs(out["analytic"]) + 1e-9


@pytest.mark.gpu
def test_host_mi_and_optimizer(tmp_path):
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=34)
    rc, out, err = run_tool(["-g", "5", "-M", "mi", "--bins", "32"], str(tmp_path), pair)
    assert rc == 0, (out, err)
    oc = oracle_case(pair)
    o = oc.mi(np.full(oc.num_coeff, 0.01, np.float32), oc.hist(32))
    assert rel_scalar(out["score"], o["score"]) < 1e-5
    rc, out, err = run_tool(["-g", "5", "-p", "optimize", "--iterations", "5"], str(tmp_path), pair)
    assert rc == 0, (out, err)
    assert out["final_score"] < out["initial_score"]


def test_host_xform_file_round_trip(tmp_path):
    """The C++ MGH_GPUIT_BSP reader / writer on the reference's own files
    (Testing/Data/xf-bspline-*.txt): the writer reproduces a file the
    reference wrote byte for byte; an older file without a direction-cosines
    line reads with the identity."""
    subprocess.run(["make", "-s", "-C", HOST], check=True)
    out = tmp_path / "xf2.txt"
    r = subprocess.run([TOOL, "--xf-roundtrip", os.path.join(ROOT, "tests/golden/xf-bspline-2.txt"), str(out)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert "num_coeff 648" in r.stdout
    assert out.read_bytes() == open(os.path.join(ROOT, "tests/golden/xf-bspline-2.txt"), "rb").read()
    out1 = tmp_path / "xf1.txt"
    r = subprocess.run([TOOL, "--xf-roundtrip", os.path.join(ROOT, "tests/golden/xf-bspline-1.txt"), str(out1)],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "num_coeff 192" in r.stdout
    lines = out1.read_text().splitlines()
    assert lines[7] == "direction_cosines = 1.000000 0.000000 0.000000 0.000000 1.000000 0.000000 0.000000 0.000000 1.000000"
    assert lines[8:] == open(os.path.join(ROOT, "tests/golden/xf-bspline-1.txt")).read().splitlines()[7:]
    bad = tmp_path / "bad.txt"
    bad.write_text("MGH_GPUIT_BSP <experimental>\nimg_dim = 4 4 4\nnonsense = 1\n0.0\n")
    r = subprocess.run([TOOL, "--xf-roundtrip", str(bad), str(tmp_path / "x.txt")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "bad header line" in r.stderr


@pytest.mark.gpu
def test_host_mi_rois_and_ranges(tmp_path):
    """MI flavor h through the C++ host: a fixed ROI file narrowed by the
    fixed intensity range, and a moving ROI made from the moving range
    (bspline_stage_rois / update_roi), against the oracle with the ROIs the
    stage's rules give."""
    from plastimatch_amd.bspline import Volume, stage_rois
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=35)
    pair.moving = synth.ct_remap(pair.moving.astype(np.float64)).astype(np.float32)
    nz, ny, nx = pair.fixed.shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    box = ((i > 4) & (i < 34) & (j > 3) & (j < 30) & (k > 2) & (k < 25)).astype(np.uint8)
    froi_fn = os.path.join(str(tmp_path), "froi.mha")
    synth.write_mha(froi_fn, box.astype(np.float32), pair.spacing, pair.origin)
    flo, fhi = (float(v) for v in np.percentile(pair.fixed, [15, 97]))
    mlo, mhi = (float(v) for v in np.percentile(pair.moving, [10, 92]))
    gfile = os.path.join(str(tmp_path), "grad.raw")
    rc, out, err = run_tool(["-g", "5", "-M", "mi", "--bins", "24", "--fixed-roi", froi_fn,
                             "--fixed-range", repr(flo), repr(fhi), "--moving-range", repr(mlo), repr(mhi),
                             "-O", gfile], str(tmp_path), pair)
    assert rc == 0, (out, err)
    fixed = Volume(pair.fixed, pair.spacing, pair.origin)
    moving = Volume(pair.moving, pair.spacing, pair.origin)
    fr, mr = stage_rois(fixed, moving, Volume(box, pair.spacing, pair.origin), None,
                        np.float32(flo), np.float32(fhi), np.float32(mlo), np.float32(mhi))
    oc = oracle_case(pair, fixed_roi=(fr.img, fr.spacing, fr.origin), moving_roi=(mr.img, mr.spacing, mr.origin))
    o = oc.mi(np.full(oc.num_coeff, 0.01, np.float32), oc.hist(24))
    assert out["num_vox"] == o["num_vox"] and o["num_vox"] < 0.6 * pair.fixed.size
    assert rel_scalar(out["score"], o["score"]) < 1e-5
    assert rel(np.fromfile(gfile, np.float64), o["grad"]) < 1e-4


@pytest.mark.gpu
def test_host_mi_vopt(tmp_path):
    """histogram_type=vopt through the C++ host: the V-optimal key tables come
    from the plugin (plm_hip_hist_vopt) and reach the context
    (plm_hip_bspline_set_hist_keys); MI against the oracle on the oracle's
    own V-optimal bins."""
    pair = small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=36)
    pair.moving = synth.ct_remap(pair.moving.astype(np.float64)).astype(np.float32)
    gfile = os.path.join(str(tmp_path), "grad.raw")
    rc, out, err = run_tool(["-g", "5", "-M", "mi", "--bins", "16", "--hist", "vopt", "-O", gfile],
                            str(tmp_path), pair)
    assert rc == 0, (out, err)
    oc = oracle_case(pair)
    o = oc.mi(np.full(oc.num_coeff, 0.01, np.float32), oc.hist(16, vopt=True))
    e = oc.mi(np.full(oc.num_coeff, 0.01, np.float32), oc.hist(16))
    assert out["num_vox"] == o["num_vox"]
    assert rel_scalar(out["score"], o["score"]) < 1e-5
    assert rel(np.fromfile(gfile, np.float64), o["grad"]) < 1e-4
    # and not the equal-spacing result
    assert rel_scalar(out["score"], e["score"]) > 1e-3
