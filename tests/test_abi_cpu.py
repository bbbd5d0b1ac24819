"""CPU checks of the C ABI library and the host side (no GPU, no compute)."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from plastimatch_amd import _abi, synth
from plastimatch_amd.bspline import Bspline_xform, Volume

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "plm_hip_bspline.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(plm_hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _abi.load()
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n
        assert n in _abi.SIGNATURES, n


def test_abi_version():
    assert _abi.load().plm_hip_abi_version() == 1


def test_argument_errors_without_device():
    lib = _abi.load()
    ctx = ctypes.c_void_p()
    rc = lib.plm_hip_bspline_create(None, None, None, None, None, 0, ctypes.byref(ctx))
    assert rc == _abi.PLM_HIP_ERR_ARG
    assert b"NULL" in lib.plm_hip_last_error()
    assert lib.plm_hip_bspline_score(None, 0, 0, None, None, None, None, None) == _abi.PLM_HIP_ERR_ARG
    assert lib.plm_hip_bspline_destroy(None) == _abi.PLM_HIP_OK
    # the batch entry points check their arguments before any HIP call
    b = ctypes.c_void_p()
    assert lib.plm_hip_batch_create(None, 1, ctypes.byref(b)) == _abi.PLM_HIP_ERR_ARG
    arr = (ctypes.c_void_p * 1)(None)
    assert lib.plm_hip_batch_create(arr, 1, ctypes.byref(b)) == _abi.PLM_HIP_ERR_ARG
    assert lib.plm_hip_batch_create(arr, 65, ctypes.byref(b)) == _abi.PLM_HIP_ERR_ARG
    assert b.value is None
    assert lib.plm_hip_batch_score_device(None, 0, None) == _abi.PLM_HIP_ERR_ARG
    assert lib.plm_hip_batch_launch(None, None) == _abi.PLM_HIP_ERR_ARG
    assert lib.plm_hip_batch_set_tuning(None, 2) == _abi.PLM_HIP_ERR_ARG
    assert lib.plm_hip_batch_set_occupancy(None, 2) == _abi.PLM_HIP_ERR_ARG
    assert lib.plm_hip_batch_get_tuning(None, None, None) == _abi.PLM_HIP_ERR_ARG
    assert lib.plm_hip_bspline_set_occupancy(None, 2) == _abi.PLM_HIP_ERR_ARG
    assert lib.plm_hip_bspline_get_tuning(None, None, None) == _abi.PLM_HIP_ERR_ARG
    assert lib.plm_hip_batch_destroy(None) == _abi.PLM_HIP_OK


def test_bad_geometry_rejected_before_device():
    lib = _abi.load()
    img = np.zeros((8, 8, 8), np.float32)
    v = Volume(img, (1.0, 1.0, 1.0))
    geom = Bspline_xform.for_volume(v, (4, 4, 4)).geom()
    geom.vox_per_rgn[0] = 0
    ctx = ctypes.c_void_p()
    d = v.desc()
    rc = lib.plm_hip_bspline_create(ctypes.byref(geom), ctypes.byref(d), ctypes.byref(d), None,
                                    None, 0, ctypes.byref(ctx))
    assert rc == _abi.PLM_HIP_ERR_GEOMETRY


@pytest.mark.parametrize("name", list(synth.CONFIGS))
def test_host_grid_matches_oracle(name):
    cfg = synth.CONFIGS[name]
    v = Volume(np.zeros((2, 2, 2), np.float32), cfg["spacing"])
    bxf = Bspline_xform(v.origin, cfg["spacing"], cfg["dim"], cfg["vox_per_rgn"])
    r, c, nk = oracle.header(oracle.geom(cfg["dim"], cfg["spacing"], cfg["vox_per_rgn"]))
    assert bxf.rdims == tuple(r) and bxf.cdims == tuple(c) and bxf.num_knots == nk
    assert bxf.num_coeff == 3 * nk


def test_jitter_if_zero():
    bxf = Bspline_xform((0, 0, 0), (1, 1, 1), (20, 20, 20), (5, 5, 5))
    bxf.jitter_if_zero()
    assert np.all(bxf.coeff == np.float32(0.01))
    bxf.coeff[3] = 2.0
    bxf.jitter_if_zero()
    assert bxf.coeff[3] == 2.0


def test_synth_is_seeded():
    a = synth.make_pair("a", (20, 18, 16), (1, 1, 2), (5, 5, 5), 2.0, 5, seed=3)
    b = synth.make_pair("b", (20, 18, 16), (1, 1, 2), (5, 5, 5), 2.0, 5, seed=3)
    assert np.array_equal(a.fixed, b.fixed) and np.array_equal(a.moving, b.moving)
    assert np.array_equal(a.true_coeff, b.true_coeff)


def test_bench_rejects_multi_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
