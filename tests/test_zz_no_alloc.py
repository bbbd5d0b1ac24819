"""The allocation-free MSE entry point (plm_hip_bspline_score_noalloc),
through the C ABI, against plm_hip_bspline_score_device on a context built
from the same device images: the same gradient bits and overlap count and
the score to summation order, for flavors g and p and the score alone, on
two calls in a row with the second on a smaller image and grid (the static
scratch keeps the first call's contents, so a slot the second call failed
to clear would show in its gradient).  Also: a grid beyond the reserved
scratch is refused before anything is launched, and no device memory is
allocated by the calls."""
import ctypes

import numpy as np
import pytest

from plastimatch_amd import _abi, synth
from plastimatch_amd.bspline import GRAD_MODES, BsplineHip, Bspline_xform, Volume, volume_gradient

pytestmark = pytest.mark.gpu


def device_stage(dim, vpr, seed):
    import torch
    pair = synth.make_pair("noalloc", dim, (1.0, 1.0, 2.0), vpr, 2.0, 10, seed=seed, ct=True,
                           origin=(-3.0, 5.0, 1.5))
    dev = torch.device("cuda:0")
    fixed = Volume(torch.from_numpy(pair.fixed).to(dev), pair.spacing, pair.origin)
    moving = Volume(torch.from_numpy(pair.moving).to(dev), pair.spacing, pair.origin)
    mgrad = volume_gradient(moving)
    bxf = Bspline_xform.for_volume(fixed, vpr)
    coeff = torch.from_numpy(synth.random_coeff(pair.dim, vpr, 2.0, seed=seed + 1)).to(dev)
    return fixed, moving, mgrad, bxf, coeff


def outputs(n):
    import torch
    return (torch.zeros(1, dtype=torch.float64, device="cuda"),
            torch.full((n,), 7.0, dtype=torch.float64, device="cuda"),
            torch.zeros(1, dtype=torch.int64, device="cuda"))


def noalloc(lib, bxf, fixed, moving, mgrad, coeff, mode, out):
    s, g, n = out
    fd, md, gd = fixed.desc(), moving.desc(), mgrad.desc()
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    return lib.plm_hip_bspline_score_noalloc(
        ctypes.byref(bxf.geom()), ctypes.byref(fd), ctypes.byref(md), ctypes.byref(gd),
        GRAD_MODES[mode], 0, ptr(coeff), ptr(s), ptr(g) if mode != "none" else None, ptr(n), None)


def test_noalloc_matches_context_on_two_calls(hip_lib):
    import torch
    lib = _abi.load()
    big = device_stage((64, 56, 40), (8, 8, 8), 3)
    small = device_stage((37, 30, 22), (5, 6, 4), 5)
    refs = {}
    for name, (fixed, moving, mgrad, bxf, coeff) in (("big", big), ("small", small)):
        with BsplineHip(bxf, fixed, moving, mgrad) as st:
            st.set_tuning(2)
            for mode in ("ref_nn", "exact", "none"):
                out = outputs(bxf.num_coeff)
                st.score_device(coeff, out[0], out[1] if mode != "none" else None, out[2], "mse", mode)
                torch.cuda.synchronize()
                refs[name, mode] = [t.clone() for t in out]
    outs = {(name, mode): outputs(st[3].num_coeff) for name, st in (("big", big), ("small", small))
            for mode in ("ref_nn", "exact", "none")}

    def one_round():
        # two calls in a row per mode, the second on the smaller grid
        for mode in ("ref_nn", "exact", "none"):
            for name, (fixed, moving, mgrad, bxf, coeff) in (("big", big), ("small", small)):
                out = outs[name, mode]
                _abi.check(noalloc(lib, bxf, fixed, moving, mgrad, coeff, mode, out))
                torch.cuda.synchronize()
                s, g, n = out
                rs, rg, rn = refs[name, mode]
                assert int(n.item()) == int(rn.item()), (name, mode)
                rel = abs(s.item() - rs.item()) / abs(rs.item())
                assert rel < 1e-12, (name, mode, rel)
                if mode != "none":
                    assert torch.equal(g, rg), (name, mode, (g - rg).abs().max().item())

    # The first call on the device loads the library's scratch array (its
    # code object's static memory), and first uses of kernels -- ours and
    # torch's comparisons -- can load code or grow the runtime's own pools;
    # over a second identical round the calls take no device memory.
    one_round()
    torch.cuda.synchronize()
    taken = lambda: -torch.cuda.mem_get_info()[0] - torch.cuda.memory_reserved()
    mem0 = taken()
    one_round()
    torch.cuda.synchronize()
    assert taken() == mem0, taken() - mem0


def test_noalloc_refuses_a_grid_beyond_the_scratch(hip_lib):
    import torch
    lib = _abi.load()
    fixed, moving, mgrad, _, _ = device_stage((64, 56, 40), (8, 8, 8), 7)
    # one-voxel regions: 67 x 59 x 43 knots need ~130 MB of condense slots
    bxf = Bspline_xform.for_volume(fixed, (1, 1, 1))
    coeff = torch.zeros(bxf.num_coeff, dtype=torch.float32, device="cuda")
    out = outputs(bxf.num_coeff)
    rc = noalloc(lib, bxf, fixed, moving, mgrad, coeff, "ref_nn", out)
    assert rc == _abi.PLM_HIP_ERR_UNSUPPORTED
    assert b"scratch" in lib.plm_hip_last_error()
    torch.cuda.synchronize()
    assert out[1][0].item() == 7.0
