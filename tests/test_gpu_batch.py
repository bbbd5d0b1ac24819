"""Batched evaluation of several stages (the multi-atlas form): one
persistent tile-kernel launch over the flattened tiles of every pair, one
reduce launch, and its hipGraph replay.

Per pair the batch must give what the single-pair entry point gives on the
same inputs: the same gradient bits (the condense and the reduce do not
depend on which workgroup took a tile), the same overlap count, and the
score to summation order (1e-12 relative).  Pair 0 is also checked against
the oracle to the 1e-5 / 1e-4 bars.
"""
import numpy as np
import pytest

from plastimatch_amd import synth
from plastimatch_amd.bspline import BsplineBatch, stage_for_pair
from tests.helpers import assert_count, assert_parity, bf16_round, oracle_case, rel_scalar

pytestmark = pytest.mark.gpu

DIMS = [(40, 36, 28), (33, 29, 31), (52, 24, 20), (25, 41, 22)]


def make_pairs(vpr, seed0, storage="f32"):
    pairs = [synth.make_pair(f"b{i}", d, (1.0, 1.0, 1.5), vpr, 2.0, 8, seed=seed0 + i, ct=True)
             for i, d in enumerate(DIMS)]
    if storage == "bf16":
        for p in pairs:
            p.fixed = bf16_round(p.fixed)
            p.moving = bf16_round(p.moving)
    return pairs


def single_and_batch(pairs, grad_mode, storage="f32", vpt=None):
    import torch
    dev = torch.device("cuda:0")
    stages = [stage_for_pair(p, storage=storage) for p in pairs]
    coeffs = [torch.from_numpy(synth.random_coeff(p.dim, p.vox_per_rgn, 2.0, seed=50 + i)).to(dev)
              for i, p in enumerate(pairs)]
    n = len(pairs)
    single = []
    for st, c in zip(stages, coeffs):
        if vpt:
            st.set_tuning(vpt)
        s = torch.zeros(1, dtype=torch.float64, device=dev)
        g = torch.zeros(c.numel(), dtype=torch.float64, device=dev)
        nv = torch.zeros(1, dtype=torch.int64, device=dev)
        st.score_device(c, s, g, nv, "mse", grad_mode)
        torch.cuda.synchronize()
        single.append((float(s.item()), g.cpu().numpy(), int(nv.item())))
    batch = BsplineBatch(stages)
    if vpt:
        batch.set_tuning(vpt)
    grads = [torch.zeros(c.numel(), dtype=torch.float64, device=dev) for c in coeffs]
    score = torch.zeros(n, dtype=torch.float64, device=dev)
    nvox = torch.zeros(n, dtype=torch.int64, device=dev)
    batch.set_buffers(coeffs, grads, score, nvox)
    return stages, batch, coeffs, grads, score, nvox, single


def check_same(single, grads, score, nvox, label):
    for i, (s1, g1, n1) in enumerate(single):
        g = grads[i].cpu().numpy()
        assert int(nvox[i].item()) == n1, (label, i)
        assert np.array_equal(g, g1), (label, i, np.abs(g - g1).max())
        assert rel_scalar(float(score[i].item()), s1) < 1e-12, (label, i)


@pytest.mark.parametrize("vpr,grad_mode", [((6, 6, 6), "ref_nn"), ((8, 8, 8), "ref_nn"),
                                           ((8, 8, 8), "exact"), ((5, 7, 6), "ref_nn")])
def test_batch_matches_single_pair(hip_lib, vpr, grad_mode):
    import torch
    pairs = make_pairs(vpr, 100)
    stages, batch, coeffs, grads, score, nvox, single = single_and_batch(pairs, grad_mode)
    with batch:
        batch.score_device(grad_mode)
        torch.cuda.synchronize()
        check_same(single, grads, score, nvox, f"vpr {vpr} {grad_mode}")
    # pair 0 against the oracle
    oc = oracle_case(pairs[0])
    c0 = coeffs[0].cpu().numpy()
    flavor = "g" if grad_mode == "ref_nn" else "p"
    s_ref, g_ref, n_ref, _ = oc.mse(c0, flavor)
    s64, g64, n64 = oc.mse_f64(c0, 0 if grad_mode == "ref_nn" else 1)
    assert_count(int(nvox[0].item()), n_ref, n64, f"batch pair 0 vpr {vpr}:")
    assert_parity(float(score[0].item()), grads[0].cpu().numpy(), (s_ref, g_ref), (s64, g64),
                  label=f"batch pair 0 vpr {vpr}:")
    for st in stages:
        st.close()


@pytest.mark.parametrize("vpt", [1, 2, 4])
def test_batch_launch_shapes(hip_lib, vpt):
    import torch
    pairs = make_pairs((6, 6, 6), 200)
    stages, batch, coeffs, grads, score, nvox, single = single_and_batch(pairs, "ref_nn", vpt=vpt)
    with batch:
        batch.score_device("ref_nn")
        torch.cuda.synchronize()
        check_same(single, grads, score, nvox, f"vpt {vpt}")
        # fewer workgroups per CU: the same bits
        batch.set_occupancy(3)
        assert batch.get_tuning() == (vpt, 3)
        batch.score_device("ref_nn")
        torch.cuda.synchronize()
        check_same(single, grads, score, nvox, f"vpt {vpt}, 3 workgroups per CU")
    for st in stages:
        st.close()


def test_batch_bf16_storage(hip_lib):
    import torch
    pairs = make_pairs((8, 8, 8), 300, storage="bf16")
    stages, batch, coeffs, grads, score, nvox, single = single_and_batch(pairs, "ref_nn",
                                                                         storage="bf16")
    with batch:
        batch.score_device("ref_nn")
        torch.cuda.synchronize()
        check_same(single, grads, score, nvox, "bf16")
    for st in stages:
        st.close()


def test_batch_graph_replays(hip_lib):
    """The captured evaluation, replayed several times with eager work in
    between and new coefficients written in place before each replay, gives
    the eager evaluation's results every time."""
    import torch
    pairs = make_pairs((6, 6, 6), 400)
    stages, batch, coeffs, grads, score, nvox, single = single_and_batch(pairs, "ref_nn")
    with batch:
        batch.capture("ref_nn")
        for rep in range(3):
            for i, c in enumerate(coeffs):
                c.copy_(torch.from_numpy(synth.random_coeff(pairs[i].dim, pairs[i].vox_per_rgn, 2.0,
                                                            seed=1000 * rep + i)))
            torch.cuda.synchronize()
            batch.score_device("ref_nn")
            torch.cuda.synchronize()
            eager = ([g.clone() for g in grads], score.clone(), nvox.clone())
            for g in grads:
                g.zero_()
            score.zero_()
            nvox.zero_()
            torch.cuda.synchronize()
            batch.launch()
            torch.cuda.synchronize()
            for i in range(len(pairs)):
                assert torch.equal(grads[i], eager[0][i]), (rep, i)
            assert torch.equal(score, eager[1]) and torch.equal(nvox, eager[2]), rep
    for st in stages:
        st.close()


def test_batch_counted_tiles(hip_lib):
    """Four pairs of about 64^3 at a 4-voxel grid: at one workgroup per CU
    each workgroup walks more than 12 tiles, so the batched launch takes its
    tiles from the XCD counters; with as many workgroups as fit it walks
    them at a fixed stride.  Both give every pair's single-pair gradient
    bits and overlap count and its score to summation order; the counted
    launch gives the same bits on every call and from a captured graph
    replayed three times (each launch leaves the counters zero)."""
    import torch
    dims = [(64, 60, 56), (58, 66, 62), (70, 52, 60), (60, 64, 58)]
    pairs = [synth.make_pair(f"c{i}", d, (1.0, 1.0, 1.0), (4, 4, 4), 2.0, 12, seed=700 + i, ct=True)
             for i, d in enumerate(dims)]
    stages, batch, coeffs, grads, score, nvox, single = single_and_batch(pairs, "ref_nn", vpt=1)
    with batch:
        runs = []
        for w in (0, 1, 1):
            batch.set_occupancy(w)
            batch.score_device("ref_nn")
            torch.cuda.synchronize()
            check_same(single, grads, score, nvox, f"{w} workgroups per CU")
            runs.append((score.clone(), [g.clone() for g in grads]))
        assert torch.equal(runs[1][0], runs[2][0])
        batch.capture("ref_nn")
        for rep in range(3):
            for g in grads:
                g.zero_()
            score.zero_()
            nvox.zero_()
            torch.cuda.synchronize()
            batch.launch()
            torch.cuda.synchronize()
            assert torch.equal(score, runs[1][0]), rep
            for i in range(len(pairs)):
                assert torch.equal(grads[i], runs[1][1][i]), (rep, i)
            assert [int(v) for v in nvox.cpu()] == [r[2] for r in single], rep
    for st in stages:
        st.close()


def test_full_size_batch16_bf16(hip_lib):
    """The multi-atlas workload: 16 pairs with edges drawn from 96..256 per
    axis (10-voxel grid, N(0, 3 mm) fields), BF16 storage, one batched
    evaluation captured in a hipGraph and replayed.  Every pair against the
    oracle on the same BF16-rounded images and BF16-stored moving gradient:
    N equal to FP64's, score within 1e-5 and gradient within 1e-4 of FP64;
    and the replay equal to the eager batched evaluation."""
    import torch
    from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Volume, volume_gradient
    dev = torch.device("cuda:0")
    pairs = synth.make_batch(16, seed=0)
    stages = []
    for p in pairs:
        p.fixed = bf16_round(p.fixed)
        p.moving = bf16_round(p.moving)
        fixed = Volume(torch.from_numpy(p.fixed).to(dev), p.spacing, p.origin)
        moving = Volume(torch.from_numpy(p.moving).to(dev), p.spacing, p.origin)
        bxf = Bspline_xform.for_volume(fixed, p.vox_per_rgn)
        stages.append(BsplineHip(bxf, fixed, moving, volume_gradient(moving), storage="bf16"))
    coeffs = [torch.from_numpy(synth.random_coeff(p.dim, p.vox_per_rgn, 3.0, seed=200 + i)).to(dev)
              for i, p in enumerate(pairs)]
    grads = [torch.zeros(c.numel(), dtype=torch.float64, device=dev) for c in coeffs]
    score = torch.zeros(16, dtype=torch.float64, device=dev)
    nvox = torch.zeros(16, dtype=torch.int64, device=dev)
    with BsplineBatch(stages) as batch:
        batch.set_buffers(coeffs, grads, score, nvox)
        batch.score_device("ref_nn")
        torch.cuda.synchronize()
        eager = ([g.clone() for g in grads], score.clone(), nvox.clone())
        batch.capture("ref_nn")
        for g in grads:
            g.zero_()
        score.zero_()
        torch.cuda.synchronize()
        batch.launch()
        torch.cuda.synchronize()
        for i in range(16):
            assert torch.equal(grads[i], eager[0][i]), i
        assert torch.equal(score, eager[1]) and torch.equal(nvox, eager[2])
    for i, p in enumerate(pairs):
        oc = oracle_case(p)
        oc._m_grad = bf16_round(oc.m_grad)
        c = coeffs[i].cpu().numpy()
        s_ref, g_ref, n_ref, _ = oc.mse(c, "g")
        s64, g64, n64 = oc.mse_f64(c, 0)
        label = f"batch pair {i} {p.dim}:"
        assert_count(int(nvox[i].item()), n_ref, n64, label)
        assert_parity(float(score[i].item()), grads[i].cpu().numpy(), (s_ref, g_ref), (s64, g64), label=label)
    for st in stages:
        st.close()


def test_batch_rejects_mixed_tile_edges(hip_lib):
    from plastimatch_amd import _abi
    a = stage_for_pair(make_pairs((6, 6, 6), 500)[0])
    b = stage_for_pair(make_pairs((8, 8, 8), 500)[1])
    with pytest.raises(_abi.PlmHipError) as e:
        BsplineBatch([a, b])
    assert e.value.code == _abi.PLM_HIP_ERR_ARG
    a.close()
    b.close()
