"""The optimizer loop on the device (plm_hip_bspline_optimize_lbfgs)
against the oracle's L-BFGS (oracle/lbfgs_oracle.c, itself equal bit for
bit to the reference's liblbfgs build: tests/test_oracle_lbfgs.py).

Both optimizers are driven by the same objective -- the library's own
score (+ regularization) of float-rounded coefficients, as the reference's
evaluate callback rounds them (src/register/bspline_optimize_liblbfgs.cxx:
33-36) -- so the test isolates the optimizer.  The device reduces its dot
products in a fixed tree order, liblbfgs sums them sequentially; the
iterates then agree to rounding, and a float rounding of a coefficient can
flip, so values are compared to 1e-6 and the final x to 1e-5 (relative),
while the discrete record -- status, evaluation count, iteration number,
which evaluations were line-search trials -- must be identical.

A second check runs the oracle optimizer on the reference's float flavor-g
score itself (the drop-in comparison).  There the objectives differ -- the
score at ~1e-7, the gradient up to ~1e-4 where flavor g's nearest-voxel
pick flips -- so the record must still be identical but the values agree
to 1e-3 (the objective values to ~1e-5).
"""
import numpy as np
import pytest

import oracle
from plastimatch_amd.bspline import BsplineHip, Bspline_xform, Joint_histogram, Volume, volume_gradient
from tests.helpers import oracle_case, rel, small_pair

pytestmark = pytest.mark.gpu

REG = dict(curvature_penalty=0.05, diffusion_penalty=0.02, lame_coefficient_1=0.01,
           lame_coefficient_2=0.02)


def stage(roi=None, bins=None, seed=4, dim=(40, 36, 28), vpr=(5, 5, 5)):
    pair = small_pair(dim=dim, vpr=vpr, seed=seed)
    fixed = Volume(pair.fixed, pair.spacing)
    moving = Volume(pair.moving, pair.spacing)
    kw = {"roi_offset": roi[0], "roi_dim": roi[1]} if roi else {}
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn, **kw)
    hist = Joint_histogram(bins, bins).initialize(fixed, moving) if bins else None
    st = BsplineHip(bxf, fixed, moving, volume_gradient(moving), hist=hist)
    return pair, bxf, st


def gpu_objective(st, n, metric, grad_mode, reg):
    """The library's score (+ regularization) of (float) x, synchronously."""
    import torch
    c = torch.zeros(n, dtype=torch.float32, device="cuda")
    s = torch.zeros(1, dtype=torch.float64, device="cuda")
    rs = torch.zeros(1, dtype=torch.float64, device="cuda")
    g = torch.zeros(n, dtype=torch.float64, device="cuda")
    nv = torch.zeros(1, dtype=torch.int64, device="cuda")

    def fun(x):
        fun.last_x = x.copy()
        c.copy_(torch.from_numpy(x.astype(np.float32)))
        torch.cuda.synchronize()
        st.score_device(c, s, g, nv, metric, grad_mode)
        if reg:
            st.regularize(c, rs, g, accumulate=True, **reg)
        torch.cuda.synchronize()
        f = s.item() + (rs.item() if reg else 0.0)
        return f, g.cpu().numpy().copy()
    return fun


def device_run(st, bxf, metric, grad_mode, reg, x0=None, progress=None, resident=False, **parms):
    import torch
    x0 = np.zeros(bxf.num_coeff) if x0 is None else x0
    d_x = torch.from_numpy(np.array(x0, np.float64)).cuda()
    d_c = torch.full((bxf.num_coeff,), 7.0, dtype=torch.float32, device="cuda")
    r = st.optimize_lbfgs(d_x, d_c, metric, grad_mode, reg, progress=progress, resident=resident, **parms)
    return r, d_x.cpu().numpy(), d_c.cpu().numpy()


def compare(dev, ora, tol_f=1e-6):
    assert dev.status == ora.status, (dev.status, ora.status)
    assert dev.fevals == ora.feval
    assert dev.iterations == ora.it
    assert len(dev.reports) == len(ora.reports)
    for a, b in zip(dev.reports, ora.reports):
        assert a[4:] == b[4:]                    # k, ls: which evaluations were trials
        for u, v in zip(a[:4], b[:4]):
            assert abs(u - v) <= tol_f * max(abs(v), 1e-30), (a, b)


@pytest.mark.parametrize("resident", [False, True])
@pytest.mark.parametrize("variant", ["ref_nn", "exact_reg_roi", "mi", "max_feval"])
def test_device_lbfgs_matches_oracle(hip_lib, variant, resident):
    metric, grad_mode, reg, roi, bins, lim = "mse", "ref_nn", None, None, None, dict(max_its=8, max_feval=30)
    if variant == "exact_reg_roi":
        grad_mode, reg, roi = "exact", REG, ((3, 4, 2), (31, 27, 22))
    elif variant == "mi":
        metric, bins, lim = "mi", 32, dict(max_its=5, max_feval=20)
    elif variant == "max_feval":
        lim = dict(max_its=50, max_feval=6)
    pair, bxf, st = stage(roi=roi, bins=bins)
    fun = gpu_objective(st, bxf.num_coeff, metric, grad_mode, reg)
    ora = oracle.lbfgs(fun, np.zeros(bxf.num_coeff), **lim)
    dev, x, c = device_run(st, bxf, metric, grad_mode, reg, resident=resident, **lim)
    print(f"{variant}: status {dev.status} it {dev.iterations} feval {dev.fevals} "
          f"fx {ora.evals[0][0]:.6g} -> {dev.fx:.6g} (oracle {ora.fx:.6g}), x rel {rel(x, ora.x):.2e}")
    compare(dev, ora)
    assert abs(dev.fx - ora.fx) <= 1e-6 * abs(ora.fx)
    assert rel(x, ora.x) < 1e-5
    # the coefficients of the last evaluation (Bspline_xform::coeff at the end)
    assert rel(c, fun.last_x.astype(np.float32)) < 1e-5
    if dev.status >= 0:
        assert np.array_equal(c, x.astype(np.float32))
    if variant == "max_feval":
        assert dev.status == oracle.ORC_LBFGSERR_MAXIMUMLINESEARCH
    st.close()


@pytest.mark.parametrize("resident", [False, True])
@pytest.mark.parametrize("m", [3, 11])
def test_device_lbfgs_history_lengths(hip_lib, m, resident):
    """Histories shorter and longer than the six pairs whose direction solve
    runs on registers: m = 3 overwrites its oldest pair from the fourth
    iteration on, and m = 11 takes the runtime-length solve from the
    seventh iteration and overwrites its oldest pair after the eleventh;
    both against the oracle's liblbfgs restatement."""
    pair, bxf, st = stage(seed=7)
    fun = gpu_objective(st, bxf.num_coeff, "mse", "ref_nn", None)
    lim = dict(max_its=16, max_feval=60)
    ora = oracle.lbfgs(fun, np.zeros(bxf.num_coeff), m=m, **lim)
    dev, x, c = device_run(st, bxf, "mse", "ref_nn", None, resident=resident, m=m, **lim)
    print(f"m {m}: status {dev.status} it {dev.iterations} feval {dev.fevals} fx {dev.fx:.6g} "
          f"(oracle {ora.fx:.6g}), x rel {rel(x, ora.x):.2e}")
    assert dev.iterations > 12
    compare(dev, ora)
    assert abs(dev.fx - ora.fx) <= 1e-6 * abs(ora.fx)
    assert rel(x, ora.x) < 1e-5
    st.close()


def test_device_lbfgs_deterministic_and_progress_stop(hip_lib):
    pair, bxf, st = stage()
    runs = [device_run(st, bxf, "mse", "ref_nn", REG, max_its=5, max_feval=20) for _ in range(2)]
    assert runs[0][0].reports == runs[1][0].reports
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])

    # a caller's progress callback that stops at iteration 3, inside its search
    stop = lambda px, pg, fx, xn, gn, step, n, k, ls: 1 if k >= 3 else 0
    dev, x, c = device_run(st, bxf, "mse", "ref_nn", None, progress=stop, max_its=-1, max_feval=-1)
    fun = gpu_objective(st, bxf.num_coeff, "mse", "ref_nn", None)
    ora = oracle.lbfgs(fun, np.zeros(bxf.num_coeff), max_its=2)
    compare(dev, ora)
    assert dev.status == oracle.ORC_LBFGSERR_MAXIMUMLINESEARCH and dev.iterations == 3
    st.close()


def test_device_lbfgs_against_reference_objective(hip_lib):
    """Drop-in check: the oracle optimizer on the reference's float flavor-g
    score (the objective the reference's run sees) against the device run."""
    pair, bxf, st = stage(dim=(32, 30, 24), seed=6)
    case = oracle_case(pair)

    def fun(x):
        s, g, _, _ = case.mse(x.astype(np.float32), "g")
        return float(s), g.astype(np.float64)

    lim = dict(max_its=6, max_feval=20)
    ora = oracle.lbfgs(fun, np.zeros(case.num_coeff), **lim)
    dev, x, c = device_run(st, bxf, "mse", "ref_nn", None, **lim)
    print(f"status {dev.status}/{ora.status} feval {dev.fevals}/{ora.feval} fx {dev.fx:.8g} vs {ora.fx:.8g}, "
          f"x rel {rel(x, ora.x):.2e}")
    compare(dev, ora, tol_f=1e-3)
    assert abs(dev.fx - ora.fx) <= 1e-4 * abs(ora.fx)
    assert rel(x, ora.x) < 1e-3
    st.close()


@pytest.mark.parametrize("variant", ["mse", "exact_reg", "mi"])
def test_resident_loop_equals_host_loop(hip_lib, variant):
    """The loop run on the device (its scalar logic on one GPU thread, the
    evaluation chain replayed from a hipGraph) and the host-driven loop
    (the same logic on the host between evaluations) make the same run bit
    for bit: every report, the status, the counts, x and the coefficients."""
    metric, grad_mode, reg, bins = "mse", "ref_nn", None, None
    if variant == "exact_reg":
        grad_mode, reg = "exact", REG
    elif variant == "mi":
        metric, bins = "mi", 32
    pair, bxf, st = stage(bins=bins, seed=8)
    runs = []
    for resident in (True, False):
        import torch
        d_x = torch.zeros(bxf.num_coeff, dtype=torch.float64, device="cuda")
        d_c = torch.full((bxf.num_coeff,), 7.0, dtype=torch.float32, device="cuda")
        r = st.optimize_lbfgs(d_x, d_c, metric, grad_mode, reg, resident=resident, max_its=6, max_feval=25)
        runs.append((r, d_x.cpu().numpy(), d_c.cpu().numpy()))
    (a, xa, ca), (b, xb, cb) = runs
    print(f"{variant}: status {a.status} it {a.iterations} feval {a.fevals} reports {len(a.reports)}")
    assert (a.status, a.iterations, a.fevals, a.fx) == (b.status, b.iterations, b.fevals, b.fx)
    assert a.reports == b.reports and len(a.reports) > 3
    assert np.array_equal(xa, xb) and np.array_equal(ca, cb)
    st.close()
