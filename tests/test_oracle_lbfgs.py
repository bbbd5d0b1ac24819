"""The oracle's L-BFGS (oracle/lbfgs_oracle.c) against the reference's own
liblbfgs 1.9, compiled from libs/liblbfgs-1.9/lib/lbfgs.c into
oracle/_ref by oracle/Makefile, and against what the mathematics fixes.

The restatement keeps liblbfgs's arithmetic order, so the two must give
the same trajectory bit for bit: every evaluation (value, step), every
progress report, the status and the final x.  Independently of the
reference build, the optimizer must find the minimizers the problems have
in closed form: the reference's own sample problem (libs/liblbfgs-1.9/
sample/sample.c: the extended Rosenbrock function, minimum at x = 1) and
a strictly convex quadratic (minimum A^-1 b).  Where the reference build is
absent the bitwise checks are skipped, the closed-form ones still run.
"""
import numpy as np
import pytest

import oracle
from tests.helpers import oracle_case, small_pair

have_ref = oracle.ref_lbfgs_lib() is not None
need_ref = pytest.mark.skipif(not have_ref, reason="reference liblbfgs not built (oracle/_ref)")


def rosenbrock(x):
    """libs/liblbfgs-1.9/sample/sample.c:4-23, same arithmetic order."""
    g = np.zeros_like(x)
    f = 0.0
    for i in range(0, len(x), 2):
        t1 = 1.0 - x[i]
        t2 = 10.0 * (x[i + 1] - x[i] * x[i])
        g[i + 1] = 20.0 * t2
        g[i] = -2.0 * (x[i] * g[i + 1] + t1)
        f += t1 * t1 + t2 * t2
    return f, g


def quadratic(n=40, seed=0):
    rng = np.random.default_rng(seed)
    q = rng.normal(size=(n, n))
    A = q @ q.T / n + np.diag(np.linspace(0.5, 4.0, n))
    b = rng.normal(size=n)

    def fun(x):
        Ax = A @ x
        return float(0.5 * x @ Ax - b @ x), Ax - b
    return fun, np.linalg.solve(A, b)


def same_run(a, b):
    assert a.status == b.status, (a.status, b.status)
    assert a.evals == b.evals
    assert a.reports == b.reports
    assert np.array_equal(a.x, b.x)
    assert a.fx == b.fx


def test_rosenbrock_sample_converges():
    x0 = np.tile([-1.2, 1.0], 50)
    r = oracle.lbfgs(rosenbrock, x0)
    assert r.status == oracle.ORC_LBFGS_SUCCESS
    assert np.abs(r.x - 1.0).max() < 1e-4
    assert r.fx < 1e-10
    # every evaluation after the first is a line-search trial with a positive step
    assert r.evals[0][1] == 0.0 and all(s > 0 for _, s in r.evals[1:])


def test_quadratic_minimizer():
    fun, xstar = quadratic()
    r = oracle.lbfgs(fun, np.zeros(40), epsilon=1e-8)
    assert r.status == oracle.ORC_LBFGS_SUCCESS
    assert np.abs(r.x - xstar).max() < 2e-8 * np.abs(xstar).max()


def test_already_minimized():
    fun, xstar = quadratic(n=10, seed=1)
    r = oracle.lbfgs(lambda x: (0.0, np.zeros_like(x)), xstar)
    assert r.status == oracle.ORC_LBFGS_ALREADY_MINIMIZED and len(r.evals) == 1


@need_ref
@pytest.mark.parametrize("case", ["rosenbrock", "quadratic", "m3", "past", "max_iterations",
                                  "max_its", "max_feval", "bad_epsilon", "minimized"])
def test_matches_reference_liblbfgs(case):
    x0 = np.tile([-1.2, 1.0], 50)
    fun, kw, ext = rosenbrock, {}, {}
    if case == "quadratic":
        fun, _ = quadratic()
        x0 = np.zeros(40)
    elif case == "m3":
        kw = dict(m=3)
    elif case == "past":
        kw = dict(past=3, delta=1e-2)
    elif case == "max_iterations":
        kw = dict(max_iterations=7)
    elif case == "max_its":
        ext = dict(max_its=4)                    # the reference's progress rule
    elif case == "max_feval":
        ext = dict(max_feval=9)                  # stops inside a line search: x reverts
    elif case == "bad_epsilon":
        kw = dict(epsilon=-1.0)
    elif case == "minimized":
        x0 = np.ones(100)
    a = oracle.lbfgs(fun, x0, **ext, **kw)
    b = oracle.lbfgs(fun, x0, impl="ref", **ext, **kw)
    print(f"{case}: status {a.status}, {len(a.evals)} evaluations, fx {a.fx:.6e}")
    same_run(a, b)
    if case == "max_feval":
        assert a.status == oracle.ORC_LBFGSERR_MAXIMUMLINESEARCH
    if case == "max_its":
        # the line search reports each trial with the iteration number first,
        # so the limit ends a run inside the search of iteration max_its + 1
        assert a.status == oracle.ORC_LBFGSERR_MAXIMUMLINESEARCH and a.it == 5
    if case == "bad_epsilon":
        assert a.status == oracle.ORC_LBFGSERR_INVALID_EPSILON and not a.evals


@need_ref
def test_registration_objective_matches_reference_liblbfgs():
    """The reference's own evaluate (bspline_optimize_liblbfgs.cxx:20-51):
    coefficients rounded to float, the flavor-g MSE score and float gradient
    widened to double -- here the oracle's float restatement of that score."""
    pair = small_pair(dim=(24, 22, 20), vpr=(5, 5, 5), seed=4)
    case = oracle_case(pair)

    def fun(x):
        s, g, _, _ = case.mse(x.astype(np.float32), "g")
        return float(s), g.astype(np.float64)

    x0 = np.zeros(case.num_coeff)
    a = oracle.lbfgs(fun, x0, max_its=6, max_feval=12)
    b = oracle.lbfgs(fun, x0, impl="ref", max_its=6, max_feval=12)
    print(f"status {a.status}, it {a.it}, feval {a.feval}, fx {a.evals[0][0]:.4f} -> {a.fx:.4f}")
    same_run(a, b)
    assert a.fx < a.evals[0][0]
