"""Shared builders for the parity tests: one synthetic pair, seen by both
the oracle (CPU restatement) and the HIP flavor (through the C ABI)."""
import numpy as np

import oracle
from plastimatch_amd import synth
from plastimatch_amd.bspline import bf16_round  # noqa: F401  (re-exported for the tests)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def rel_scalar(a, b):
    return abs(float(a) - float(b)) / max(abs(float(b)), 1e-30)


def small_pair(dim=(40, 36, 28), spacing=(1.0, 1.0, 2.5), vpr=(5, 5, 5), seed=1, ct=True,
               coeff_sigma=2.0, n_blobs=10, origin=(0.0, 0.0, 0.0)):
    return synth.make_pair("small", dim, spacing, vpr, coeff_sigma, n_blobs, seed=seed, ct=ct,
                           origin=origin)


def oracle_case(pair, **kw):
    return oracle.Case(pair.fixed, pair.moving, pair.spacing, pair.vox_per_rgn, origin=pair.origin,
                       dc=pair.direction_cosines, **kw)




def assert_parity(score, grad, ref, ref64, tol_s=1e-5, tol_g=1e-4, label=""):
    """GPU result against the float reference flavor (ref = (score, grad)) and
    the FP64 oracle (ref64).  The bars are 1e-5 (score) and 1e-4 (gradient,
    relative L2).  Where the float flavor is itself farther than the bar from
    FP64 -- a nearest-voxel rounding flip of flavor g's round(mijk), or
    cancellation in MI's float PV derivative -- the bar becomes 1.5 times
    that own error: the GPU must then be as close to FP64 as the reference
    flavor is."""
    s_ref, g_ref = ref
    s64, g64 = ref64
    own_s = rel_scalar(s_ref, s64)
    own_g = rel(g_ref, g64)
    bar_s = max(tol_s, 1.5 * own_s)
    bar_g = max(tol_g, 1.5 * own_g)
    es, eg = rel_scalar(score, s64), rel(grad, g64)
    print(f"{label} score err vs f64 {es:.2e} (reference flavor {own_s:.2e}), "
          f"grad err vs f64 {eg:.2e} (reference flavor {own_g:.2e}), vs reference "
          f"{rel(grad, g_ref):.2e}")
    assert es < bar_s, (es, bar_s)
    assert rel_scalar(score, s_ref) < bar_s
    assert eg < bar_g, (eg, bar_g)
    assert rel(grad, g_ref) < bar_g, (rel(grad, g_ref), bar_g)
