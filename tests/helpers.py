"""Shared builders for the parity tests: one synthetic pair, seen by both
the oracle (CPU restatement) and the HIP flavor (through the C ABI)."""
import numpy as np

import oracle
from plastimatch_amd import synth


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


def bf16_round(a):
    """float32 -> bfloat16 -> float32, round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(np.shape(a))
