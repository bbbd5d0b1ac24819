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


def reference_synth(pattern, dim=38, volume_size=50.0, origin=-25.0, background=None, foreground=None,
                    center=0.0, width=10.0):
    """The reference's synthetic test images ("plastimatch synth" as the
    reference's tests call it) on its geometry: --origin given, spacing =
    volume_size / dim in float (src/cli/pcmd_synth.cxx:338-357), voxel
    positions origin + spacing * index rounded to float (ITK's
    TransformIndexToPhysicalPoint into a float point).  pattern "gauss":
    synth_gauss in double, stored as float (src/util/synthetic_mha.cxx:
    227-245); "sphere": synth_sphere in float, foreground where the
    normalized radius is <= 1, stored as uchar values (:304-328).
    Returns (image zyx, spacing, origin)."""
    sp = np.float32(volume_size) / np.float32(dim)
    o = np.float32(origin)
    phys = (np.float64(o) + np.float64(sp) * np.arange(dim, dtype=np.float64)).astype(np.float32)
    z, y, x = np.meshgrid(phys, phys, phys, indexing="ij")
    if pattern == "gauss":
        bg = -1000.0 if background is None else background
        fg = 0.0 if foreground is None else foreground
        f = np.zeros(z.shape, np.float64)
        for p in (x, y, z):
            f1 = (p.astype(np.float64) - np.float64(np.float32(center))) / np.float64(np.float32(width))
            f += f1 * f1
        f = np.exp(-0.5 * f)
        img = ((1 - f) * np.float64(np.float32(bg)) + f * np.float64(np.float32(fg))).astype(np.float32)
    elif pattern == "sphere":
        bg = 0.0 if background is None else background
        fg = 255.0 if foreground is None else foreground
        f = np.zeros(z.shape, np.float32)
        for p in (x, y, z):
            f1 = (p - np.float32(center)) / np.float32(width)
            f = (f + f1 * f1).astype(np.float32)
        img = np.where(f.astype(np.float64) > 1.0, np.float32(bg), np.float32(fg)).astype(np.float32)
    else:
        raise ValueError(pattern)
    return img, (float(sp),) * 3, (float(o),) * 3
