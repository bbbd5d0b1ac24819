"""Writes the brute-force golden vectors (tests/golden/bspline_tiny.json).

Inputs are tiny seeded volumes; expected outputs come from bruteforce.py
(float64, voxel-order scatter).  Run: python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import bruteforce  # noqa: E402


def central_gradient(img, spacing):
    """Central differences with edge replicate, float as volume_calc_grad."""
    out = np.zeros(img.shape + (3,), np.float32)
    for axis, comp in ((2, 0), (1, 1), (0, 2)):
        n = img.shape[axis]
        idx = np.arange(n)
        nxt = np.minimum(idx + 1, n - 1)
        prv = np.maximum(idx - 1, 0)
        d = (np.take(img, nxt, axis=axis) - np.take(img, prv, axis=axis)).astype(np.float32)
        out[..., comp] = (d.astype(np.float64) / 2.0 / np.float64(np.float32(spacing[comp]))).astype(np.float32)
    return out


def eqsp(img, bins):
    mn, mx = np.float32(img.min()), np.float32(img.max())
    delta = np.float32((mx - mn) / np.float32(bins - 1))
    return float(np.float32(np.float64(mn) - 0.5 * np.float64(delta))), float(delta)


def blobs(shape, rng, n=4):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    img = np.zeros(shape)
    for _ in range(n):
        c = rng.uniform(0, 1, 3) * np.array([nx, ny, nz])
        s = rng.uniform(1.5, 3.0)
        img += rng.uniform(0.5, 2.0) * np.exp(-((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) / (2 * s * s))
    return (100.0 * img).astype(np.float32)


def main():
    rng = np.random.default_rng(2024)
    cases = []
    for name, dim, vpr, spacing, sigma in [
        ("tiny_a", (9, 8, 7), (3, 4, 3), (1.0, 1.0, 2.0), 0.6),
        ("tiny_b", (7, 10, 6), (2, 3, 4), (0.5, 1.0, 1.0), 0.4),
    ]:
        shape = (dim[2], dim[1], dim[0])
        fixed = blobs(shape, rng)
        moving = blobs(shape, rng)
        _, cd = bruteforce.grid(dim, vpr)
        nc = 3 * cd[0] * cd[1] * cd[2]
        coeff = (rng.normal(0.0, sigma, nc) * 8).round() / 8     # dyadic, exact in float
        coeff = coeff.astype(np.float32)
        mg = central_gradient(fixed * 0 + moving, spacing)
        s0, g0, n0 = bruteforce.mse(fixed, moving, spacing, vpr, coeff, 0, mg)
        s1, g1, n1 = bruteforce.mse(fixed, moving, spacing, vpr, coeff, 1)
        bins = 8
        fo, fd = eqsp(fixed, bins)
        mo, md = eqsp(moving, bins)
        s2, g2, n2 = bruteforce.mi(fixed, moving, spacing, vpr, coeff, bins, fo, fd, mo, md)
        cases.append(dict(
            name=name, dim=list(dim), vox_per_rgn=list(vpr), spacing=list(spacing),
            fixed=fixed.ravel().tolist(), moving=moving.ravel().tolist(), coeff=coeff.tolist(),
            m_grad=mg.ravel().tolist(),
            mse_nn=dict(score=s0, grad=g0.tolist(), num_vox=n0),
            mse_exact=dict(score=s1, grad=g1.tolist(), num_vox=n1),
            mi=dict(bins=bins, f_offset=fo, f_delta=fd, m_offset=mo, m_delta=md,
                    score=s2, grad=g2.tolist(), num_vox=n2)))
    with open(os.path.join(HERE, "bspline_tiny.json"), "w") as f:
        json.dump(dict(generator="tests/golden/make_golden.py (bruteforce.py, float64)", cases=cases), f)
    print("wrote", len(cases), "cases")


if __name__ == "__main__":
    main()
