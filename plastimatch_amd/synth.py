"""Seeded synthetic fixed/moving pairs for the B-spline score workloads.

The reference's own fixtures are synthesized too (``plastimatch synth``,
src/util/synthetic_mha.cxx:227-245, used by Testing/CMakeLists.txt:319-384);
these generators follow the same idea at the sizes the score is measured on:
a fixed image made of random 3-D Gaussians, and a moving image obtained by
pulling the fixed image back through a seeded cubic B-spline displacement
field (moving(y) = fixed(y - d(y)), trilinear), so that the score evaluated at
the true coefficients is small but not zero.

Arrays are numpy, z-major (shape (nz, ny, nx)), i.e. x fastest in memory as
the reference's volumes are (src/base/volume_macros.h:13-23).
"""
from dataclasses import dataclass, field

import numpy as np

IDENTITY = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


@dataclass
class Pair:
    """One fixed/moving pair plus its B-spline grid and true coefficients."""
    name: str
    fixed: np.ndarray              # (nz, ny, nx) float32
    moving: np.ndarray             # (nz, ny, nx) float32
    spacing: tuple
    origin: tuple
    vox_per_rgn: tuple
    true_coeff: np.ndarray         # (3*num_knots,) float32, interleaved
    direction_cosines: tuple = IDENTITY
    extra: dict = field(default_factory=dict)

    @property
    def dim(self):
        nz, ny, nx = self.fixed.shape
        return (nx, ny, nz)


def grid_dims(dim, vox_per_rgn):
    """rdims = 1 + (dim-1)//vpr, cdims = rdims + 3 (src/base/bspline_header.cxx:28-62)."""
    rdims = tuple(1 + (d - 1) // v for d, v in zip(dim, vox_per_rgn))
    cdims = tuple(r + 3 for r in rdims)
    return rdims, cdims


def num_coeff(dim, vox_per_rgn):
    _, c = grid_dims(dim, vox_per_rgn)
    return 3 * c[0] * c[1] * c[2]


def _basis_matrix(n, vpr, ncp):
    """Dense (n x ncp) matrix of cubic B-spline weights along one axis."""
    i = np.arange(n)
    p = i // vpr
    u = (i % vpr) / vpr
    b = np.stack([(1 - u) ** 3 / 6.0,
                  (3 * u ** 3 - 6 * u ** 2 + 4) / 6.0,
                  (-3 * u ** 3 + 3 * u ** 2 + 3 * u + 1) / 6.0,
                  u ** 3 / 6.0], axis=1)
    w = np.zeros((n, ncp))
    for t in range(4):
        w[i, p + t] = b[:, t]
    return w


def displacement_field(dim, vox_per_rgn, coeff):
    """Dense displacement (nz, ny, nx, 3) in mm from interleaved coefficients (float64)."""
    _, cd = grid_dims(dim, vox_per_rgn)
    c = np.asarray(coeff, dtype=np.float64).reshape(cd[2], cd[1], cd[0], 3)
    wx = _basis_matrix(dim[0], vox_per_rgn[0], cd[0])
    wy = _basis_matrix(dim[1], vox_per_rgn[1], cd[1])
    wz = _basis_matrix(dim[2], vox_per_rgn[2], cd[2])
    t = np.einsum("xa,cbad->cbxd", wx, c)
    t = np.einsum("yb,cbxd->cyxd", wy, t)
    return np.einsum("zc,cyxd->zyxd", wz, t)


def gaussian_blobs(dim, n, rng, sigma=(4.0, 15.0), amplitude=None):
    """Sum of n separable 3-D Gaussians (sigma in voxels), truncated at 4 sigma."""
    nx, ny, nz = dim
    img = np.zeros((nz, ny, nx), dtype=np.float64)
    for _ in range(n):
        s = rng.uniform(sigma[0], sigma[1])
        cx, cy, cz = rng.uniform(0, nx), rng.uniform(0, ny), rng.uniform(0, nz)
        a = amplitude(rng) if amplitude else rng.uniform(0.0, 1.0)
        r = 4.0 * s
        sl = []
        g1 = []
        for c, n_ in ((cz, nz), (cy, ny), (cx, nx)):
            lo, hi = max(0, int(c - r)), min(n_, int(c + r) + 1)
            if hi <= lo:
                lo, hi = 0, 0
            x = np.arange(lo, hi)
            sl.append(slice(lo, hi))
            g1.append(np.exp(-0.5 * ((x - c) / s) ** 2))
        if any(len(g) == 0 for g in g1):
            continue
        img[sl[0], sl[1], sl[2]] += a * g1[0][:, None, None] * g1[1][None, :, None] * g1[2][None, None, :]
    return img


def ct_amplitude(rng):
    """Blob intensity from a CT-like mixture: air / soft tissue / bone (HU)."""
    k = rng.choice(3, p=[0.3, 0.6, 0.1])
    return (rng.normal(-1000.0, 50.0), rng.normal(40.0, 20.0), rng.normal(700.0, 100.0))[k]


def warp_pull_back(fixed, spacing, origin, disp):
    """moving(y) = fixed(y - d(y)), trilinear in voxel coordinates, edge clamped."""
    from scipy.ndimage import map_coordinates
    nz, ny, nx = fixed.shape
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    coords = np.empty((3,) + fixed.shape, dtype=np.float64)
    coords[0] = zz - disp[..., 2] / spacing[2]
    coords[1] = yy - disp[..., 1] / spacing[1]
    coords[2] = xx - disp[..., 0] / spacing[0]
    out = map_coordinates(fixed.astype(np.float64), coords, order=1, mode="nearest")
    return out.astype(np.float32)


def make_pair(name, dim, spacing, vox_per_rgn, coeff_sigma, n_blobs, seed,
              ct=False, noise=0.0, remap=None, origin=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    fixed = gaussian_blobs(dim, n_blobs, rng, amplitude=ct_amplitude if ct else None)
    if noise:
        fixed = fixed + rng.normal(0.0, noise, size=fixed.shape)
    fixed = fixed.astype(np.float32)
    nc = num_coeff(dim, vox_per_rgn)
    true_coeff = rng.normal(0.0, coeff_sigma, size=nc).astype(np.float32)
    disp = displacement_field(dim, vox_per_rgn, true_coeff)
    moving = warp_pull_back(fixed, spacing, origin, disp)
    if remap is not None:
        moving = remap(moving.astype(np.float64)).astype(np.float32)
    return Pair(name, fixed, moving, tuple(spacing), tuple(origin), tuple(vox_per_rgn), true_coeff)


def ct_remap(x):
    """Monotone intensity remap y = 0.8x + 200 + 50 tanh(x/300)."""
    return 0.8 * x + 200.0 + 50.0 * np.tanh(x / 300.0)


# The five measured workloads.  Sizes are (nx, ny, nz); spacings in mm.
CONFIGS = {
    "mse_128cube_vpr15": dict(dim=(128, 128, 128), spacing=(1.0, 1.0, 1.0), vox_per_rgn=(15, 15, 15),
                              coeff_sigma=2.0, n_blobs=40, noise=0.01),
    "mse_256x256x128_vpr10": dict(dim=(256, 256, 128), spacing=(1.0, 1.0, 2.5), vox_per_rgn=(10, 10, 10),
                                  coeff_sigma=3.0, n_blobs=120, ct=True),
    "mi_512x512x160_vpr12": dict(dim=(512, 512, 160), spacing=(0.9765625, 0.9765625, 2.5),
                                 vox_per_rgn=(12, 12, 12), coeff_sigma=4.0, n_blobs=120, ct=True,
                                 remap=ct_remap),
    "adjoint_192x192x96_vpr8": dict(dim=(192, 192, 96), spacing=(1.0, 1.0, 1.0), vox_per_rgn=(8, 8, 8),
                                    coeff_sigma=2.0, n_blobs=40, noise=0.01),
}


def make_config(name, seed=0, scale=1):
    """Build one named workload; scale>1 shrinks every edge by that factor (for tests)."""
    cfg = dict(CONFIGS[name])
    if scale != 1:
        cfg["dim"] = tuple(max(8, d // scale) for d in cfg["dim"])
        cfg["n_blobs"] = max(4, cfg["n_blobs"] // scale)
    return make_pair(name, seed=seed, **cfg)


def make_batch(n_pairs=16, seed=0, lo=96, hi=256, vox_per_rgn=(10, 10, 10), coeff_sigma=3.0):
    """Multi-atlas batch: pairs of uneven sizes, edges uniform in lo..hi per axis."""
    rng = np.random.default_rng(seed)
    pairs = []
    for p in range(n_pairs):
        dim = tuple(int(v) for v in rng.integers(lo, hi + 1, size=3))
        pairs.append(make_pair(f"pair{p}", dim, (1.0, 1.0, 1.0), vox_per_rgn, coeff_sigma,
                               n_blobs=40, seed=int(rng.integers(1 << 30)), noise=0.01))
    return pairs


def random_coeff(dim, vox_per_rgn, sigma, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(0.0, sigma, size=num_coeff(dim, vox_per_rgn)).astype(np.float32)


def write_mha(path, img, spacing, origin=(0.0, 0.0, 0.0), direction_cosines=IDENTITY):
    """Uncompressed float MetaImage of a synthetic pair, with the geometry
    written exactly (plastimatch_amd.mha.write_mha, precise)."""
    from .mha import write_mha as _write
    _write(path, np.ascontiguousarray(img, np.float32), spacing, origin, direction_cosines, precise=True)


def eqsp_hist(img, bins):
    """Equal-spacing bin parameters in float32: delta=(max-min)/(bins-1),
    offset=min-0.5*delta (src/register/joint_histogram.cxx:115-139)."""
    mn = np.float32(img.min())
    mx = np.float32(img.max())
    delta = np.float32((mx - mn) / np.float32(bins - 1))
    offset = np.float32(np.float64(mn) - 0.5 * np.float64(delta))
    return float(offset), float(delta)
