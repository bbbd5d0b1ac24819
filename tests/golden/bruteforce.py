"""Independent FP64 brute force of the B-spline MSE and PV-8 Mattes-MI
score and gradient, for tiny inputs.

Written from the mathematics, not from the C oracle: the displacement is the
tensor-product cubic B-spline sum over the 4x4x4 knots around each voxel
(src/base/bspline_interpolate.cxx:78-110), the moving image is sampled
trilinearly with the reference's clamping (src/base/interpolate.cxx:17-41),
and every gradient contribution is scattered voxel by voxel straight onto its
64 knots (the voxel-order form of src/register/bspline_score.cxx:145-170), so
none of the tile condense/reduce machinery is shared with the oracle.
Everything is float64; dyadic inputs keep positions exact.
"""
import math

import numpy as np


def basis(u):
    return np.array([(1 - u) ** 3 / 6.0, (3 * u ** 3 - 6 * u ** 2 + 4) / 6.0,
                     (-3 * u ** 3 + 3 * u ** 2 + 3 * u + 1) / 6.0, u ** 3 / 6.0])


def grid(dim, vpr):
    rd = [1 + (d - 1) // v for d, v in zip(dim, vpr)]
    return rd, [r + 3 for r in rd]


def _voxel_weights(ijk, vpr, cd):
    """[(knot_linear, weight)] for the 64 knots of voxel ijk."""
    p = [ijk[a] // vpr[a] for a in range(3)]
    u = [(ijk[a] % vpr[a]) / vpr[a] for a in range(3)]
    bx, by, bz = basis(u[0]), basis(u[1]), basis(u[2])
    out = []
    for k in range(4):
        for j in range(4):
            for i in range(4):
                kn = (p[0] + i) + cd[0] * ((p[1] + j) + cd[1] * (p[2] + k))
                out.append((kn, bx[i] * by[j] * bz[k]))
    return out


def _sample(img, mijk):
    """Clamped trilinear state of a point, or None if outside (-0.5, dim-0.5)."""
    nz, ny, nx = img.shape
    dims = (nx, ny, nz)
    f, r, l2 = [], [], []
    for a in range(3):
        m = mijk[a]
        if m <= -0.5 or m >= dims[a] - 0.5:
            return None
        dmax = dims[a] - 1
        if m < 0:
            f.append(0), r.append(0), l2.append(0.0)
        elif m >= dmax:
            f.append(dmax - 1), r.append(dmax), l2.append(1.0)
        else:
            fl = math.floor(m)
            f.append(fl), r.append(int(m + 0.5)), l2.append(m - fl)
    return f, r, l2


def _corners(img, f):
    x, y, z = f
    return np.array([img[z, y, x], img[z, y, x + 1], img[z, y + 1, x], img[z, y + 1, x + 1],
                     img[z + 1, y, x], img[z + 1, y, x + 1], img[z + 1, y + 1, x],
                     img[z + 1, y + 1, x + 1]], dtype=np.float64)


def _weights(l2):
    l1 = [1 - v for v in l2]
    return np.array([l1[0] * l1[1] * l1[2], l2[0] * l1[1] * l1[2], l1[0] * l2[1] * l1[2],
                     l2[0] * l2[1] * l1[2], l1[0] * l1[1] * l2[2], l2[0] * l1[1] * l2[2],
                     l1[0] * l2[1] * l2[2], l2[0] * l2[1] * l2[2]]), l1


def _dweights(l1, l2):
    """d w_n / d(i, j, k) for the 8 PV weights, shape (8, 3)."""
    sx = [-1, 1, -1, 1, -1, 1, -1, 1]
    sy = [-1, -1, 1, 1, -1, -1, 1, 1]
    sz = [-1, -1, -1, -1, 1, 1, 1, 1]
    fx = lambda n: l2[0] if sx[n] > 0 else l1[0]
    fy = lambda n: l2[1] if sy[n] > 0 else l1[1]
    fz = lambda n: l2[2] if sz[n] > 0 else l1[2]
    return np.array([[sx[n] * fy(n) * fz(n), fx(n) * sy[n] * fz(n), fx(n) * fy(n) * sz[n]]
                     for n in range(8)])


def _roi_at(mask, pos):
    """inside_roi of a mask on an index grid: the floor of each coordinate,
    below 0 taken as 0 and at or past the edge as the last voxel."""
    nz, ny, nx = mask.shape
    p = []
    for a, n in enumerate((nx, ny, nz)):
        v = pos[a]
        p.append(0 if v < 0 else (n - 1 if v >= n else int(math.floor(v))))
    return bool(mask[p[2], p[1], p[0]])


def mse(fixed, moving, spacing, vpr, coeff, grad_mode, m_grad=None, fmask=None, mmask=None):
    """FP64 MSE score, gradient (2/N normalized) and N; identity direction
    cosines, zero origins.  grad_mode 0: diff * m_grad[round], 1: analytic.
    fmask / mmask: optional regions of interest on the fixed and moving
    grids (a voxel counts only where the fixed mask holds at its position
    and the moving mask at its displaced position)."""
    nz, ny, nx = fixed.shape
    dim = (nx, ny, nz)
    _, cd = grid(dim, vpr)
    c = np.asarray(coeff, np.float64).reshape(-1, 3)
    grad = np.zeros_like(c)
    raw, n = 0.0, 0
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if fmask is not None and not fmask[k, j, i]:
                    continue
                ws = _voxel_weights((i, j, k), vpr, cd)
                d = sum(w * c[kn] for kn, w in ws)
                mijk = [(v * spacing[a] + d[a]) / spacing[a] for a, v in enumerate((i, j, k))]
                smp = _sample(moving, mijk)
                if smp is None:
                    continue
                if mmask is not None and not _roi_at(mmask, mijk):
                    continue
                f, r, l2 = smp
                a8 = _corners(moving, f)
                w8, l1 = _weights(l2)
                diff = float(w8 @ a8) - float(fixed[k, j, i])
                raw += diff * diff
                n += 1
                if grad_mode == 0:
                    g = np.array(m_grad[r[2], r[1], r[0]], np.float64)
                else:
                    g = (_dweights(l1, l2).T @ a8) / np.array(spacing, np.float64)
                for kn, w in ws:
                    grad[kn] += diff * g * w
    if n < 1:
        return float(np.finfo(np.float32).max), np.zeros(c.size), 0
    return raw / n, (2.0 * grad / n).reshape(-1), n


def mi(fixed, moving, spacing, vpr, coeff, bins, f_off, f_delta, m_off, m_delta, fmask=None,
       mmask=None, f_keys=None, m_keys=None):
    """FP64 PV-8 Mattes MI: score -MI, gradient (thresholds 0.001/B^2 for the
    score, 1e-4 for the gradient) and N; identity direction cosines.
    fmask / mmask as for mse.  f_keys / m_keys: V-optimal key tables (the
    bin of key floor((v - offset) / delta)), or None for equal spacing."""
    def to_bin(v, off, dl, keys):
        b = int(math.floor((float(v) - off) / dl))
        return b if keys is None else int(keys[min(max(b, 0), len(keys) - 1)])

    nz, ny, nx = fixed.shape
    dim = (nx, ny, nz)
    _, cd = grid(dim, vpr)
    c = np.asarray(coeff, np.float64).reshape(-1, 3)
    fh = np.zeros(bins)
    mh = np.zeros(bins)
    jh = np.zeros((bins, bins))
    vox = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if fmask is not None and not fmask[k, j, i]:
                    continue
                ws = _voxel_weights((i, j, k), vpr, cd)
                d = sum(w * c[kn] for kn, w in ws)
                mijk = [(v * spacing[a] + d[a]) / spacing[a] for a, v in enumerate((i, j, k))]
                smp = _sample(moving, mijk)
                if smp is None:
                    continue
                if mmask is not None and not _roi_at(mmask, mijk):
                    continue
                f, r, l2 = smp
                a8 = _corners(moving, f)
                w8, l1 = _weights(l2)
                fb = to_bin(fixed[k, j, i], f_off, f_delta, f_keys)
                mbs = [to_bin(v, m_off, m_delta, m_keys) for v in a8]
                fh[fb] += 1
                for nn in range(8):
                    mh[mbs[nn]] += w8[nn]
                    jh[fb, mbs[nn]] += w8[nn]
                vox.append((ws, fb, mbs, l1, l2))
    n = len(vox)
    thresh = 0.001 / (bins * bins)
    s = 0.0
    for a in range(bins):
        for b in range(bins):
            if jh[a, b] > thresh:
                s -= jh[a, b] * math.log(n * jh[a, b] / (mh[b] * fh[a]))
    s /= n
    grad = np.zeros_like(c)
    for ws, fb, mbs, l1, l2 in vox:
        dw = _dweights(l1, l2)
        dc = np.zeros(3)
        for nn in range(8):
            jv = jh[fb, mbs[nn]]
            if jv > 0.0001:
                dc -= dw[nn] * (math.log(n * jv / (fh[fb] * mh[mbs[nn]])) - s)
        dc = dc / n / np.array(spacing, np.float64)
        for kn, w in ws:
            grad[kn] += dc * w
    return s, grad.reshape(-1), n
