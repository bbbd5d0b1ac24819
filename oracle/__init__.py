"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle restates the reference's CPU flavors (plm_oracle.c) and is used
by tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg as
the checker.  The product package never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

c_float3 = ctypes.c_float * 3
c_int64_3 = ctypes.c_int64 * 3
c_float9 = ctypes.c_float * 9
IDENTITY = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


class OrcVolHdr(ctypes.Structure):
    _fields_ = [("dim", c_int64_3), ("origin", c_float3), ("spacing", c_float3), ("dc", c_float9)]


class OrcRois(ctypes.Structure):
    """Optional fixed / moving regions of interest (uchar volumes)."""
    _fields_ = [("fixed_hdr", ctypes.POINTER(OrcVolHdr)), ("fixed", ctypes.c_void_p),
                ("moving_hdr", ctypes.POINTER(OrcVolHdr)), ("moving", ctypes.c_void_p)]


class OrcGeom(ctypes.Structure):
    _fields_ = [("img_origin", c_float3), ("img_spacing", c_float3), ("img_dim", c_int64_3),
                ("roi_offset", c_int64_3), ("roi_dim", c_int64_3), ("vox_per_rgn", c_int64_3),
                ("dc", c_float9)]


class OrcRegParms(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in (
        "total_displacement_penalty", "diffusion_penalty", "curvature_penalty", "lame_coefficient_1",
        "lame_coefficient_2", "linear_elastic_multiplier", "third_order_penalty",
        "curvature_mixed_weight")]


REG_DEFAULTS = {"total_displacement_penalty": 0.0, "diffusion_penalty": 0.0, "curvature_penalty": 0.0,
                "lame_coefficient_1": 0.0, "lame_coefficient_2": 0.0, "linear_elastic_multiplier": 1.0,
                "third_order_penalty": 0.0, "curvature_mixed_weight": 1.0}


def reg_parms(**kw):
    """Regularization weights with the reference's defaults
    (src/register/regularization_parms.h:31-42)."""
    d = dict(REG_DEFAULTS)
    d.update(kw)
    return OrcRegParms(**d)


VOPT_KEYS = 1000


class OrcHist(ctypes.Structure):
    """orc_hist; f_keys / m_keys point at a side's V-optimal key table (its
    numpy array is kept on the instance) or are NULL (equal spacing)."""
    _fields_ = [("fixed_bins", ctypes.c_int64), ("moving_bins", ctypes.c_int64),
                ("f_offset", ctypes.c_float), ("f_delta", ctypes.c_float),
                ("m_offset", ctypes.c_float), ("m_delta", ctypes.c_float),
                ("f_keys", ctypes.c_void_p), ("m_keys", ctypes.c_void_p)]

    def set_keys(self, f_lut=None, m_lut=None):
        self._luts = (None if f_lut is None else np.ascontiguousarray(f_lut, np.int32),
                      None if m_lut is None else np.ascontiguousarray(m_lut, np.int32))
        self.f_keys = None if f_lut is None else self._luts[0].ctypes.data
        self.m_keys = None if m_lut is None else self._luts[1].ctypes.data
        return self


class OrcLbfgsParms(ctypes.Structure):
    """The lbfgs_parameter_t fields the More-Thuente path reads."""
    _fields_ = [("m", ctypes.c_int), ("epsilon", ctypes.c_double), ("past", ctypes.c_int),
                ("delta", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("max_linesearch", ctypes.c_int), ("min_step", ctypes.c_double),
                ("max_step", ctypes.c_double), ("ftol", ctypes.c_double), ("gtol", ctypes.c_double),
                ("xtol", ctypes.c_double)]


class RefLbfgsParameter(ctypes.Structure):
    """lbfgs_parameter_t of the reference's liblbfgs 1.9
    (libs/liblbfgs-1.9/include/lbfgs.h:200-360), for the build in _ref/."""
    _fields_ = [("m", ctypes.c_int), ("epsilon", ctypes.c_double), ("past", ctypes.c_int),
                ("delta", ctypes.c_double), ("max_iterations", ctypes.c_int),
                ("linesearch", ctypes.c_int), ("max_linesearch", ctypes.c_int),
                ("min_step", ctypes.c_double), ("max_step", ctypes.c_double), ("ftol", ctypes.c_double),
                ("wolfe", ctypes.c_double), ("gtol", ctypes.c_double), ("xtol", ctypes.c_double),
                ("orthantwise_c", ctypes.c_double), ("orthantwise_start", ctypes.c_int),
                ("orthantwise_end", ctypes.c_int)]


_dp = ctypes.POINTER(ctypes.c_double)
LBFGS_EVAL = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_double)
LBFGS_PROGRESS = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int)

REF_DIR = os.path.join(HERE, "_ref")
REF_LBFGS_PATH = os.path.join(REF_DIR, "liblbfgs_ref.so")

_lib = None
_ref_lbfgs = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.orc_header.restype = ctypes.c_int64
        L.orc_header.argtypes = [ctypes.POINTER(OrcGeom), P, P]
        L.orc_basis_lut.argtypes = [ctypes.c_int64, P]
        L.orc_direction_matrices.argtypes = [P, P, P, P, P]
        L.orc_interp_vf.argtypes = [ctypes.POINTER(OrcGeom), P, P]
        L.orc_warp.restype = ctypes.c_int
        L.orc_warp.argtypes = [ctypes.POINTER(OrcGeom), ctypes.POINTER(OrcVolHdr), P, P, ctypes.c_int,
                               ctypes.c_float, ctypes.c_int, P]
        L.orc_volume_gradient.argtypes = [ctypes.POINTER(OrcVolHdr), P, P]
        L.orc_hist_eqsp.argtypes = [P, ctypes.c_int64, ctypes.c_int64, P, P]
        L.orc_hist_vopt.restype = ctypes.c_int
        L.orc_hist_vopt.argtypes = [P, ctypes.c_int64, ctypes.c_int64, P, P, P, P, P]
        L.orc_vopt_partition.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int64, P]
        L.orc_pass1_position_mismatch.restype = ctypes.c_int64
        L.orc_pass1_position_mismatch.argtypes = [ctypes.POINTER(OrcVolHdr), ctypes.POINTER(OrcGeom)]
        L.orc_mse_score.restype = ctypes.c_int
        L.orc_mse_score.argtypes = [ctypes.POINTER(OrcGeom), ctypes.POINTER(OrcVolHdr), P,
                                    ctypes.POINTER(OrcVolHdr), P, P, P, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(OrcRois), P, P, P, P]
        L.orc_mse_score_f64.restype = ctypes.c_int
        L.orc_mse_score_f64.argtypes = [ctypes.POINTER(OrcGeom), ctypes.POINTER(OrcVolHdr), P,
                                        ctypes.POINTER(OrcVolHdr), P, P, P, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(OrcRois), P, P, P]
        L.orc_pv_quantize.argtypes = [P, P]
        L.orc_reg_analytic.restype = ctypes.c_int
        L.orc_reg_analytic.argtypes = [ctypes.POINTER(OrcGeom), ctypes.POINTER(OrcRegParms), P, ctypes.c_int,
                                       P, P]
        L.orc_reg_matrices.restype = ctypes.c_int
        L.orc_reg_matrices.argtypes = [ctypes.POINTER(OrcGeom), ctypes.POINTER(OrcRegParms), P]
        L.orc_mi_score.restype = ctypes.c_int
        L.orc_mi_score.argtypes = [ctypes.POINTER(OrcGeom), ctypes.POINTER(OrcVolHdr), P,
                                   ctypes.POINTER(OrcVolHdr), P, P, ctypes.POINTER(OrcHist),
                                   ctypes.c_int, ctypes.POINTER(OrcRois), P, P, P, P, P, P, P, P]
        L.orc_mi_score_f64.restype = ctypes.c_int
        L.orc_mi_score_f64.argtypes = [ctypes.POINTER(OrcGeom), ctypes.POINTER(OrcVolHdr), P,
                                       ctypes.POINTER(OrcVolHdr), P, P, ctypes.POINTER(OrcHist),
                                       ctypes.c_int, ctypes.POINTER(OrcRois), P, P, P]
        L.orc_mse_f64_support.restype = ctypes.c_int
        L.orc_mse_f64_support.argtypes = [ctypes.POINTER(OrcGeom), ctypes.POINTER(OrcVolHdr), P,
                                          ctypes.POINTER(OrcVolHdr), P, P, ctypes.c_int64,
                                          ctypes.c_double, ctypes.POINTER(OrcRois), P, P, P]
        L.orc_mi_score_f64_hist.restype = ctypes.c_int
        L.orc_mi_score_f64_hist.argtypes = [ctypes.POINTER(OrcGeom), ctypes.POINTER(OrcVolHdr), P,
                                            ctypes.POINTER(OrcVolHdr), P, P, ctypes.POINTER(OrcHist),
                                            ctypes.c_int, ctypes.POINTER(OrcRois), P, ctypes.c_int64, P,
                                            P, P]
        L.orc_resample_geometry.restype = ctypes.c_int
        L.orc_resample_geometry.argtypes = [ctypes.c_int, ctypes.POINTER(OrcVolHdr), P, P, P, P]
        L.orc_volume_resample.argtypes = [ctypes.POINTER(OrcVolHdr), P, P, P, P, ctypes.c_int, P]
        L.orc_lbfgs_parms_init.argtypes = [ctypes.POINTER(OrcLbfgsParms)]
        L.orc_lbfgs.restype = ctypes.c_int
        L.orc_lbfgs.argtypes = [ctypes.c_int, P, P, LBFGS_EVAL, LBFGS_PROGRESS, P,
                                ctypes.POINTER(OrcLbfgsParms)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def vol_hdr(img, spacing, origin=(0.0, 0.0, 0.0), dc=IDENTITY):
    nz, ny, nx = img.shape[:3]
    return OrcVolHdr(c_int64_3(nx, ny, nz), c_float3(*origin), c_float3(*spacing), c_float9(*dc))


def geom(dim, spacing, vox_per_rgn, origin=(0.0, 0.0, 0.0), roi_offset=(0, 0, 0), roi_dim=None,
         dc=IDENTITY):
    roi_dim = roi_dim or dim
    return OrcGeom(c_float3(*origin), c_float3(*spacing), c_int64_3(*dim), c_int64_3(*roi_offset),
                   c_int64_3(*roi_dim), c_int64_3(*vox_per_rgn), c_float9(*dc))


def header(g):
    r = np.zeros(3, np.int64)
    c = np.zeros(3, np.int64)
    nk = lib().orc_header(ctypes.byref(g), _ptr(r), _ptr(c))
    return r, c, nk


def basis_lut(vpr):
    out = np.zeros(4 * vpr, np.float32)
    lib().orc_basis_lut(vpr, _ptr(out))
    return out.reshape(vpr, 4)


def direction_matrices(dc, spacing):
    dc = np.asarray(dc, np.float32)
    sp = np.asarray(spacing, np.float32)
    step = np.zeros(9, np.float32)
    proj = np.zeros(9, np.float32)
    inv = np.zeros(9, np.float32)
    lib().orc_direction_matrices(_ptr(dc), _ptr(sp), _ptr(step), _ptr(proj), _ptr(inv))
    return step, proj, inv


def interp_vf(g, coeff):
    nx, ny, nz = g.img_dim
    vf = np.zeros((nz, ny, nx, 3), np.float32)
    lib().orc_interp_vf(ctypes.byref(g), _ptr(np.ascontiguousarray(coeff, np.float32)), _ptr(vf))
    return vf


def warp(g, moving, m_spacing, coeff, linear=True, default_val=0.0, m_origin=(0.0, 0.0, 0.0),
         m_dc=IDENTITY, nthreads=0):
    """Warped moving image on the fixed grid of g (orc_warp)."""
    moving = np.ascontiguousarray(moving, np.float32)
    nx, ny, nz = g.img_dim
    out = np.zeros((nz, ny, nx), np.float32)
    mh = vol_hdr(moving, m_spacing, m_origin, m_dc)
    rc = lib().orc_warp(ctypes.byref(g), ctypes.byref(mh), _ptr(moving),
                        _ptr(np.ascontiguousarray(coeff, np.float32)), 1 if linear else 0,
                        float(default_val), int(nthreads), _ptr(out))
    if rc:
        raise ValueError(f"orc_warp failed ({rc})")
    return out


RESAMPLE_KINDS = {"voxel_rate": 0, "voxel_legacy": 1, "mm": 2, "pct": 3}


def resample_geometry(kind, shape_zyx, spacing, origin, rate):
    """registration_resample_volume's output geometry: (identity, dim xyz,
    origin, spacing); identity means the reference clones the volume."""
    nz, ny, nx = shape_zyx
    h = OrcVolHdr(c_int64_3(nx, ny, nz), c_float3(*origin), c_float3(*spacing), c_float9(*IDENTITY))
    r = np.ascontiguousarray(rate, np.float32)
    dim = np.zeros(3, np.int64)
    o = np.zeros(3, np.float32)
    s = np.zeros(3, np.float32)
    rc = lib().orc_resample_geometry(RESAMPLE_KINDS[kind], ctypes.byref(h), _ptr(r), _ptr(dim), _ptr(o), _ptr(s))
    if rc < 0:
        raise ValueError("resample kind or rate the reference cannot use")
    return bool(rc), tuple(int(v) for v in dim), tuple(float(v) for v in o), tuple(float(v) for v in s)


def volume_resample(img, spacing, origin, dim, out_origin, out_spacing, nearest=False):
    """volume_resample_float_li / _nn of img (z, y, x) onto dim (x, y, z)."""
    img = np.ascontiguousarray(img, np.float32)
    nz, ny, nx = img.shape
    h = OrcVolHdr(c_int64_3(nx, ny, nz), c_float3(*origin), c_float3(*spacing), c_float9(*IDENTITY))
    d = np.ascontiguousarray(dim, np.int64)
    out = np.zeros((d[2], d[1], d[0]), np.float32)
    lib().orc_volume_resample(ctypes.byref(h), _ptr(img), _ptr(d), _ptr(np.asarray(out_origin, np.float32)),
                              _ptr(np.asarray(out_spacing, np.float32)), 1 if nearest else 0, _ptr(out))
    return out


def stage_resample(img, spacing, origin, kind, rate, nearest=False):
    """registration_resample_volume: (resampled image, origin, spacing)."""
    ident, dim, o, s = resample_geometry(kind, img.shape, spacing, origin, rate)
    if ident:
        return np.array(img, np.float32, copy=True), o, s
    return volume_resample(img, spacing, origin, dim, o, s, nearest), o, s


def reg_analytic(g, rp, coeff, nthreads=0):
    """Analytic regularization, flavor c: (score, gradient) in double."""
    _, _, nk = header(g)
    grad = np.zeros(3 * nk, np.float64)
    sc = np.zeros(1, np.float64)
    rc = lib().orc_reg_analytic(ctypes.byref(g), ctypes.byref(rp),
                                _ptr(np.ascontiguousarray(coeff, np.float32)), int(nthreads), _ptr(sc),
                                _ptr(grad))
    if rc:
        raise ValueError(f"orc_reg_analytic failed ({rc})")
    return float(sc[0]), grad


def reg_matrices(g, rp):
    """The 32 V matrices, shape (32, 64, 64)."""
    V = np.zeros((32, 64, 64), np.float64)
    lib().orc_reg_matrices(ctypes.byref(g), ctypes.byref(rp), _ptr(V))
    return V


def volume_gradient(img, spacing, origin=(0.0, 0.0, 0.0), dc=IDENTITY):
    img = np.ascontiguousarray(img, np.float32)
    out = np.zeros(img.shape + (3,), np.float32)
    h = vol_hdr(img, spacing, origin, dc)
    lib().orc_volume_gradient(ctypes.byref(h), _ptr(img), _ptr(out))
    return out


def hist_eqsp(img, bins):
    img = np.ascontiguousarray(img, np.float32)
    off = np.zeros(1, np.float32)
    dl = np.zeros(1, np.float32)
    lib().orc_hist_eqsp(_ptr(img), img.size, bins, _ptr(off), _ptr(dl))
    return float(off[0]), float(dl[0])


def hist_vopt(img, bins, with_keys=False):
    """V-optimal histogram of an image: (offset, delta, key table) of the
    VOPT_KEYS-key grid (and, with_keys, the per-key counts and sums of
    v - offset); None where the reference would not build one (bins
    outside [2, VOPT_KEYS], constant image)."""
    img = np.ascontiguousarray(img, np.float32)
    off = np.zeros(1, np.float32)
    dl = np.zeros(1, np.float32)
    lut = np.zeros(VOPT_KEYS, np.int32)
    cnt = np.zeros(VOPT_KEYS, np.float64)
    sm = np.zeros(VOPT_KEYS, np.float64)
    if lib().orc_hist_vopt(_ptr(img), img.size, bins, _ptr(off), _ptr(dl), _ptr(lut), _ptr(cnt), _ptr(sm)):
        return None
    if with_keys:
        return float(off[0]), float(dl[0]), lut, cnt, sm
    return float(off[0]), float(dl[0]), lut


def vopt_partition(s, ssq, cnt, bins):
    """The partition step of the V-optimal histogram on given prefix tables."""
    s = np.ascontiguousarray(s, np.float64)
    ssq = np.ascontiguousarray(ssq, np.float64)
    cnt = np.ascontiguousarray(cnt, np.float64)
    lut = np.zeros(s.size, np.int32)
    lib().orc_vopt_partition(_ptr(s), _ptr(ssq), _ptr(cnt), s.size, bins, _ptr(lut))
    return lut


def pass1_position_mismatch(fixed, spacing, g, origin=(0.0, 0.0, 0.0), dc=IDENTITY):
    h = vol_hdr(fixed, spacing, origin, dc)
    return lib().orc_pass1_position_mismatch(ctypes.byref(h), ctypes.byref(g))


def pv_quantize(w):
    w = np.ascontiguousarray(w, np.float32)
    q = np.zeros(8, np.int64)
    lib().orc_pv_quantize(_ptr(w), _ptr(q))
    return q


class Case:
    """A fixed/moving pair on one B-spline grid, in the oracle's terms."""

    def __init__(self, fixed, moving, spacing, vox_per_rgn, origin=(0.0, 0.0, 0.0),
                 roi_offset=(0, 0, 0), roi_dim=None, dc=IDENTITY, m_spacing=None, m_origin=None,
                 m_dc=None, fixed_roi=None, moving_roi=None):
        """fixed_roi / moving_roi: optional uchar region-of-interest volumes,
        each (image zyx, spacing, origin, direction cosines)."""
        self.fixed = np.ascontiguousarray(fixed, np.float32)
        self.moving = np.ascontiguousarray(moving, np.float32)
        nz, ny, nx = self.fixed.shape
        self.dim = (nx, ny, nz)
        self.spacing = tuple(spacing)
        self.origin = tuple(origin)
        self.dc = tuple(dc)
        self.m_spacing = tuple(m_spacing or spacing)
        self.m_origin = tuple(m_origin or origin)
        self.m_dc = tuple(m_dc or dc)
        self.g = geom(self.dim, spacing, vox_per_rgn, origin, roi_offset, roi_dim, dc)
        self.fh = vol_hdr(self.fixed, spacing, origin, dc)
        self.mh = vol_hdr(self.moving, self.m_spacing, self.m_origin, self.m_dc)
        self.rdims, self.cdims, self.num_knots = header(self.g)
        self.num_coeff = 3 * self.num_knots
        self._m_grad = None
        self._roi_keep = []
        self.rois = None
        if fixed_roi is not None or moving_roi is not None:
            r = OrcRois()
            for kind, roi in (("fixed", fixed_roi), ("moving", moving_roi)):
                if roi is None:
                    continue
                img = np.ascontiguousarray(roi[0], np.uint8)
                hdr = vol_hdr(img, *roi[1:])
                self._roi_keep += [img, hdr]
                setattr(r, kind + "_hdr", ctypes.pointer(hdr))
                setattr(r, kind, img.ctypes.data)
            self.rois = r

    def _rois(self):
        return ctypes.byref(self.rois) if self.rois is not None else None

    @property
    def m_grad(self):
        if self._m_grad is None:
            self._m_grad = volume_gradient(self.moving, self.m_spacing, self.m_origin, self.m_dc)
        return self._m_grad

    def mse(self, coeff, flavor="g", nthreads=0):
        """Float-faithful flavor g/p/c: returns (score, grad float32, num_vox, raw)."""
        coeff = np.ascontiguousarray(coeff, np.float32)
        grad = np.zeros(self.num_coeff, np.float32)
        raw = np.zeros(1, np.float64)
        sc = np.zeros(1, np.float32)
        nv = np.zeros(1, np.int64)
        mg = self.m_grad if flavor in ("g", "c") else None
        rc = lib().orc_mse_score(ctypes.byref(self.g), ctypes.byref(self.fh), _ptr(self.fixed),
                                 ctypes.byref(self.mh), _ptr(self.moving), _ptr(mg), _ptr(coeff),
                                 ord(flavor), nthreads, self._rois(), _ptr(raw), _ptr(sc), _ptr(grad), _ptr(nv))
        assert rc == 0, rc
        return float(sc[0]), grad, int(nv[0]), float(raw[0])

    def mse_f64(self, coeff, grad_mode=0, nthreads=0):
        """FP64 yardstick: returns (score, grad float64, num_vox)."""
        coeff = np.ascontiguousarray(coeff, np.float32)
        grad = np.zeros(self.num_coeff, np.float64)
        sc = np.zeros(1, np.float64)
        nv = np.zeros(1, np.int64)
        mg = self.m_grad if grad_mode == 0 else None
        rc = lib().orc_mse_score_f64(ctypes.byref(self.g), ctypes.byref(self.fh), _ptr(self.fixed),
                                     ctypes.byref(self.mh), _ptr(self.moving), _ptr(mg), _ptr(coeff),
                                     grad_mode, nthreads, self._rois(), _ptr(sc), _ptr(grad), _ptr(nv))
        assert rc == 0, rc
        return float(sc[0]), grad, int(nv[0])

    def mse_f64_support(self, coeff, ci, delta):
        """FP64 squared-difference sum over coefficient ci's support with it
        moved by delta: returns (raw, n inside, n clamped)."""
        coeff = np.ascontiguousarray(coeff, np.float32)
        raw = np.zeros(1, np.float64)
        n = np.zeros(1, np.int64)
        nc = np.zeros(1, np.int64)
        rc = lib().orc_mse_f64_support(ctypes.byref(self.g), ctypes.byref(self.fh), _ptr(self.fixed),
                                       ctypes.byref(self.mh), _ptr(self.moving), _ptr(coeff), int(ci),
                                       float(delta), self._rois(), _ptr(raw), _ptr(n), _ptr(nc))
        assert rc == 0, rc
        return float(raw[0]), int(n[0]), int(nc[0])

    def fd_mse_f64(self, coeff, idx, h):
        """Central differences of the FP64 MSE (flavor p's score) at
        coefficients idx with step h (mm), from the support sums; returns
        (fd, border): border counts, per coefficient and over both steps,
        the support's clamped voxels plus the voxels that enter or leave the
        overlap (where the difference is not that of the smooth score)."""
        _, _, n_all = self.mse_f64(coeff, 1)
        fd, border = [], []
        for i in idx:
            _, n0, _ = self.mse_f64_support(coeff, i, 0.0)
            rp, np_, cp = self.mse_f64_support(coeff, i, h)
            rm, nm, cm = self.mse_f64_support(coeff, i, -h)
            fd.append((rp - rm) / (2.0 * h) / n_all)
            border.append(cp + cm + abs(np_ - n0) + abs(nm - n0))
        return np.array(fd), np.array(border)

    def hist(self, bins=32, f_bins=None, vopt=False):
        """Histogram parameters of the pair: equal spacing, or V-optimal on
        both sides (vopt=True; Joint_histogram::initialize with
        HIST_VOPT)."""
        fb = f_bins or bins
        if vopt:
            fo, fd, fl = hist_vopt(self.fixed, fb)
            mo, md, ml = hist_vopt(self.moving, bins)
            return OrcHist(fb, bins, fo, fd, mo, md).set_keys(fl, ml)
        fo, fd = hist_eqsp(self.fixed, fb)
        mo, md = hist_eqsp(self.moving, bins)
        return OrcHist(fb, bins, fo, fd, mo, md)

    def mi(self, coeff, hp, nthreads=0):
        """Float-faithful flavor g MI: returns dict with score, grad, num_vox,
        double histograms and the fixed-point (2^-32) histograms."""
        coeff = np.ascontiguousarray(coeff, np.float32)
        fb, mb = hp.fixed_bins, hp.moving_bins
        grad = np.zeros(self.num_coeff, np.float32)
        sc = np.zeros(1, np.float32)
        nv = np.zeros(1, np.int64)
        fh = np.zeros(fb, np.float64)
        mh = np.zeros(mb, np.float64)
        jh = np.zeros(fb * mb, np.float64)
        fxf = np.zeros(fb, np.int64)
        fxj = np.zeros(fb * mb, np.int64)
        rc = lib().orc_mi_score(ctypes.byref(self.g), ctypes.byref(self.fh), _ptr(self.fixed),
                                ctypes.byref(self.mh), _ptr(self.moving), _ptr(coeff), ctypes.byref(hp),
                                nthreads, self._rois(), _ptr(sc), _ptr(grad), _ptr(nv), _ptr(fh), _ptr(mh),
                                _ptr(jh),
                                _ptr(fxf), _ptr(fxj))
        assert rc == 0, rc
        return dict(score=float(sc[0]), grad=grad, num_vox=int(nv[0]), f_hist=fh, m_hist=mh,
                    j_hist=jh, fx_f_hist=fxf, fx_j_hist=fxj)

    def mi_f64(self, coeff, hp, nthreads=0, hist=None):
        """FP64 MI: returns (score, grad float64, num_vox).  With hist (the
        dict of mi()) the score and pass 2 run in FP64 on that float flavor's
        histograms instead of FP64 ones."""
        coeff = np.ascontiguousarray(coeff, np.float32)
        grad = np.zeros(self.num_coeff, np.float64)
        sc = np.zeros(1, np.float64)
        nv = np.zeros(1, np.int64)
        if hist is None:
            rc = lib().orc_mi_score_f64(ctypes.byref(self.g), ctypes.byref(self.fh), _ptr(self.fixed),
                                        ctypes.byref(self.mh), _ptr(self.moving), _ptr(coeff),
                                        ctypes.byref(hp), nthreads, self._rois(), _ptr(sc), _ptr(grad),
                                        _ptr(nv))
        else:
            h = np.ascontiguousarray(np.concatenate([hist["f_hist"], hist["m_hist"], hist["j_hist"]]),
                                     np.float64)
            rc = lib().orc_mi_score_f64_hist(ctypes.byref(self.g), ctypes.byref(self.fh),
                                             _ptr(self.fixed), ctypes.byref(self.mh), _ptr(self.moving),
                                             _ptr(coeff), ctypes.byref(hp), nthreads, self._rois(), _ptr(h),
                                             int(hist["num_vox"]), _ptr(sc), _ptr(grad), _ptr(nv))
        assert rc == 0, rc
        return float(sc[0]), grad, int(nv[0])


def ref_lbfgs_lib():
    """The reference's liblbfgs built from its source (oracle/_ref), or None
    when it was not built (no reference sources where it was compiled)."""
    global _ref_lbfgs
    if _ref_lbfgs is None and os.path.exists(REF_LBFGS_PATH):
        L = ctypes.CDLL(REF_LBFGS_PATH)
        L.lbfgs_parameter_init.argtypes = [ctypes.POINTER(RefLbfgsParameter)]
        L.lbfgs.restype = ctypes.c_int
        L.lbfgs.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, LBFGS_EVAL, LBFGS_PROGRESS,
                            ctypes.c_void_p, ctypes.POINTER(RefLbfgsParameter)]
        _ref_lbfgs = L
    return _ref_lbfgs


ORC_LBFGS_SUCCESS = 0
ORC_LBFGS_STOP = 1
ORC_LBFGS_ALREADY_MINIMIZED = 2
ORC_LBFGSERR_INVALID_EPSILON = -1017
ORC_LBFGSERR_MAXIMUMLINESEARCH = -998
ORC_LBFGSERR_INCREASEGRADIENT = -994

LBFGS_DEFAULTS = dict(m=6, epsilon=1e-5, past=0, delta=1e-5, max_iterations=0, max_linesearch=40,
                      min_step=1e-20, max_step=1e20, ftol=1e-4, gtol=0.9, xtol=1e-16)


class LbfgsRun:
    """Outcome of one L-BFGS run: liblbfgs status, x, fx at exit, and the
    trace -- one entry per evaluation (f, step) and per progress report
    (fx, xnorm, gnorm, step, k, ls)."""

    def __init__(self):
        self.status = None
        self.x = None
        self.fx = None
        self.evals = []
        self.reports = []
        self.it = 0
        self.feval = 0


def lbfgs(fun, x0, impl="oracle", max_its=None, max_feval=None, **parms):
    """Minimize fun (x -> (f, g), float64) from x0 with the oracle's L-BFGS
    (impl="oracle") or the reference's liblbfgs build (impl="ref").

    max_its / max_feval add the reference's progress rule
    (src/register/bspline_optimize_liblbfgs.cxx:55-80): stop once the
    iteration number exceeds max_its or the evaluations exceed max_feval."""
    p = dict(LBFGS_DEFAULTS)
    p.update(parms)
    n = len(x0)
    x = np.array(x0, dtype=np.float64, copy=True)
    fx = np.zeros(1, np.float64)
    run = LbfgsRun()

    def _eval(inst, xp, gp, nn, step):
        xa = np.ctypeslib.as_array(xp, shape=(nn,))
        f, g = fun(xa.copy())
        np.ctypeslib.as_array(gp, shape=(nn,))[:] = g
        run.feval += 1
        run.evals.append((float(f), float(step)))
        return float(f)

    def _progress(inst, xp, gp, f, xnorm, gnorm, step, nn, k, ls):
        run.reports.append((float(f), float(xnorm), float(gnorm), float(step), int(k), int(ls)))
        run.it = k
        if (max_its is not None and run.it > max_its) or (max_feval is not None and run.feval > max_feval):
            return 1
        return 0

    ev, pr = LBFGS_EVAL(_eval), LBFGS_PROGRESS(_progress)
    if impl == "oracle":
        op = OrcLbfgsParms(**p)
        run.status = lib().orc_lbfgs(n, _ptr(x), _ptr(fx), ev, pr, None, ctypes.byref(op))
    elif impl == "ref":
        L = ref_lbfgs_lib()
        if L is None:
            raise FileNotFoundError(REF_LBFGS_PATH)
        rp = RefLbfgsParameter()
        L.lbfgs_parameter_init(ctypes.byref(rp))
        for k, v in p.items():
            setattr(rp, k, v)
        run.status = L.lbfgs(n, _ptr(x), _ptr(fx), ev, pr, None, ctypes.byref(rp))
    else:
        raise ValueError(impl)
    run.x, run.fx = x, float(fx[0])
    return run
