"""Python host side of the ``threading=hip`` B-spline flavor.

Mirrors the reference objects the score path reads and writes:

* ``Volume``          -- src/base/volume.h:39-57 (dim, origin, spacing,
                         direction cosines, voxels x-fastest);
* ``Bspline_xform``   -- src/base/bspline_xform.h:22-96 / bspline_header.cxx:28-62
                         (grid geometry and interleaved coefficients);
* ``Joint_histogram`` -- src/register/joint_histogram.h:14-46 (EQSP bins);
* ``Bspline_score``   -- src/register/bspline_score.h:33-95 (curr_smetric,
                         curr_smetric_grad, curr_num_vox);
* ``update_roi`` / ``stage_rois`` -- the stage's region-of-interest
                         preparation (src/register/bspline_stage.cxx:71-205);
* ``BsplineHip``      -- the device state the reference keeps behind
                         ``Bspline_state::dev_ptrs`` (bspline_state.cxx:165-271),
                         created once per stage, scored once per evaluation.

Every evaluation runs on the GPU through libplmregisterhip.so; there is no
CPU path here.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _abi

IDENTITY = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)

METRICS = {"mse": _abi.PLM_METRIC_MSE, "mi": _abi.PLM_METRIC_MI}
GRAD_MODES = {"ref_nn": _abi.PLM_GRAD_REF_NN, "exact": _abi.PLM_GRAD_EXACT, "none": _abi.PLM_GRAD_NONE}
STORAGE = {"f32": _abi.PLM_STORE_F32, "bf16": _abi.PLM_STORE_BF16, "u8": _abi.PLM_STORE_U8}


def _is_torch(x):
    return type(x).__module__.startswith("torch")


class Volume:
    """A float image (numpy array or CUDA torch tensor), shape (nz, ny, nx[, 3]);
    a uint8 array is a region of interest (PT_UCHAR)."""

    def __init__(self, img, spacing, origin=(0.0, 0.0, 0.0), direction_cosines=IDENTITY):
        if _is_torch(img):
            self.img = img
        elif np.asarray(img).dtype == np.uint8:
            self.img = np.ascontiguousarray(img)
        else:
            self.img = np.ascontiguousarray(img, np.float32)
        nz, ny, nx = img.shape[:3]
        self.dim = (int(nx), int(ny), int(nz))
        self.spacing = tuple(float(s) for s in spacing)
        self.origin = tuple(float(o) for o in origin)
        self.direction_cosines = tuple(float(d) for d in direction_cosines)

    @property
    def npix(self):
        return self.dim[0] * self.dim[1] * self.dim[2]

    @property
    def is_uchar(self):
        return str(self.img.dtype).endswith("uint8")

    def desc(self, storage="f32"):
        if self.is_uchar:
            storage = "u8"
        if _is_torch(self.img):
            if not self.img.is_cuda or not self.img.is_contiguous():
                raise ValueError("device images must be contiguous CUDA tensors")
            ptr, loc = self.img.data_ptr(), _abi.PLM_MEM_DEVICE
        else:
            ptr, loc = self.img.ctypes.data, _abi.PLM_MEM_HOST
        return _abi.PlmImgDesc(ctypes.c_void_p(ptr), loc, STORAGE[storage], _abi.c_int64_3(*self.dim),
                               _abi.c_float3(*self.origin), _abi.c_float3(*self.spacing),
                               _abi.c_float9(*self.direction_cosines))


class Bspline_xform:
    """B-spline grid over a fixed image (Bspline_header::set semantics)."""

    def __init__(self, img_origin, img_spacing, img_dim, vox_per_rgn, roi_offset=(0, 0, 0),
                 roi_dim=None, direction_cosines=IDENTITY):
        self.img_origin = tuple(float(v) for v in img_origin)
        self.img_spacing = tuple(float(v) for v in img_spacing)
        self.img_dim = tuple(int(v) for v in img_dim)
        self.vox_per_rgn = tuple(int(v) for v in vox_per_rgn)
        self.roi_offset = tuple(int(v) for v in roi_offset)
        self.roi_dim = tuple(int(v) for v in (roi_dim or img_dim))
        self.direction_cosines = tuple(direction_cosines)
        self.grid_spac = tuple(v * abs(s) for v, s in zip(self.vox_per_rgn, self.img_spacing))
        self.rdims = tuple(1 + (r - 1) // v for r, v in zip(self.roi_dim, self.vox_per_rgn))
        self.cdims = tuple(r + 3 for r in self.rdims)
        self.num_knots = self.cdims[0] * self.cdims[1] * self.cdims[2]
        self.num_coeff = 3 * self.num_knots
        self.coeff = np.zeros(self.num_coeff, np.float32)

    @classmethod
    def for_volume(cls, vol, vox_per_rgn, **kw):
        return cls(vol.origin, vol.spacing, vol.dim, vox_per_rgn,
                   direction_cosines=vol.direction_cosines, **kw)

    def jitter_if_zero(self):
        """All-zero coefficients become 0.01 (src/base/bspline_xform.cxx:630-644)."""
        if not np.any(self.coeff != 0.0):
            self.coeff[:] = np.float32(0.01)

    def save(self, path):
        """Write the MGH_GPUIT_BSP text file of Bspline_xform::save
        (src/base/bspline_xform.cxx:101-145): the header, then the
        coefficients planar (all x, all y, all z), one "%.20f" per line."""
        lines = ["MGH_GPUIT_BSP <experimental>",
                 "img_origin = %f %f %f" % self.img_origin,
                 "img_spacing = %f %f %f" % self.img_spacing,
                 "img_dim = %u %u %u" % self.img_dim,
                 "roi_offset = %d %d %d" % self.roi_offset,
                 "roi_dim = %d %d %d" % self.roi_dim,
                 "vox_per_rgn = %d %d %d" % self.vox_per_rgn,
                 "direction_cosines = " + " ".join("%f" % v for v in self.direction_cosines)]
        c = np.asarray(self.coeff, np.float32).reshape(-1, 3)
        for d in range(3):
            lines.extend("%.20f" % v for v in c[:, d].tolist())
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")

    @classmethod
    def load(cls, path):
        """Read an MGH_GPUIT_BSP file as bspline_xform_load does
        (src/base/bspline_xform.cxx:147-279): header keys in any order, the
        direction cosines optional (identity), then the planar
        coefficients.  Raises ValueError where the reference returns 0."""
        with open(path) as f:
            lines = f.read().split("\n")
        if not lines or not lines[0].startswith("MGH_GPUIT_BSP"):
            raise ValueError(f"{path}: not an MGH_GPUIT_BSP file")
        h = {"img_origin": (0.0, 0.0, 0.0), "img_spacing": (1.0, 1.0, 1.0), "img_dim": (0, 0, 0),
             "roi_offset": (0, 0, 0), "roi_dim": (0, 0, 0), "vox_per_rgn": (0, 0, 0),
             "direction_cosines": IDENTITY}
        counts = {"img_origin": 3, "img_spacing": 3, "img_dim": 3, "roi_offset": 3, "roi_dim": 3,
                  "vox_per_rgn": 3, "direction_cosines": 9}
        i = 1
        while True:
            if i >= len(lines) or (i == len(lines) - 1 and lines[i] == ""):
                raise ValueError(f"{path}: header ends without coefficients")
            line = lines[i]
            if "=" not in line:
                break
            key, _, val = line.partition("=")
            key = key.strip()
            if key not in counts:
                raise ValueError(f"{path}: unknown header line {line!r}")
            vals = val.split()
            if len(vals) < counts[key]:
                raise ValueError(f"{path}: short header line {line!r}")
            vals = vals[:counts[key]]
            if key in ("img_origin", "img_spacing", "direction_cosines"):
                h[key] = tuple(float(np.float32(v)) for v in vals)
            else:
                h[key] = tuple(int(v) for v in vals)
            i += 1
        bxf = cls(h["img_origin"], h["img_spacing"], h["img_dim"], h["vox_per_rgn"], h["roi_offset"],
                  h["roi_dim"], h["direction_cosines"])
        if bxf.num_coeff < 1:
            raise ValueError(f"{path}: no coefficients")
        body = lines[i:i + bxf.num_coeff]
        if len(body) < bxf.num_coeff or any(not t.strip() for t in body):
            raise ValueError(f"{path}: {len(body)} coefficient lines for {bxf.num_coeff}")
        planar = np.array([np.float32(float(t.split()[0])) for t in body], np.float32)
        bxf.coeff = planar.reshape(3, -1).T.reshape(-1).copy()
        return bxf

    def geom(self):
        return _abi.PlmBspGeom(_abi.c_float3(*self.img_origin), _abi.c_float3(*self.img_spacing),
                               _abi.c_int64_3(*self.img_dim), _abi.c_int64_3(*self.roi_offset),
                               _abi.c_int64_3(*self.roi_dim), _abi.c_int64_3(*self.vox_per_rgn),
                               _abi.c_float9(*self.direction_cosines))


class Joint_histogram:
    """Joint histogram parameters (src/register/histogram.h:14-35,
    joint_histogram.cxx:115-139, 195-346): equal spacing ("eqsp") or
    V-optimal ("vopt", the stage's histogram_type), per side offset and
    delta and, for V-optimal bins, the key table."""

    def __init__(self, fixed_bins=32, moving_bins=32, hist_type="eqsp"):
        if hist_type not in ("eqsp", "vopt"):
            raise ValueError("hist_type is 'eqsp' or 'vopt'")
        self.fixed_bins = int(fixed_bins)
        self.moving_bins = int(moving_bins)
        self.fixed_type = self.moving_type = hist_type
        self.fixed_offset = self.fixed_delta = 0.0
        self.moving_offset = self.moving_delta = 0.0
        self.fixed_keys = self.moving_keys = None

    def initialize(self, fixed, moving, gpuid=0):
        """Bin offsets and widths (and V-optimal key tables) from each image,
        computed on the GPU (bspline_initialize_mi_hist).  As the reference
        does, a side asking for more V-optimal bins than there are keys falls
        back to equal spacing."""
        out = []
        for vol, bins, kind in ((fixed, self.fixed_bins, self.fixed_type),
                                (moving, self.moving_bins, self.moving_type)):
            if kind == "vopt" and bins > _abi.VOPT_KEYS:
                print(f"WARNING: Falling back to EQSP histograms.\n         (Reason: # bins > {_abi.VOPT_KEYS})")
                kind = "eqsp"
            if kind == "vopt":
                out.append((kind,) + hist_vopt(vol, bins, gpuid))
            else:
                out.append((kind,) + hist_eqsp(vol, bins, gpuid) + (None,))
        (self.fixed_type, self.fixed_offset, self.fixed_delta, self.fixed_keys), \
            (self.moving_type, self.moving_offset, self.moving_delta, self.moving_keys) = out
        return self

    def desc(self):
        return _abi.PlmHistDesc(self.fixed_bins, self.moving_bins, self.fixed_offset,
                                self.fixed_delta, self.moving_offset, self.moving_delta)


@dataclass
class Bspline_score:
    """One evaluation's result: curr_smetric / curr_smetric_grad / curr_num_vox."""
    score: float
    grad: np.ndarray          # float64, interleaved, already normalized
    num_vox: int
    hist: np.ndarray = None   # MI: Bf*Bm joint bins (2^-32 units) then Bf counts
    status: int = 0

    @property
    def curr_smetric(self):
        return np.float32(self.score)

    @property
    def curr_smetric_grad(self):
        return self.grad.astype(np.float32)


def bf16_round(a):
    """float32 -> bfloat16 -> float32 with round to nearest even: what
    PLM_STORE_BF16 keeps of an image (NaN payloads aside)."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(np.shape(a))


def hist_eqsp(vol, bins, gpuid=0):
    """Equal-spacing histogram (offset, delta) of a volume: min/max scan on
    the GPU (src/register/joint_histogram.cxx:115-139)."""
    lib = _abi.load()
    d = vol.desc()
    off, dl = ctypes.c_float(), ctypes.c_float()
    _abi.check(lib.plm_hip_hist_eqsp(ctypes.byref(d), int(bins), ctypes.byref(off), ctypes.byref(dl), gpuid))
    return float(off.value), float(dl.value)


def hist_vopt(vol, bins, gpuid=0):
    """V-optimal histogram of a volume (src/register/joint_histogram.cxx:
    195-324): (offset, delta, key table) on the GPU."""
    lib = _abi.load()
    d = vol.desc()
    off, dl = ctypes.c_float(), ctypes.c_float()
    lut = np.zeros(_abi.VOPT_KEYS, np.int32)
    _abi.check(lib.plm_hip_hist_vopt(ctypes.byref(d), int(bins), ctypes.byref(off), ctypes.byref(dl),
                                     lut.ctypes.data_as(ctypes.c_void_p), gpuid))
    return float(off.value), float(dl.value), lut


def resample_geometry(vol, kind, rate):
    """registration_resample_volume's output geometry for a stage:
    (identity, dim xyz, origin, spacing)."""
    lib = _abi.load()
    d = vol.desc()
    r = np.ascontiguousarray(rate, np.float32)
    dim = np.zeros(3, np.int64)
    o = np.zeros(3, np.float32)
    sp = np.zeros(3, np.float32)
    ident = ctypes.c_int()
    _abi.check(lib.plm_hip_resample_geometry(_abi.PLM_RESAMPLE[kind], ctypes.byref(d), r.ctypes.data,
                                             dim.ctypes.data, o.ctypes.data, sp.ctypes.data, ctypes.byref(ident)))
    return bool(ident.value), tuple(int(v) for v in dim), tuple(float(v) for v in o), tuple(float(v) for v in sp)


def _new_like(vol, shape):
    if _is_torch(vol.img):
        import torch
        return torch.empty(shape, dtype=torch.float32, device=vol.img.device), _abi.PLM_MEM_DEVICE
    return np.empty(shape, np.float32), _abi.PLM_MEM_HOST


def _ptr_of(a):
    return ctypes.c_void_p(a.data_ptr() if _is_torch(a) else a.ctypes.data)


def volume_resample(vol, dim, origin, spacing, nearest=False, gpuid=0):
    """vol resampled onto dim (x, y, z) / origin / spacing on the GPU,
    trilinear or nearest voxel (volume_resample_float_li / _nn)."""
    lib = _abi.load()
    d = vol.desc()
    out, loc = _new_like(vol, (dim[2], dim[1], dim[0]))
    # named arrays: they must outlive the call that reads them
    a_dim = np.asarray(dim, np.int64)
    a_org = np.asarray(origin, np.float32)
    a_sp = np.asarray(spacing, np.float32)
    _abi.check(lib.plm_hip_volume_resample(ctypes.byref(d), a_dim.ctypes.data, a_org.ctypes.data,
                                           a_sp.ctypes.data, 1 if nearest else 0, _ptr_of(out), loc, gpuid))
    return Volume(out, spacing, origin, vol.direction_cosines)


def stage_resample(vol, kind, rate, nearest=False, gpuid=0):
    """registration_resample_volume on the GPU: the stage's subsampled
    volume (kind: voxel_rate, voxel_legacy, mm, pct)."""
    lib = _abi.load()
    ident, dim, origin, spacing = resample_geometry(vol, kind, rate)
    d = vol.desc()
    out, loc = _new_like(vol, (dim[2], dim[1], dim[0]))
    a_rate = np.asarray(rate, np.float32)
    _abi.check(lib.plm_hip_stage_resample(_abi.PLM_RESAMPLE[kind], a_rate.ctypes.data, ctypes.byref(d),
                                          1 if nearest else 0, _ptr_of(out), loc, gpuid))
    return Volume(out, spacing, origin, vol.direction_cosines)


def volume_gradient(vol, gpuid=0):
    """Moving-image gradient on the GPU (src/base/volume_grad.cxx:9-76)."""
    lib = _abi.load()
    d = vol.desc()
    if _is_torch(vol.img):
        import torch
        out = torch.empty(vol.img.shape + (3,), dtype=torch.float32, device=vol.img.device)
        _abi.check(lib.plm_hip_volume_gradient(ctypes.byref(d), ctypes.c_void_p(out.data_ptr()),
                                               _abi.PLM_MEM_DEVICE, gpuid))
        return Volume(out, vol.spacing, vol.origin, vol.direction_cosines)
    out = np.empty(vol.img.shape + (3,), np.float32)
    _abi.check(lib.plm_hip_volume_gradient(ctypes.byref(d), out.ctypes.data_as(ctypes.c_void_p),
                                           _abi.PLM_MEM_HOST, gpuid))
    return Volume(out, vol.spacing, vol.origin, vol.direction_cosines)


# Regularization_parms defaults (src/register/regularization_parms.h:31-42)
REG_DEFAULTS = {"total_displacement_penalty": 0.0, "diffusion_penalty": 0.0, "curvature_penalty": 0.0,
                "lame_coefficient_1": 0.0, "lame_coefficient_2": 0.0, "linear_elastic_multiplier": 1.0,
                "third_order_penalty": 0.0, "curvature_mixed_weight": 1.0}


@dataclass
class LbfgsResult:
    """Outcome of plm_hip_bspline_optimize_lbfgs: liblbfgs's status, the
    iteration number and evaluation count at exit (Bspline_state::it and
    ::feval), the objective lbfgs() returns, and every progress report
    (fx, xnorm, gnorm, step, k, ls)."""
    status: int
    iterations: int
    fevals: int
    fx: float
    reports: list


class BsplineHip:
    """Device state for one stage: images resident in HBM, scored per evaluation."""

    def __init__(self, bxf, fixed, moving, moving_grad=None, hist=None, gpuid=0, storage="f32"):
        self.lib = _abi.load()
        self.bxf = bxf
        self.hist = hist
        self.storage = storage
        self.gpuid = gpuid
        fd = fixed.desc(storage)
        md = moving.desc(storage)
        gd = moving_grad.desc(storage) if moving_grad is not None else None
        hd = hist.desc() if hist is not None else None
        ctx = ctypes.c_void_p()
        _abi.check(self.lib.plm_hip_bspline_create(
            ctypes.byref(bxf.geom()), ctypes.byref(fd), ctypes.byref(md),
            ctypes.byref(gd) if gd is not None else None,
            ctypes.byref(hd) if hd is not None else None, gpuid, ctypes.byref(ctx)))
        self.ctx = ctx
        n = ctypes.c_int64()
        _abi.check(self.lib.plm_hip_bspline_num_knots(ctx, ctypes.byref(n)))
        if n.value != bxf.num_knots:
            raise RuntimeError("library and host disagree on the knot count")
        _abi.check(self.lib.plm_hip_bspline_num_tiles(ctx, ctypes.byref(n)))
        self.num_tiles = n.value
        if hist is not None and (hist.fixed_keys is not None or hist.moving_keys is not None):
            keys = [None if k is None else np.ascontiguousarray(k, np.int32)
                    for k in (hist.fixed_keys, hist.moving_keys)]
            _abi.check(self.lib.plm_hip_bspline_set_hist_keys(
                ctx, *[None if k is None else k.ctypes.data_as(ctypes.c_void_p) for k in keys]))

    @property
    def fixed_voxels(self):
        """Voxels of the fixed image (a displacement field holds 3 floats each)."""
        d = self.bxf.img_dim
        return d[0] * d[1] * d[2]

    @property
    def hist_size(self):
        if self.hist is None:
            return 0
        return self.hist.fixed_bins * self.hist.moving_bins + self.hist.fixed_bins

    def score(self, coeff=None, metric="mse", grad_mode="ref_nn", want_hist=False):
        """One evaluation with host buffers (the CUDA_bspline_mse_j contract)."""
        coeff = self.bxf.coeff if coeff is None else coeff
        coeff = np.ascontiguousarray(coeff, np.float32)
        if coeff.size != self.bxf.num_coeff:
            raise ValueError("coefficient vector has the wrong length")
        grad = np.empty(self.bxf.num_coeff, np.float64)
        sc = ctypes.c_double()
        nv = ctypes.c_int64()
        hist = np.empty(self.hist_size, np.int64) if want_hist else None
        rc = _abi.check(self.lib.plm_hip_bspline_score(
            self.ctx, METRICS[metric], GRAD_MODES[grad_mode], coeff.ctypes.data_as(ctypes.c_void_p),
            ctypes.byref(sc), grad.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nv),
            hist.ctypes.data_as(ctypes.c_void_p) if hist is not None else None))
        return Bspline_score(sc.value, grad, nv.value, hist, rc)

    def set_roi(self, fixed_roi=None, moving_roi=None):
        """Set (or, with neither, clear) the fixed and moving regions of
        interest, uint8 Volumes with geometries of their own
        (Bspline_state::fixed_roi / moving_roi): MSE then follows flavor i
        (r for grad_mode "exact"), MI flavor h."""
        for r in (fixed_roi, moving_roi):
            if r is not None and not r.is_uchar:
                raise ValueError("a region of interest is a uint8 Volume")
        fd = fixed_roi.desc() if fixed_roi is not None else None
        md = moving_roi.desc() if moving_roi is not None else None
        _abi.check(self.lib.plm_hip_bspline_set_roi(self.ctx, ctypes.byref(fd) if fd is not None else None,
                                                   ctypes.byref(md) if md is not None else None))
        self.fixed_roi, self.moving_roi = fixed_roi, moving_roi

    def score_device(self, d_coeff, d_score, d_grad, d_num_vox, metric="mse", grad_mode="ref_nn",
                     d_hist=None, stream=None):
        """Enqueue one evaluation with device tensors; returns immediately."""
        ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
        st = ctypes.c_void_p(stream) if stream else None
        return _abi.check(self.lib.plm_hip_bspline_score_device(
            self.ctx, METRICS[metric], GRAD_MODES[grad_mode], ptr(d_coeff), ptr(d_score),
            ptr(d_grad), ptr(d_num_vox), ptr(d_hist), st))

    def interpolate_vf(self, d_coeff, d_vf, stream=None):
        """Dense displacement field (mm, float3 per fixed voxel) into d_vf."""
        st = ctypes.c_void_p(stream) if stream else None
        return _abi.check(self.lib.plm_hip_bspline_interpolate_vf(
            self.ctx, ctypes.c_void_p(d_coeff.data_ptr()), ctypes.c_void_p(d_vf.data_ptr()), st))

    def regularize(self, d_coeff, d_score, d_grad, accumulate=False, stream=None, **weights):
        """Analytic smoothness regularization (flavor c): score into d_score,
        gradient into (or added to) d_grad.  Weights by their
        Regularization_parms names; unset ones take the reference defaults
        (0, except linear_elastic_multiplier = curvature_mixed_weight = 1)."""
        w = dict(REG_DEFAULTS)
        unknown = set(weights) - set(w)
        if unknown:
            raise ValueError(f"unknown regularization weights: {sorted(unknown)}")
        w.update(weights)
        rp = _abi.PlmRegParms(**w)
        st = ctypes.c_void_p(stream) if stream else None
        ptr = lambda t: ctypes.c_void_p(t.data_ptr())
        return _abi.check(self.lib.plm_hip_bspline_regularize(
            self.ctx, ctypes.byref(rp), ptr(d_coeff), ptr(d_score), ptr(d_grad), 1 if accumulate else 0, st))

    def warp(self, d_coeff, d_out, default_val=0.0, linear=True, stream=None):
        """Warped moving image (one float per fixed voxel) into d_out: trilinear
        or nearest voxel, default_val outside the moving image and the ROI
        (bspline_warp)."""
        st = ctypes.c_void_p(stream) if stream else None
        return _abi.check(self.lib.plm_hip_bspline_warp(
            self.ctx, ctypes.c_void_p(d_coeff.data_ptr()), ctypes.c_void_p(d_out.data_ptr()),
            float(default_val), 1 if linear else 0, st))

    def optimize_lbfgs(self, d_x, d_coeff=None, metric="mse", grad_mode="ref_nn", reg=None, progress=None,
                       record=True, resident=False, max_reports=4096, **parms):
        """L-BFGS on the device: minimize the score (+ the analytic
        regularization with weights ``reg``, a dict of Regularization_parms
        names) from the float64 device tensor d_x, which holds the x liblbfgs
        returns at exit; d_coeff (float32 device tensor, optional) gets the
        coefficients of the last evaluation.  ``parms``: plm_lbfgs_parms
        fields (liblbfgs names, max_its, max_feval; unset ones take liblbfgs's
        defaults and max_its = max_feval = 10).  ``progress``: callable
        (d_x_ptr, d_g_ptr, fx, xnorm, gnorm, step, n, k, ls) -> int, nonzero
        stops (the host-driven loop, plm_hip_bspline_optimize_lbfgs).
        ``resident``: the whole loop on the device, the evaluation chain
        replayed from a hipGraph (plm_hip_bspline_optimize_lbfgs_resident),
        its reports read back afterwards (the first ``max_reports``; no
        progress callable).  ``record``: keep every progress report in the
        result.  Returns an LbfgsResult."""
        if d_x.numel() != self.bxf.num_coeff:
            raise ValueError("d_x has the wrong length")
        p = _abi.PlmLbfgsParms()
        self.lib.plm_hip_lbfgs_parms_init(ctypes.byref(p))
        for k, v in parms.items():
            if not hasattr(p, k) or k.startswith("_"):
                raise ValueError(f"unknown L-BFGS parameter {k}")
            setattr(p, k, v)
        rp = None
        if reg is not None:
            w = dict(REG_DEFAULTS)
            unknown = set(reg) - set(w)
            if unknown:
                raise ValueError(f"unknown regularization weights: {sorted(unknown)}")
            w.update(reg)
            rp = _abi.PlmRegParms(**w)
        reports = []
        if resident:
            if progress is not None:
                raise ValueError("a progress callable needs the host loop (resident=False)")
            cap = int(max_reports) if record else 0
            buf = (_abi.PlmLbfgsReport * max(cap, 1))()
            nrep = ctypes.c_int()
            res = _abi.PlmLbfgsResult()
            _abi.check(self.lib.plm_hip_bspline_optimize_lbfgs_resident(
                self.ctx, METRICS[metric], GRAD_MODES[grad_mode], ctypes.byref(rp) if rp is not None else None,
                ctypes.byref(p), ctypes.c_void_p(d_x.data_ptr()),
                ctypes.c_void_p(d_coeff.data_ptr()) if d_coeff is not None else None,
                ctypes.cast(buf, ctypes.c_void_p) if cap else None, cap, ctypes.byref(nrep), ctypes.byref(res)))
            reports = [(r.fx, r.xnorm, r.gnorm, r.step, r.k, r.ls) for r in buf[:min(cap, nrep.value)]]
            return LbfgsResult(res.status, res.iterations, res.fevals, res.fx, reports)

        def _progress(inst, px, pg, fx, xnorm, gnorm, step, n, k, ls):
            reports.append((fx, xnorm, gnorm, step, k, ls))
            return int(progress(px, pg, fx, xnorm, gnorm, step, n, k, ls)) if progress else 0

        cb = _abi.PLM_LBFGS_PROGRESS(_progress) if (progress or record) else _abi.PLM_LBFGS_PROGRESS()
        res = _abi.PlmLbfgsResult()
        _abi.check(self.lib.plm_hip_bspline_optimize_lbfgs(
            self.ctx, METRICS[metric], GRAD_MODES[grad_mode], ctypes.byref(rp) if rp is not None else None,
            ctypes.byref(p), ctypes.c_void_p(d_x.data_ptr()),
            ctypes.c_void_p(d_coeff.data_ptr()) if d_coeff is not None else None, cb, None, ctypes.byref(res)))
        return LbfgsResult(res.status, res.iterations, res.fevals, res.fx, reports)

    def score_vf_device(self, d_vf, d_score, d_grad, d_num_vox, stream=None):
        """MSE score and exact-derivative gradient from a stored displacement field."""
        ptr = lambda t: ctypes.c_void_p(t.data_ptr())
        st = ctypes.c_void_p(stream) if stream else None
        return _abi.check(self.lib.plm_hip_bspline_score_vf_device(
            self.ctx, ptr(d_vf), ptr(d_score), ptr(d_grad), ptr(d_num_vox), st))

    def set_tuning(self, vox_per_thread):
        _abi.check(self.lib.plm_hip_bspline_set_tuning(self.ctx, int(vox_per_thread)))

    def set_occupancy(self, workgroups_per_cu):
        """Tile-kernel workgroups per compute unit (0: as many as fit)."""
        _abi.check(self.lib.plm_hip_bspline_set_occupancy(self.ctx, int(workgroups_per_cu)))

    def get_tuning(self):
        """The current launch shape: (vox_per_thread, workgroups_per_cu)."""
        v = ctypes.c_int()
        w = ctypes.c_int()
        _abi.check(self.lib.plm_hip_bspline_get_tuning(self.ctx, ctypes.byref(v), ctypes.byref(w)))
        return v.value, w.value

    def autotune(self, metric="mse", grad_mode="ref_nn", d_coeff=None, reps=5):
        """Pick the fastest launch shape on this stage's data; returns (choice, ms/eval)."""
        ch = ctypes.c_int()
        ms = ctypes.c_double()
        ptr = ctypes.c_void_p(d_coeff.data_ptr()) if d_coeff is not None else None
        _abi.check(self.lib.plm_hip_bspline_autotune(self.ctx, METRICS[metric], GRAD_MODES[grad_mode],
                                                      ptr, reps, ctypes.byref(ch), ctypes.byref(ms)))
        return ch.value, ms.value

    def set_timing(self, enable=True):
        _abi.check(self.lib.plm_hip_bspline_set_timing(self.ctx, 1 if enable else 0))

    def get_timing(self, max_stages=6):
        """Summed per-stage milliseconds since the last call, and the eval count."""
        ms = np.zeros(max_stages, np.float64)
        n = ctypes.c_int64()
        _abi.check(self.lib.plm_hip_bspline_get_timing(self.ctx, ms.ctypes.data_as(ctypes.c_void_p),
                                                        max_stages, ctypes.byref(n)))
        return ms, n.value

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.plm_hip_bspline_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def update_roi(roi, image, min_val, max_val):
    """update_roi (src/register/bspline_stage.cxx:71-93): with no ROI, a new
    one on the image's grid that holds 1 where min_val <= value <= max_val;
    otherwise the ROI with its voxels outside that range set to 0 (the ROI
    shares the image's grid, voxel for voxel)."""
    img = np.asarray(image.img)
    in_range = (img >= min_val) & (img <= max_val)
    if roi is None:
        return Volume(in_range.astype(np.uint8), image.spacing, image.origin, image.direction_cosines)
    r = np.asarray(roi.img)
    if r.shape != img.shape:
        raise ValueError("update_roi walks the ROI and the image voxel for voxel: dims must match")
    out = np.where(in_range, r, np.uint8(0)).astype(np.uint8)
    return Volume(out, roi.spacing, roi.origin, roi.direction_cosines)


def stage_rois(fixed, moving, fixed_roi=None, moving_roi=None, mi_fixed_image_minVal=0.0,
               mi_fixed_image_maxVal=0.0, mi_moving_image_minVal=0.0, mi_moving_image_maxVal=0.0):
    """The ROIs a stage hands the score path (src/register/bspline_stage.cxx:
    129-192): each image's user intensity range, when set (min or max
    non-zero, min < max), narrows that image's ROI, or makes one; returns
    (fixed_roi, moving_roi), either possibly None.  (The stage then
    subsamples them by the image's rate, nearest voxel: stage_resample with
    nearest=True.)"""
    def ranged(lo, hi):
        return (lo != 0 or hi != 0) and lo < hi
    if ranged(mi_moving_image_minVal, mi_moving_image_maxVal):
        moving_roi = update_roi(moving_roi, moving, mi_moving_image_minVal, mi_moving_image_maxVal)
    if ranged(mi_fixed_image_minVal, mi_fixed_image_maxVal):
        fixed_roi = update_roi(fixed_roi, fixed, mi_fixed_image_minVal, mi_fixed_image_maxVal)
    return fixed_roi, moving_roi


def stage_for_pair(pair, gpuid=0, storage="f32", with_grad=True, bins=None, moving_grad=None,
                   hist_type="eqsp"):
    """Build Volumes, grid, optional histogram (hist_type "eqsp" or "vopt")
    and the device state for a synth.Pair."""
    fixed = Volume(pair.fixed, pair.spacing, pair.origin, pair.direction_cosines)
    moving = Volume(pair.moving, pair.spacing, pair.origin, pair.direction_cosines)
    bxf = Bspline_xform.for_volume(fixed, pair.vox_per_rgn)
    mg = None
    if with_grad:
        mg = moving_grad if moving_grad is not None else volume_gradient(moving, gpuid)
    hist = Joint_histogram(bins, bins, hist_type).initialize(fixed, moving) if bins else None
    return BsplineHip(bxf, fixed, moving, mg, hist, gpuid, storage)


class BsplineBatch:
    """Many stages evaluated together: one persistent tile-kernel launch over
    the flattened tiles of every pair plus one reduce launch, optionally
    captured once in a hipGraph and replayed (the multi-atlas case; the
    C ABI's plm_hip_batch_*).  The stages must share the device, the storage
    type and vox_per_rgn, and must outlive the batch."""

    def __init__(self, stages):
        self.lib = _abi.load()
        self.stages = list(stages)
        arr = (ctypes.c_void_p * len(self.stages))(*[s.ctx.value for s in self.stages])
        b = ctypes.c_void_p()
        _abi.check(self.lib.plm_hip_batch_create(arr, len(self.stages), ctypes.byref(b)))
        self.b = b
        self._keep = None

    def set_buffers(self, d_coeffs, d_grads, d_score, d_num_vox):
        """Device tensors: per-pair coefficient and FP64 gradient vectors, and
        score (float64) and voxel-count (int64) vectors of one entry per pair."""
        n = len(self.stages)
        if len(d_coeffs) != n or len(d_grads) != n or d_score.numel() != n or d_num_vox.numel() != n:
            raise ValueError("one buffer per pair")
        for s, c, g in zip(self.stages, d_coeffs, d_grads):
            if c.numel() != s.bxf.num_coeff or g.numel() != s.bxf.num_coeff:
                raise ValueError("buffer length differs from the pair's coefficient count")
        ca = (ctypes.c_void_p * n)(*[c.data_ptr() for c in d_coeffs])
        ga = (ctypes.c_void_p * n)(*[g.data_ptr() for g in d_grads])
        _abi.check(self.lib.plm_hip_batch_set_buffers(
            self.b, ca, ga, ctypes.c_void_p(d_score.data_ptr()), ctypes.c_void_p(d_num_vox.data_ptr())))
        self._keep = (list(d_coeffs), list(d_grads), d_score, d_num_vox)

    def score_device(self, grad_mode="ref_nn", stream=None):
        st = ctypes.c_void_p(stream) if stream else None
        return _abi.check(self.lib.plm_hip_batch_score_device(self.b, GRAD_MODES[grad_mode], st))

    def capture(self, grad_mode="ref_nn"):
        _abi.check(self.lib.plm_hip_batch_capture(self.b, GRAD_MODES[grad_mode]))

    def launch(self, stream=None):
        st = ctypes.c_void_p(stream) if stream else None
        return _abi.check(self.lib.plm_hip_batch_launch(self.b, st))

    def set_tuning(self, vox_per_thread):
        _abi.check(self.lib.plm_hip_batch_set_tuning(self.b, int(vox_per_thread)))

    def set_occupancy(self, workgroups_per_cu):
        """Tile-kernel workgroups per compute unit (0: as many as fit)."""
        _abi.check(self.lib.plm_hip_batch_set_occupancy(self.b, int(workgroups_per_cu)))

    def get_tuning(self):
        """The current launch shape: (vox_per_thread, workgroups_per_cu)."""
        v = ctypes.c_int()
        w = ctypes.c_int()
        _abi.check(self.lib.plm_hip_batch_get_tuning(self.b, ctypes.byref(v), ctypes.byref(w)))
        return v.value, w.value

    def autotune(self, grad_mode="ref_nn", reps=5):
        ch = ctypes.c_int()
        ms = ctypes.c_double()
        _abi.check(self.lib.plm_hip_batch_autotune(self.b, GRAD_MODES[grad_mode], reps,
                                                    ctypes.byref(ch), ctypes.byref(ms)))
        return ch.value, ms.value

    def set_timing(self, enable=True):
        _abi.check(self.lib.plm_hip_batch_set_timing(self.b, 1 if enable else 0))

    def get_timing(self, max_stages=6):
        ms = np.zeros(max_stages, np.float64)
        n = ctypes.c_int64()
        _abi.check(self.lib.plm_hip_batch_get_timing(self.b, ms.ctypes.data_as(ctypes.c_void_p),
                                                      max_stages, ctypes.byref(n)))
        return ms, n.value

    def close(self):
        if getattr(self, "b", None):
            self.lib.plm_hip_batch_destroy(self.b)
            self.b = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
