"""The oracle's stage subsampling (oracle/resample_oracle.c) and EQSP
scan, pinned by what the mathematics fixes.

* Geometry: the closed forms of volume_subsample_vox / _legacy,
  volume_resample_spacing / _percent (src/base/volume_resample.cxx:
  298-413) on cases worked by hand, the clone at rate 1 and the refusal of
  rates that divide by zero in the reference.
* Samples on input voxel centres: a voxel-rate subsample lands every
  output voxel on an input voxel, so trilinear and nearest must both return
  img[::r] exactly -- a transposed axis or a wrong stride fails it (the
  shapes and rates are anisotropic).
* Linear reproduction: trilinear interpolation is exact on a linear image;
  the legacy subsample samples between voxels (the origin moves by half a
  voxel) and must reproduce a ramp inside the volume.
* Outside: a geometry reaching beyond the input gives 0 there.
* EQSP: offset and delta from the image's min and max.
"""
import numpy as np
import pytest

import oracle


def vol(shape=(9, 12, 14), seed=0):
    return np.random.default_rng(seed).normal(100, 40, shape).astype(np.float32)


@pytest.mark.parametrize("kind,rate,expect", [
    # dim, origin, spacing for a 14 x 12 x 9 volume, spacing (1, 1.5, 2.5), origin (-3, 2, 10)
    ("voxel_rate", (4, 4, 1), ((4, 3, 9), (-3.0, 2.0, 10.0), (4.0, 6.0, 2.5))),
    ("voxel_rate", (2, 3, 2), ((7, 4, 5), (-3.0, 2.0, 10.0), (2.0, 4.5, 5.0))),
    ("voxel_legacy", (2, 2, 1), ((7, 6, 9), (-2.5, 2.75, 10.0), (2.0, 3.0, 2.5))),
    ("voxel_legacy", (4, 5, 3), ((3, 2, 3), (-3.5 + 7 / 3, 1.25 + 4.5, 8.75 + 3.75), (14 / 3, 9.0, 7.5))),
    ("mm", (2.0, 2.0, 5.0), ((8, 9, 5), (-3.0, 2.0, 10.0), (2.0, 2.0, 5.0))),
    ("pct", (0.5, 0.5, 1.0), ((8, 7, 9), (-3.0, 2.0, 10.0), (13 / 7, 16.5 / 6, 2.5))),
])
def test_geometry_closed_forms(kind, rate, expect):
    ident, dim, origin, spacing = oracle.resample_geometry(kind, (9, 12, 14), (1.0, 1.5, 2.5), (-3.0, 2.0, 10.0), rate)
    assert not ident
    assert dim == expect[0]
    np.testing.assert_allclose(origin, expect[1], rtol=1e-6)
    np.testing.assert_allclose(spacing, expect[2], rtol=1e-6)


def test_geometry_clone_and_refusals():
    assert oracle.resample_geometry("voxel_rate", (9, 12, 14), (1, 1, 1), (0, 0, 0), (1, 1, 1))[0]
    assert oracle.resample_geometry("voxel_legacy", (9, 12, 14), (1, 1, 1), (0, 0, 0), (1, 1, 1))[0]
    for kind, rate in (("voxel_rate", (0.4, 1, 1)), ("voxel_legacy", (0.9, 1, 2)), ("mm", (0, 1, 1))):
        with pytest.raises(ValueError):
            oracle.resample_geometry(kind, (9, 12, 14), (1, 1, 1), (0, 0, 0), rate)


@pytest.mark.parametrize("nearest", [False, True])
@pytest.mark.parametrize("rate", [(2, 3, 2), (3, 1, 4), (1, 2, 1)])
def test_voxel_rate_samples_on_centres(nearest, rate):
    img = vol()
    out, o, s = oracle.stage_resample(img, (1.0, 1.5, 2.5), (-3.0, 2.0, 10.0), "voxel_rate", rate, nearest)
    assert np.array_equal(out, img[::rate[2], ::rate[1], ::rate[0]])


def test_legacy_subsample_reproduces_a_ramp():
    nz, ny, nx = 10, 12, 16
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    img = (3.0 + 0.5 * x - 1.25 * y + 2.0 * z).astype(np.float32)
    sp, org = (1.0, 2.0, 0.5), (4.0, -1.0, 0.0)
    out, o, s = oracle.stage_resample(img, sp, org, "voxel_legacy", (2, 2, 2))
    # output voxel (i, j, k) sits at input index 2i + 0.5 (per axis)
    k, j, i = np.meshgrid(np.arange(out.shape[0]), np.arange(out.shape[1]), np.arange(out.shape[2]), indexing="ij")
    xi, yi, zi = 2 * i + 0.5, 2 * j + 0.5, 2 * k + 0.5
    ramp = 3.0 + 0.5 * xi - 1.25 * yi + 2.0 * zi
    np.testing.assert_allclose(out, ramp, rtol=0, atol=2e-5 * np.abs(ramp).max())


def test_outside_is_zero_and_inside_continues():
    img = vol((6, 7, 8), seed=2)
    # a grid shifted by 3 voxels in x: the last 3 output columns round outside
    out = oracle.volume_resample(img, (1, 1, 1), (0, 0, 0), (8, 7, 6), (3.0, 0.0, 0.0), (1, 1, 1))
    assert np.array_equal(out[:, :, :5], img[:, :, 3:])
    assert not np.any(out[:, :, 5:])


def test_hist_eqsp():
    img = vol(seed=5)
    off, delta = oracle.hist_eqsp(img, 32)
    mn, mx = img.min(), img.max()
    assert delta == np.float32((mx - mn) / np.float32(31))
    assert off == np.float32(np.float64(mn) - 0.5 * np.float64(delta))


def _bins(img, off, delta):
    """add_pvi_8's bin of each value, floor((v - offset) / delta) in float
    (src/register/joint_histogram.cxx:458-482)."""
    return np.floor((img.astype(np.float32) - np.float32(off)) / np.float32(delta)).astype(np.int64)


@pytest.mark.parametrize("bins", [2, 7, 20, 32, 64])
def test_hist_eqsp_half_bins_at_the_ends(bins):
    """The equal-spacing bins put the image minimum at bin coordinate 0.5
    and the maximum at bins - 0.5 -- the top and bottom bins are half full
    (src/register/joint_histogram.cxx:136-138) -- so every voxel bins into
    0 .. bins-1; a plain min/max scaling (no half-bin offset) would put the
    maximum in bin `bins`, and a delta of (max - min) / bins would put it at
    bins - 1 - 0.5 (bins - 1) / bins."""
    for seed in range(3):
        img = vol((9, 8, 7), seed=seed) * np.float32(37.0) - np.float32(600.0)
        off, delta = oracle.hist_eqsp(img, bins)
        mn, mx = np.float64(img.min()), np.float64(img.max())
        lo = (mn - off) / delta
        hi = (mx - off) / delta
        assert abs(lo - 0.5) < 1e-4 and abs(hi - (bins - 0.5)) < 1e-4 * bins
        b = _bins(img, off, delta)
        assert b.min() == 0 and b.max() == bins - 1


def test_hist_eqsp_two_values_fill_the_end_bins():
    """A two-valued image fills exactly bins 0 and bins - 1."""
    img = np.where(vol((10, 9, 8), seed=3) > 0, np.float32(1234.5), np.float32(-1000.0)).astype(np.float32)
    for bins in (2, 16, 32):
        off, delta = oracle.hist_eqsp(img, bins)
        counts = np.bincount(_bins(img, off, delta).ravel(), minlength=bins)
        assert counts.size == bins
        assert counts[0] == (img < 0).sum() and counts[-1] == (img > 0).sum()
        assert counts[1:-1].sum() == 0
