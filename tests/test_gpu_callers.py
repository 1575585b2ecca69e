"""The in-crate callers of GaussianFilter, on the GPU path, each at the shape
it uses the filter with, bit-identical (EXACT) to the oracle's Gaussian
applied the caller's way:

- BSplineDownsampleFilter (src/filters/spatial.rs:1128-1170): sigma 1.0 over
  the whole largest region, every second pixel;
- SliceBySliceFilter (spatial.rs:1284-1340): a 2-D Gaussian per z slice (the
  oracle's 2-D filter slice by slice);
- DisplacementFieldToBSplineFilter (spatial.rs:1489-1525) and the field
  smoothing of DiffeomorphicDemonsRegistrationFilter
  (src/filters/registration_methods.rs:125-158): VecPixel<f32, 2> fields,
  each component smoothed alone (the oracle's scalar filter per component);
- N4BiasFieldCorrectionFilter (src/filters/bias_correction.rs:49-67): an f32
  2-D log image at sigma 30 (radius 90).

The reference has no tests of these adapters; the expected values are the
oracle's Gaussian (pinned against the reference's own Gaussian tests and the
dense operator) composed as each adapter composes it."""
import numpy as np
import pytest
import torch

RNG = np.random.default_rng(4242)


def _leaf(itk, x_np, index, spacing, dev, D):
    size = list(reversed(x_np.shape[:D]))
    return itk.Image(itk.Region(list(index), size), tuple(spacing), (0.5,) * D,
                     torch.from_numpy(np.ascontiguousarray(x_np)).to(dev))


@pytest.fixture(params=["auto", "per_axis"])
def path(request, itk):
    itk.set_path(itk.PATH_AUTO if request.param == "auto" else itk.PATH_PER_AXIS)
    yield request.param
    itk.set_path(itk.PATH_AUTO)


def test_caller_regions_and_spacing(itk):
    # host-side bookkeeping only: no kernel runs
    x = torch.zeros((5, 7, 9))
    img = itk.Image(itk.Region([2, -3, 4], [9, 7, 5]), (1.0, 0.5, 2.0), (1.0, 2.0, 3.0), x)
    ds = itk.BSplineDownsampleFilter(img)
    assert ds.largest_region() == itk.Region([2, -3, 4], [5, 4, 3])
    assert ds.spacing == (2.0, 1.0, 4.0) and ds.origin == (1.0, 2.0, 3.0)
    sbs = itk.SliceBySliceFilter(img, 1.5)
    assert sbs.largest_region() == img.region and sbs.spacing == img.spacing
    assert itk.DisplacementFieldToBSplineFilter(img).sigma == 2.0


@pytest.mark.gpu
@pytest.mark.parametrize("shape,index,spacing,ncomp,dtype", [
    ((21, 30, 37), (2, -3, 5), (1.0, 0.8, 1.3), 1, np.float32),
    ((20, 31, 36), (0, 0, 0), (1.0, 1.0, 1.0), 1, np.float32),
    ((45, 61), (4, 1), (0.7, 1.1), 1, np.float64),
    ((12, 17, 23), (1, 1, 1), (1.0, 1.0, 1.0), 3, np.float32),
])
def test_bspline_downsample(itk, orc, dev, path, shape, index, spacing, ncomp, dtype):
    D = len(shape)
    x = RNG.random(tuple(shape) + ((ncomp,) if ncomp > 1 else ())).astype(dtype)
    src = _leaf(itk, x, index, spacing, dev, D)
    f = itk.BSplineDownsampleFilter(src)
    out = f.generate_region(f.largest_region())
    size = list(reversed(shape))
    smooth = orc.generate_region(x, list(index), list(spacing), 1.0, list(index), size, ncomp)
    ref = smooth[(slice(None, None, 2),) * D]
    assert out.region == f.largest_region()
    assert out.spacing == tuple(2.0 * s for s in spacing)
    assert np.array_equal(out.data.cpu().numpy(), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,spacing,sigma,req,ncomp,dtype", [
    ((7, 33, 40), (0.9, 1.2, 3.0), 1.5, None, 1, np.float32),
    ((5, 48, 52), (1.0, 1.0, 1.0), 2.0, ((3, 2, 1), (40, 30, 3)), 1, np.float32),
    ((4, 29, 31), (1.3, 0.6, 1.0), 1.1, None, 1, np.float64),
    ((3, 20, 26), (1.0, 1.0, 1.0), 1.7, None, 3, np.float32),
    ((6, 64, 64), (1.0, 1.0, 1.0), 3.0, None, 1, np.float32),
])
def test_slice_by_slice(itk, orc, dev, path, shape, spacing, sigma, req, ncomp, dtype):
    x = RNG.random(tuple(shape) + ((ncomp,) if ncomp > 1 else ())).astype(dtype)
    src = _leaf(itk, x, (0, 0, 0), spacing, dev, 3)
    f = itk.SliceBySliceFilter(src, sigma)
    region = f.largest_region() if req is None else itk.Region(list(req[0]), list(req[1]))
    out = f.generate_region(region)
    assert out.region == region
    (ix, iy, iz), (sx, sy, sz) = region.index, region.size
    sub = x[iz:iz + sz, iy:iy + sy, ix:ix + sx]
    ref = np.stack([orc.generate_region(sub[z], [ix, iy], list(spacing[:2]), sigma, [ix, iy],
                                        [sx, sy], ncomp) for z in range(sz)])
    assert np.array_equal(out.data.cpu().numpy(), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,sigma,req", [
    ((41, 50), 2.0, None),
    ((64, 70), 2.0, ((5, 9), (50, 40))),
    ((33, 27), 1.0, None),
])
def test_displacement_field_smoothing(itk, orc, dev, path, shape, sigma, req):
    x = (RNG.random(tuple(shape) + (2,)) * 4.0 - 2.0).astype(np.float32)
    src = _leaf(itk, x, (0, 0), (1.0, 1.0), dev, 2)
    f = itk.DisplacementFieldToBSplineFilter(src, sigma)
    region = src.region if req is None else itk.Region(list(req[0]), list(req[1]))
    out = f.generate_region(region)
    assert out.region == region
    (ix, iy), (sx, sy) = region.index, region.size
    sub = x[iy:iy + sy, ix:ix + sx]
    # the reference splits the field into two scalar images and smooths each
    ref = np.stack([orc.generate_region(np.ascontiguousarray(sub[..., c]), [ix, iy], [1.0, 1.0],
                                        sigma, [ix, iy], [sx, sy]) for c in range(2)], axis=-1)
    assert np.array_equal(out.data.cpu().numpy(), ref)


@pytest.mark.gpu
def test_demons_field_smoothing(itk, orc, dev, path):
    # DiffeomorphicDemonsRegistrationFilter: sigma_diff 1.0 on the whole field
    x = (RNG.random((60, 80, 2)) - 0.5).astype(np.float32)
    src = _leaf(itk, x, (0, 0), (1.0, 1.0), dev, 2)
    g = itk.GaussianFilter(src, 1.0)
    out = g.generate_region(g.largest_region()).data.cpu().numpy()
    for c in range(2):
        ref = orc.generate_region(np.ascontiguousarray(x[..., c]), [0, 0], [1.0, 1.0], 1.0,
                                  [0, 0], [80, 60])
        assert np.array_equal(out[..., c], ref)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(150, 200), (96, 257)])
def test_n4_log_image_smoothing(itk, orc, dev, path, shape):
    # N4BiasFieldCorrectionFilter: an f32 log image, sigma 30 at unit spacing
    x = np.log(np.maximum(RNG.random(shape), 1e-6)).astype(np.float32)
    src = _leaf(itk, x, (0, 0), (1.0, 1.0), dev, 2)
    g = itk.GaussianFilter(src, 30.0)
    assert g._radii() == [90, 90]
    out = g.generate_region(g.largest_region()).data.cpu().numpy()
    ref = orc.generate_region(x, [0, 0], [1.0, 1.0], 30.0, [0, 0], list(reversed(shape)))
    assert np.array_equal(out, ref)
