"""
Drop-in ``fastbarnes.interpolationS2``: Barnes interpolation on the unit sphere S^2,
backed by HIP kernels (MeteoSwiss fast-barnes-py 2.0.0, fastbarnes/interpolationS2.py).

'optimized_convolution_S2'  the sample points are mapped by the Lambert conformal conic
    projection create_proj(11.5, 34.5, 42.5, 65.5) onto the fixed Lambert grid
    x0 = (-32, -2), size (int(64/step), int(44/step)); the optimized convolution runs
    there; the Lambert field is resampled bilinearly to the requested lon/lat grid
    (resample=True) or returned as is (resample=False).  float32 result.
'naive_S2'  O(N*W*H) Barnes with the spherical (great-circle) distance in degrees.
    float64 result.

Everything runs on the GPU through libfastbarnes.so (fb_barnes_s2).  The device's
trigonometric functions may differ from the host libm's in the last ulp, so results
agree with a CPU evaluation to rounding rather than bit for bit.
"""
import ctypes
from math import exp

import numpy as np

from fastbarnes import _lib
from fastbarnes._lib import check, lib

__all__ = ['barnes_S2', 'get_lambert_proj']

_S2_METHODS = {
    'optimized_convolution_S2': _lib.FB_METHOD_OPTIMIZED_CONVOLUTION,
    'naive_S2': _lib.FB_METHOD_NAIVE,
    # the reference's signature defaults to 'optimized_convolution', which its dispatch then
    # rejects (interpolationS2.py:31, :146-154); here it selects the optimized S2 method
    'optimized_convolution': _lib.FB_METHOD_OPTIMIZED_CONVOLUTION,
}


def get_lambert_proj():
    """The Lambert projection of barnes_S2, (center_lon, n, n_inv, F, rho0) as returned by
    lambert_conformal.create_proj(11.5, 34.5, 42.5, 65.5) (interpolationS2.py:214-216)."""
    from math import cos, log, pi, tan
    rad = pi / 180.0
    lat1, lat2, clat = 42.5, 65.5, 34.5
    n = log(cos(lat1 * rad) / cos(lat2 * rad)) / log(tan((90.0 + lat2) * rad / 2) / tan((90.0 + lat1) * rad / 2))
    F = cos(lat1 * rad) * tan((90.0 + lat1) * rad / 2) ** n / n
    return (11.5, n, 1.0 / n, F, F / tan((90.0 + clat) * rad / 2) ** n)


def _vec(name, x, dim):
    if isinstance(x, (list, tuple, np.ndarray)):
        if len(x) != dim:
            raise RuntimeError('specified ' + name + ' with invalid length: ' + str(len(x)))
        return np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return np.full(dim, x, dtype=np.float64)


def barnes_S2(pts, val, sigma, x0, step, size, method='optimized_convolution', num_iter=4, max_dist=3.5,
              resample=True, *, precision='fp64'):
    """
    Barnes interpolation on S^2 of the observations `val` at the sample points `pts`
    ((lon, lat) in degrees, N x 2) onto the lon/lat grid (x0, step, size), on the GPU.

    Arguments and results are those of the reference's barnes_S2 (interpolationS2.py:31-154):
    float32 for 'optimized_convolution_S2' (shape size[::-1], or the Lambert grid's shape
    when resample=False), float64 for 'naive_S2'.  precision ('fp64' or 'fp32c',
    keyword-only) selects the arithmetic of the convolution on the Lambert grid.
    """
    dim = pts.shape[1]
    sigma = _vec('sigma', sigma, dim)
    x0 = _vec('x0', x0, dim)
    step = _vec('step', step, dim)
    if isinstance(size, (list, tuple, np.ndarray)):
        if len(size) != dim:
            raise RuntimeError('specified size with invalid length: ' + str(len(size)))
        size = tuple(size)
    elif dim != 1:
        raise RuntimeError('array size should be array-like of length: ' + str(dim))
    else:
        size = (size, )
    max_dist_weight = exp(-max_dist**2/2)
    if method not in _S2_METHODS:
        raise RuntimeError('encountered invalid Barnes interpolation method: ' + str(method))
    if dim != 2:
        raise RuntimeError('Barnes interpolation on S^2 needs (lon, lat) sample points')
    mcode = _S2_METHODS[method]
    prec = _lib.PRECISIONS.get(precision)
    if prec is None:
        raise RuntimeError('unknown precision: ' + str(precision) + " (expected 'fp64' or 'fp32c')")
    p = np.ascontiguousarray(pts, dtype=np.float64)
    v = np.ascontiguousarray(val, dtype=np.float64)
    if p.shape[0] != v.shape[0] or p.shape[0] < 1:
        raise RuntimeError('pts and val arrays have inconsistent shapes: ' + str(p.shape) + ' vs. ' + str(v.shape))
    if np.isnan(p).any():
        raise RuntimeError('sample point coordinates contain NaN')
    sz = np.asarray(size, dtype=np.int64)
    if mcode == _lib.FB_METHOD_NAIVE:
        out = np.empty((int(sz[1]), int(sz[0])), dtype=np.float64)
    elif resample:
        out = np.empty((int(sz[1]), int(sz[0])), dtype=np.float32)
    else:
        out = np.empty((int(44.0 / step[1]), int(64.0 / step[0])), dtype=np.float32)
    dp = ctypes.POINTER(ctypes.c_double)
    check(lib().fb_barnes_s2(mcode, prec, p.shape[0], p.ctypes.data, v.ctypes.data, sigma.ctypes.data_as(dp),
                             x0.ctypes.data_as(dp), step.ctypes.data_as(dp),
                             sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(num_iter),
                             float(max_dist_weight), int(bool(resample)), out.ctypes.data, 0, None))
    return out
