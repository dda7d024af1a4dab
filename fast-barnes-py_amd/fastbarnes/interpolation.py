"""
Drop-in ``fastbarnes.interpolation`` backed by HIP kernels on AMD Instinct MI355X.

``barnes()`` keeps the reference's signature, argument handling, RuntimeError
messages and return arrays (MeteoSwiss fast-barnes-py 2.0.0,
fastbarnes/interpolation.py:31-199).  The interpolation itself always runs on
the GPU through libfastbarnes.so; there is no CPU code path.

Methods
-------
'optimized_convolution'  tailed n-fold box filter (interpolation.py:328-367)
'convolution'            plain n-fold box filter   (interpolation.py:575-612)
'naive'                  exact Gaussian Barnes, FP64 (interpolation.py:861-938)
'radius'                 Gaussian Barnes within the min_weight radius, 2D (interpolation.py:809-855)

Keyword-only ``precision`` selects the arithmetic of the convolution methods:
'fp64' (default) follows the reference's own operation order and returns the
reference's float32 result bit for bit; 'fp32c' is the fast path (FP32 storage,
FP64 window sums and quotient), within RMSE 1e-5 of 'fp64'.

The module also exposes GPU-backed equivalents of the reference's private
helpers that its unit tests call (``_accumulate_tail_array``,
``_convolve_tail_1d`` ...), with the same signatures.
"""
import ctypes
from math import exp

import numpy as np

from fastbarnes import _lib
from fastbarnes._lib import check, lib

__all__ = ['barnes', 'get_half_kernel_size_opt', 'get_tail_value', 'get_half_kernel_size', 'get_sigma_effective']

_METHODS = {
    'optimized_convolution': _lib.FB_METHOD_OPTIMIZED_CONVOLUTION,
    'convolution': _lib.FB_METHOD_CONVOLUTION,
    'naive': _lib.FB_METHOD_NAIVE,
    'radius': _lib.FB_METHOD_RADIUS,
}


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _vec(x, dim):
    """sigma/x0/step normalisation of interpolation.py:128-153 (length already checked)."""
    if isinstance(x, (list, tuple, np.ndarray)):
        return np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return np.full(dim, x, dtype=np.float64)


def _precision_code(precision):
    try:
        return _lib.PRECISIONS[precision]
    except KeyError:
        raise RuntimeError('unknown precision: ' + str(precision) + " (expected 'fp64' or 'fp32c')")


###############################################################################

def barnes(pts, val, sigma, x0, step, size, method='optimized_convolution',
           num_iter=4, max_dist=3.5, min_weight=0.001, *, precision='fp64'):
    """
    Barnes interpolation of the observations `val` at the sample points `pts`
    onto the regular grid (x0, step, size), computed on the GPU.

    Parameters and return value are those of the reference's
    ``fastbarnes.interpolation.barnes``: `size` is given in (x, y, z) order and
    the result is indexed [y, x] / [z, y, x]; it is float32 for the convolution
    methods (NaN beyond `max_dist`) and float64 for 'naive' and 'radius'.

    precision : {'fp64', 'fp32c'}, keyword-only
        Arithmetic of the convolution methods (see the module docstring).

    Raises RuntimeError on invalid arguments, with the reference's messages
    (interpolation.py:105-199).  Beyond the reference, NaN sample coordinates,
    ``num_iter < 1``, non-positive `step` or `sigma` and an empty `pts` raise
    RuntimeError too (the reference's behaviour there is undefined).
    """
    # pts: (N, M) or (N,)  -- interpolation.py:104-115
    if not isinstance(pts, np.ndarray):
        raise RuntimeError('specified pts is not a numpy ndarray')
    if len(pts.shape) > 2:
        raise RuntimeError('expected pts array of shape (N, M) but was: ' + str(pts.shape))
    if len(pts.shape) == 1:
        pts = np.reshape(pts, (-1, 1))
    dim = pts.shape[1]
    if dim < 1 or dim > 3:
        raise RuntimeError('Barnes interpolation supports only sample points in dimensions 1, 2 or 3')

    # val  -- interpolation.py:118-123 (the caller's arrays are never modified)
    if not isinstance(val, np.ndarray):
        raise RuntimeError('specified val is not a numpy ndarray')
    if len(val.shape) > 1:
        raise RuntimeError('expected val array of shape (N) but was: ' + str(val.shape))
    if val.shape[0] != pts.shape[0]:
        raise RuntimeError('pts and val arrays have inconsistent shapes: ' + str(pts.shape) + ' vs. ' + str(val.shape))

    # sigma, x0, step  -- interpolation.py:128-153
    for name, arg in (('sigma', sigma), ('x0', x0), ('step', step)):
        if isinstance(arg, (list, tuple, np.ndarray)) and len(arg) != dim:
            raise RuntimeError('specified ' + name + ' with invalid length: ' + str(len(arg)))
    sigma = _vec(sigma, dim)
    x0 = _vec(x0, dim)
    step = _vec(step, dim)

    # size  -- interpolation.py:156-164
    if isinstance(size, (list, tuple, np.ndarray)):
        if len(size) != dim:
            raise RuntimeError('specified size with invalid length: ' + str(len(size)))
        size = tuple(size)
    elif dim != 1:
        raise RuntimeError('array size should be array-like of length: ' + str(dim))
    else:
        size = (size, )

    max_dist_weight = exp(-max_dist**2/2)   # interpolation.py:167

    if method not in _METHODS:
        raise RuntimeError('encountered invalid Barnes interpolation method: ' + method)
    mcode = _METHODS[method]

    if mcode in (_lib.FB_METHOD_OPTIMIZED_CONVOLUTION, _lib.FB_METHOD_CONVOLUTION):
        # kernel must be smaller than the grid  -- interpolation.py:170-175, :179-184
        if mcode == _lib.FB_METHOD_OPTIMIZED_CONVOLUTION:
            kernel_size = 2*_get_half_kernel_size_opt(sigma, step, num_iter) + 1
        else:
            kernel_size = 2*_get_half_kernel_size(sigma, step, num_iter) + 1
        for m in range(dim):
            if kernel_size[m] >= size[m]:
                raise RuntimeError('resulting rectangular kernel size should be smaller w.r.t. specified grid: '
                                   + str(kernel_size) + ' vs. ' + str(size))
    elif mcode == _lib.FB_METHOD_RADIUS:
        # interpolation.py:189-192
        if dim != 2:
            raise RuntimeError('radius algorithm works only in 2D but data is: ' + str(dim) + 'D')
        if sigma[0] != sigma[1]:
            raise RuntimeError('radius algorithm in 2D works only for scalar sigma value but sigma is: ' + str(sigma))

    return _run(mcode, dim, pts, val, sigma, x0, step, size, num_iter, max_dist_weight, min_weight,
                _precision_code(precision))


def _run(mcode, dim, pts, val, sigma, x0, step, size, num_iter, max_dist_weight, min_weight, prec):
    """Hands the validated arguments to fb_barnes (host buffers in, new host array out)."""
    n = pts.shape[0]
    if n < 1:
        raise RuntimeError('at least one sample point is required')
    if not np.all(np.isfinite(step)) or np.any(step <= 0) or not np.all(np.isfinite(sigma)) or np.any(sigma <= 0):
        raise RuntimeError('sigma and step must be finite and positive')
    if int(num_iter) < 1 and mcode in (_lib.FB_METHOD_OPTIMIZED_CONVOLUTION, _lib.FB_METHOD_CONVOLUTION):
        raise RuntimeError('num_iter must be >= 1')
    p = np.ascontiguousarray(pts, dtype=np.float64)
    if np.isnan(p).any():
        raise RuntimeError('sample point coordinates contain NaN')
    v = np.ascontiguousarray(val, dtype=np.float64)
    sz = np.asarray(size, dtype=np.int64)
    if np.any(sz < 1):
        raise RuntimeError('grid size must be positive: ' + str(size))
    shape = tuple(int(s) for s in sz[::-1])
    out = np.empty(shape, dtype=np.float32 if mcode <= _lib.FB_METHOD_CONVOLUTION else np.float64)
    check(lib().fb_barnes(mcode, prec, dim, n, p.ctypes.data, v.ctypes.data, _dptr(sigma), _dptr(x0), _dptr(step),
                          sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), int(num_iter), float(max_dist_weight),
                          float(min_weight), out.ctypes.data, 0, None))
    return out


# -----------------------------------------------------------------------------
# kernel parameters (interpolation.py:538-569, :777-803)

def _to_np(value):
    """ Converts a scalar to a numpy array. """
    return np.asarray([value], dtype=np.float64)


def _params_args(sigma, step):
    sigma = np.ascontiguousarray(np.asarray(sigma, dtype=np.float64).reshape(-1))
    step = np.ascontiguousarray(np.asarray(step, dtype=np.float64).reshape(-1))
    if len(step) == 1 and len(sigma) > 1:
        step = np.full(len(sigma), step[0])
    if len(sigma) == 1 and len(step) > 1:
        sigma = np.full(len(step), sigma[0])
    return sigma, step


def _get_half_kernel_size_opt(sigma, step, num_iter):
    """ Half kernel size T of the optimized convolution, per axis (int32 array). """
    sigma, step = _params_args(sigma, step)
    out = np.empty(len(sigma), dtype=np.int32)
    check(lib().fb_half_kernel_size(_lib.FB_METHOD_OPTIMIZED_CONVOLUTION, len(sigma), _dptr(sigma), _dptr(step),
                                    int(num_iter), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return out


def _get_half_kernel_size(sigma, step, num_iter):
    """ Half kernel size T of the plain convolution, per axis (int32 array). """
    sigma, step = _params_args(sigma, step)
    out = np.empty(len(sigma), dtype=np.int32)
    check(lib().fb_half_kernel_size(_lib.FB_METHOD_CONVOLUTION, len(sigma), _dptr(sigma), _dptr(step),
                                    int(num_iter), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
    return out


def _get_tail_value(sigma, step, num_iter):
    """ Tail value alpha of the optimized convolution, per axis. """
    sigma, step = _params_args(sigma, step)
    out = np.empty(len(sigma), dtype=np.float64)
    check(lib().fb_tail_value(len(sigma), _dptr(sigma), _dptr(step), int(num_iter), _dptr(out)))
    return out


def _get_sigma_effective(sigma, step, num_iter):
    """ Effective sigma of the num_iter-fold plain rect kernel, per axis. """
    sigma, step = _params_args(sigma, step)
    out = np.empty(len(sigma), dtype=np.float64)
    check(lib().fb_sigma_effective(len(sigma), _dptr(sigma), _dptr(step), int(num_iter), _dptr(out)))
    return out


def get_half_kernel_size_opt(sigma, step, num_iter):
    """ Computes the half kernel size T for the optimized convolution algorithm. Version for scalar arguments. """
    return _get_half_kernel_size_opt(_to_np(sigma), _to_np(step), num_iter)[0]


def get_tail_value(sigma, step, num_iter):
    """ Computes the tail value alpha for the optimized convolution algorithm. Version for scalar arguments. """
    return _get_tail_value(_to_np(sigma), _to_np(step), num_iter)[0]


def get_half_kernel_size(sigma, step, num_iter):
    """ Computes the half kernel size T for the convolution algorithm. Version for scalar arguments. """
    return _get_half_kernel_size(_to_np(sigma), _to_np(step), num_iter)[0]


def get_sigma_effective(sigma, step, num_iter):
    """ Effective sigma of the num_iter-fold convolved rect-kernel of length 2*T+1. Scalar version. """
    return _get_sigma_effective(_to_np(sigma), _to_np(step), num_iter)[0]


# -----------------------------------------------------------------------------
# GPU-backed equivalents of the reference's private line/grid helpers

_precision_private = 'fp64'


def _accumulate_lines(in_arr, arr_len, rect_len, num_iter, alpha, tail):
    a = np.ascontiguousarray(np.asarray(in_arr, dtype=np.float64)[:arr_len])
    out = np.empty_like(a)
    check(lib().fb_accumulate_lines(int(tail), _precision_code(_precision_private), 1, int(arr_len), int(rect_len),
                                    int(num_iter), float(alpha), a.ctypes.data, out.ctypes.data, 0, None))
    return out


def _accumulate_tail_array(in_arr, h_arr, arr_len, rect_len, num_iter, alpha):
    """num_iter-fold convolution of a line with [alpha, 1 x rect_len, alpha] (interpolation.py:484-533).
    Returns a new array; h_arr is not needed by the GPU version and left untouched."""
    return _accumulate_lines(in_arr, arr_len, rect_len, num_iter, alpha, True)


def _accumulate_array(in_arr, h_arr, arr_len, rect_len, num_iter):
    """num_iter-fold convolution of a line with a rect kernel of length rect_len (interpolation.py:729-772)."""
    return _accumulate_lines(in_arr, arr_len, rect_len, num_iter, 0.0, False)


def _convolve_grid(tail, vg, wg, sigma, step, size, kernel_size, num_iter, tail_value, max_dist_weight):
    dim = len(size)
    sigma, step = _params_args(sigma, step)
    ks = np.ascontiguousarray(np.asarray(kernel_size, dtype=np.int32).reshape(-1))
    tv = np.ascontiguousarray(np.asarray(tail_value if tail else np.zeros(dim), dtype=np.float64).reshape(-1))
    thr = ctypes.c_double(0.0)
    method = _lib.FB_METHOD_OPTIMIZED_CONVOLUTION if tail else _lib.FB_METHOD_CONVOLUTION
    check(lib().fb_threshold(method, dim, _dptr(sigma), _dptr(step), ks.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                             int(num_iter), _dptr(tv), float(max_dist_weight), ctypes.byref(thr)))
    for a in (vg, wg):
        if a.dtype != np.float64 or not a.flags.c_contiguous:
            raise RuntimeError('grid arrays must be C-contiguous float64')
    sz = np.asarray(size, dtype=np.int64)
    check(lib().fb_convolve_grid(int(tail), _precision_code(_precision_private), dim,
                                 sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 ks.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), int(num_iter), _dptr(tv),
                                 thr.value, vg.ctypes.data, wg.ctypes.data, 0, None))


def _convolve_tail_1d(vg, wg, sigma, step, size, kernel_size, num_iter, tail_value, max_dist_weight):
    """In-place tailed convolution + NaN mask of 1D fields (interpolation.py:372-394)."""
    _convolve_grid(True, vg, wg, sigma, step, size, kernel_size, num_iter, tail_value, max_dist_weight)


def _convolve_tail_2d(vg, wg, sigma, step, size, kernel_size, num_iter, tail_value, max_dist_weight):
    """In-place tailed convolution + NaN mask of 2D fields (interpolation.py:397-430)."""
    _convolve_grid(True, vg, wg, sigma, step, size, kernel_size, num_iter, tail_value, max_dist_weight)


def _convolve_tail_3d(vg, wg, sigma, step, size, kernel_size, num_iter, tail_value, max_dist_weight):
    """In-place tailed convolution + NaN mask of 3D fields (interpolation.py:433-479)."""
    _convolve_grid(True, vg, wg, sigma, step, size, kernel_size, num_iter, tail_value, max_dist_weight)


def _convolve_1d(vg, wg, sigma, step, size, kernel_size, num_iter, max_dist_weight):
    """In-place plain convolution + NaN mask of 1D fields (interpolation.py:617-639)."""
    _convolve_grid(False, vg, wg, sigma, step, size, kernel_size, num_iter, None, max_dist_weight)


def _convolve_2d(vg, wg, sigma, step, size, kernel_size, num_iter, max_dist_weight):
    """In-place plain convolution + NaN mask of 2D fields (interpolation.py:642-675)."""
    _convolve_grid(False, vg, wg, sigma, step, size, kernel_size, num_iter, None, max_dist_weight)


def _convolve_3d(vg, wg, sigma, step, size, kernel_size, num_iter, max_dist_weight):
    """In-place plain convolution + NaN mask of 3D fields (interpolation.py:678-724)."""
    _convolve_grid(False, vg, wg, sigma, step, size, kernel_size, num_iter, None, max_dist_weight)


def _normalize_values(val):
    """Offsets the values, in place, so that their range is centred on 0; returns the offset
    (interpolation.py:204-212; np.amin/np.amax propagate NaN)."""
    offset = (np.amin(val) + np.amax(val)) / 2.0
    val -= offset
    return offset


def _inject_data(dim, vg, wg, pts, val, x0, step, size):
    for a in (vg, wg):
        if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
            raise RuntimeError('grid arrays must be C-contiguous float64')
        if a.size != int(np.prod(np.asarray(size, dtype=np.int64))):
            raise RuntimeError('grid arrays do not match size')
    p = np.ascontiguousarray(np.asarray(pts, dtype=np.float64).reshape(len(pts), -1)[:, :dim])
    v = np.ascontiguousarray(val, dtype=np.float64)
    x0v = np.ascontiguousarray(np.asarray(x0, dtype=np.float64).reshape(-1)[:dim])
    st = np.ascontiguousarray(np.asarray(step, dtype=np.float64).reshape(-1)[:dim])
    sz = np.ascontiguousarray(np.asarray(size, dtype=np.int64).reshape(-1)[:dim])
    check(lib().fb_inject(dim, len(v), p.ctypes.data, v.ctypes.data, _dptr(x0v), _dptr(st),
                          sz.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), vg.ctypes.data, wg.ctypes.data, 0,
                          None))


def _inject_data_1d(vg, wg, pts, val, x0, step, size):
    """Adds the linear corner weights of the samples to vg (times val) and wg, in place
    (interpolation.py:218-238), on the GPU; the sums equal the reference's bit for bit."""
    _inject_data(1, vg, wg, pts, val, x0, step, size)


def _inject_data_2d(vg, wg, pts, val, x0, step, size):
    """Adds the bilinear corner weights of the samples to vg (times val) and wg, in place
    (interpolation.py:241-272)."""
    _inject_data(2, vg, wg, pts, val, x0, step, size)


def _inject_data_3d(vg, wg, pts, val, x0, step, size):
    """Adds the trilinear corner weights of the samples to vg (times val) and wg, in place
    (interpolation.py:275-322)."""
    _inject_data(3, vg, wg, pts, val, x0, step, size)


def _interpolate_conv_private(mcode, pts, val, sigma, x0, step, size, num_iter, max_dist_weight):
    dim = len(size)
    orig = np.array(val, dtype=np.float64, copy=True)
    out = _run(mcode, dim, np.asarray(pts, dtype=np.float64).reshape(len(orig), -1), orig, _vec(sigma, dim),
               _vec(x0, dim), _vec(step, dim), tuple(size), num_iter, max_dist_weight, 0.0,
               _precision_code(_precision_private))
    # like the reference, the caller's val is left normalised (the GPU formed the same offset
    # from the unmodified copy)
    _normalize_values(val)
    return out


def _interpolate_opt_convol(pts, val, sigma, x0, step, size, num_iter, max_dist_weight):
    """Algorithm B with the tailed rectangular kernel (interpolation.py:328-367): the float32
    field of shape size[::-1].  As in the reference, val is normalised in place."""
    return _interpolate_conv_private(_lib.FB_METHOD_OPTIMIZED_CONVOLUTION, pts, val, sigma, x0, step, size,
                                     num_iter, max_dist_weight)


def _interpolate_convol(pts, val, sigma, x0, step, size, num_iter, max_dist_weight):
    """Algorithm B with the plain rectangular kernel (interpolation.py:575-612); val is
    normalised in place."""
    return _interpolate_conv_private(_lib.FB_METHOD_CONVOLUTION, pts, val, sigma, x0, step, size, num_iter,
                                     max_dist_weight)
