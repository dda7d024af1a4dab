"""
CPU oracle for fast Barnes interpolation -- TEST INFRASTRUCTURE ONLY.

ctypes front end of ``oracle/liboracle.so`` (built from ``barnes_oracle.c`` by
``oracle/Makefile``), an FP64 restatement of the reference's
``fastbarnes/interpolation.py`` that keeps its loop and operation order.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module, and only to check or time against it; the product package
never does.

Every wrapper takes and returns numpy arrays with the reference's conventions
(``size`` in (x, y, z) order, grids indexed [y, x] / [z, y, x]).
"""

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, 'liboracle.so')
_OMP_PATH = os.path.join(_HERE, 'liboracle_omp.so')
_lib = None
_omp = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compiles liboracle.so with the committed Makefile."""
    import subprocess
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


def lib():
    """The checker build (liboracle.so)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(_LIB_PATH))
    return _lib


def omp_lib():
    """The OpenMP build of the same source (liboracle_omp.so), timed as the CPU baseline;
    orc_set_threads selects its thread count."""
    global _omp
    if _omp is None:
        if not os.path.exists(_OMP_PATH):
            build()
        _omp = _bind(ctypes.CDLL(_OMP_PATH))
    return _omp


def _bind(L):
    """Argument and result types of the exported functions."""
    L.orc_half_kernel_size_opt.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, _i32p]
    L.orc_half_kernel_size.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, _i32p]
    L.orc_tail_value.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, _dp]
    L.orc_sigma_effective.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, _dp]
    L.orc_threshold.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _i32p, ctypes.c_int64, _dp,
                                ctypes.c_double]
    L.orc_threshold.restype = ctypes.c_double
    L.orc_normalize_values.argtypes = [_dp, ctypes.c_int64]
    L.orc_normalize_values.restype = ctypes.c_double
    L.orc_inject.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int64, _dp, _dp, _i64p]
    L.orc_accumulate_tail_array.argtypes = [_dp, _dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_double]
    L.orc_accumulate_tail_array.restype = _dp
    L.orc_accumulate_array.argtypes = [_dp, _dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
    L.orc_accumulate_array.restype = _dp
    L.orc_convolve.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _i64p, _i32p,
                               ctypes.c_int64, _dp, ctypes.c_double]
    L.orc_interpolate_convol.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int64, _dp, _dp,
                                         _dp, _i64p, ctypes.c_int64, ctypes.c_double, _fp, _dp, _dp]
    L.orc_interpolate_naive.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int64, _dp, _dp, _dp, _i64p, _dp]
    L.orc_interpolate_radius_bf.argtypes = [_dp, _dp, ctypes.c_int64, ctypes.c_double, _dp, _dp, _i64p,
                                            ctypes.c_double, ctypes.c_double, _dp]
    L.orc_filter_grid.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _i64p, ctypes.c_int64]
    L.orc_adjoint.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int64, _dp, _dp, _dp, _i64p,
                              ctypes.c_int64, ctypes.c_double, _dp, _dp, _dp]
    L.orc_lambert_create.argtypes = [ctypes.c_double] * 4 + [_dp]
    L.orc_lambert_scale.argtypes = [ctypes.c_double, _dp]
    L.orc_lambert_scale.restype = ctypes.c_double
    L.orc_lambert_to_map.argtypes = [_dp, _dp, ctypes.c_int64, _dp]
    L.orc_lambert_to_geo.argtypes = [_dp, _dp, ctypes.c_int64, _dp]
    L.orc_resample.argtypes = [_fp, _dp, _i64p, _dp, _dp, _i64p, _dp, _fp]
    L.orc_barnes_s2_opt.argtypes = [_dp, _dp, ctypes.c_int64, _dp, _dp, _dp, _i64p, ctypes.c_int64,
                                    ctypes.c_double, ctypes.c_int, _fp]
    L.orc_naive_s2.argtypes = [_dp, _dp, ctypes.c_int64, _dp, _dp, _dp, _i64p, _dp]
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_set_threads.restype = ctypes.c_int
    return L


def _d(a):
    return a.ctypes.data_as(_dp)


def _vec(x, dim, dtype=np.float64):
    if isinstance(x, (list, tuple, np.ndarray)):
        return np.ascontiguousarray(np.asarray(x, dtype=dtype).reshape(dim))
    return np.full(dim, x, dtype=dtype)


def _pts(pts):
    pts = np.asarray(pts, dtype=np.float64)
    if pts.ndim == 1:
        pts = pts.reshape(-1, 1)
    return np.ascontiguousarray(pts)


# ---- scalar / parameter helpers --------------------------------------------

def half_kernel_size_opt(sigma, step, num_iter, dim=1):
    out = np.empty(dim, dtype=np.int32)
    lib().orc_half_kernel_size_opt(dim, _d(_vec(sigma, dim)), _d(_vec(step, dim)), num_iter,
                                   out.ctypes.data_as(_i32p))
    return out


def half_kernel_size(sigma, step, num_iter, dim=1):
    out = np.empty(dim, dtype=np.int32)
    lib().orc_half_kernel_size(dim, _d(_vec(sigma, dim)), _d(_vec(step, dim)), num_iter,
                               out.ctypes.data_as(_i32p))
    return out


def tail_value(sigma, step, num_iter, dim=1):
    out = np.empty(dim, dtype=np.float64)
    lib().orc_tail_value(dim, _d(_vec(sigma, dim)), _d(_vec(step, dim)), num_iter, _d(out))
    return out


def sigma_effective(sigma, step, num_iter, dim=1):
    out = np.empty(dim, dtype=np.float64)
    lib().orc_sigma_effective(dim, _d(_vec(sigma, dim)), _d(_vec(step, dim)), num_iter, _d(out))
    return out


def threshold(tail, sigma, step, kernel_size, num_iter, tail_value_arr, max_dist_weight, dim):
    ks = _vec(kernel_size, dim, np.int32)
    tv = _vec(tail_value_arr if tail_value_arr is not None else 0.0, dim)
    return lib().orc_threshold(int(tail), dim, _d(_vec(sigma, dim)), _d(_vec(step, dim)),
                               ks.ctypes.data_as(_i32p), num_iter, _d(tv), max_dist_weight)


# ---- line passes ------------------------------------------------------------

def accumulate_tail_array(in_arr, rect_len, num_iter, alpha):
    """Returns the num_iter-fold tailed box filter of a 1-D line (new array)."""
    a = np.array(in_arr, dtype=np.float64, copy=True)
    b = np.empty_like(a)
    r = lib().orc_accumulate_tail_array(_d(a), _d(b), len(a), rect_len, num_iter, alpha)
    return a.copy() if ctypes.addressof(r.contents) == a.ctypes.data else b.copy()


def accumulate_array(in_arr, rect_len, num_iter):
    a = np.array(in_arr, dtype=np.float64, copy=True)
    b = np.empty_like(a)
    r = lib().orc_accumulate_array(_d(a), _d(b), len(a), rect_len, num_iter)
    return a.copy() if ctypes.addressof(r.contents) == a.ctypes.data else b.copy()


def convolve(tail, vg, wg, sigma, step, size, kernel_size, num_iter, tail_value_arr, max_dist_weight):
    """In-place convolution + NaN mask of vg, wg (float64, C order, shape size[::-1])."""
    dim = len(size)
    assert vg.dtype == np.float64 and vg.flags.c_contiguous and wg.flags.c_contiguous
    sz = np.asarray(size, dtype=np.int64)
    ks = _vec(kernel_size, dim, np.int32)
    tv = _vec(tail_value_arr if tail_value_arr is not None else 0.0, dim)
    lib().orc_convolve(int(tail), dim, _d(vg), _d(wg), _d(_vec(sigma, dim)), _d(_vec(step, dim)),
                       sz.ctypes.data_as(_i64p), ks.ctypes.data_as(_i32p), num_iter, _d(tv), max_dist_weight)


# ---- whole interpolations ---------------------------------------------------

def _common(pts, val, sigma, x0, step, size):
    pts = _pts(pts)
    dim = pts.shape[1]
    val = np.ascontiguousarray(np.asarray(val, dtype=np.float64))
    if not isinstance(size, (list, tuple, np.ndarray)):
        size = (size,)
    sz = np.asarray(size, dtype=np.int64)
    return pts, dim, val, _vec(sigma, dim), _vec(x0, dim), _vec(step, dim), sz


def interpolate_convol(pts, val, sigma, x0, step, size, num_iter=4, max_dist=3.5, tail=True,
                       return_fields=False, threads=None):
    """FP64 restatement of _interpolate_opt_convol (tail=True) / _interpolate_convol (tail=False).
    threads: run the OpenMP build (liboracle_omp.so) on that many threads instead of the
    checker build; the same bits either way."""
    L = lib()
    if threads is not None:
        L = omp_lib()
        L.orc_set_threads(int(threads))
    pts, dim, val, sg, x0v, st, sz = _common(pts, val, sigma, x0, step, size)
    mdw = math.exp(-max_dist ** 2 / 2)
    shape = tuple(int(s) for s in sz[::-1])
    out = np.empty(shape, dtype=np.float32)
    vg = np.empty(shape, dtype=np.float64) if return_fields else None
    wg = np.empty(shape, dtype=np.float64) if return_fields else None
    L.orc_interpolate_convol(int(tail), dim, _d(pts), _d(val), len(val), _d(sg), _d(x0v), _d(st),
                             sz.ctypes.data_as(_i64p), num_iter, mdw, out.ctypes.data_as(_fp),
                             _d(vg) if return_fields else None, _d(wg) if return_fields else None)
    if return_fields:
        return out, vg, wg
    return out


def interpolate_naive(pts, val, sigma, x0, step, size):
    pts, dim, val, sg, x0v, st, sz = _common(pts, val, sigma, x0, step, size)
    out = np.empty(tuple(int(s) for s in sz[::-1]), dtype=np.float64)
    lib().orc_interpolate_naive(dim, _d(pts), _d(val), len(val), _d(sg), _d(x0v), _d(st),
                                sz.ctypes.data_as(_i64p), _d(out))
    return out


def interpolate_radius(pts, val, sigma, x0, step, size, max_dist=3.5, min_weight=0.001):
    pts, dim, val, sg, x0v, st, sz = _common(pts, val, sigma, x0, step, size)
    assert dim == 2
    mdw = math.exp(-max_dist ** 2 / 2)
    out = np.empty((int(sz[1]), int(sz[0])), dtype=np.float64)
    lib().orc_interpolate_radius_bf(_d(pts), _d(val), len(val), float(sg[0]), _d(x0v), _d(st),
                                    sz.ctypes.data_as(_i64p), mdw, min_weight, _d(out))
    return out


def inject_raw(vg, wg, pts, val, x0, step, size):
    """_inject_data_*d as the reference calls it: adds into the existing vg, wg (float64,
    shape size[::-1]) with val used as given."""
    pts = _pts(pts)
    dim = pts.shape[1]
    v = np.ascontiguousarray(val, dtype=np.float64)
    sz = np.asarray(size, dtype=np.int64).reshape(-1)
    assert vg.dtype == np.float64 and wg.dtype == np.float64 and vg.flags.c_contiguous and wg.flags.c_contiguous
    lib().orc_inject(dim, _d(vg), _d(wg), _d(pts), _d(v), len(v), _d(_vec(x0, dim)), _d(_vec(step, dim)),
                     sz.ctypes.data_as(_i64p))


def inject(pts, val, x0, step, size):
    """Returns (vg, wg) after _normalize_values + _inject_data_*d, and the offset."""
    pts = _pts(pts)
    dim = pts.shape[1]
    v = np.array(val, dtype=np.float64, copy=True)
    off = lib().orc_normalize_values(_d(v), len(v))
    if not isinstance(size, (list, tuple, np.ndarray)):
        size = (size,)
    sz = np.asarray(size, dtype=np.int64)
    shape = tuple(int(s) for s in sz[::-1])
    vg = np.zeros(shape)
    wg = np.zeros(shape)
    lib().orc_inject(dim, _d(vg), _d(wg), _d(pts), _d(v), len(v), _d(_vec(x0, dim)), _d(_vec(step, dim)),
                     sz.ctypes.data_as(_i64p))
    return vg, wg, off


def filter_grid(grid, sigma, step, num_iter=4, tail=True):
    """B x: the unscaled per-axis line filters of the convolution applied to a grid
    (shape size[::-1]); the operator the adjoint transposes."""
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.float64))
    dim = g.ndim
    sz = np.asarray(g.shape[::-1], dtype=np.int64)
    out = np.empty_like(g)
    lib().orc_filter_grid(int(tail), dim, _d(g), _d(out), _d(_vec(sigma, dim)), _d(_vec(step, dim)),
                          sz.ctypes.data_as(_i64p), num_iter)
    return out


def adjoint(pts, val, sigma, x0, step, size, g, num_iter=4, max_dist=3.5, tail=True):
    """FP64 adjoint of interpolate_convol: for g = dL/df (grid, shape size[::-1]) returns
    (dL/dval (n,), dL/dpts (n, dim)).  No reference counterpart; derived from the forward."""
    pts, dim, val, sg, x0v, st, sz = _common(pts, val, sigma, x0, step, size)
    mdw = math.exp(-max_dist ** 2 / 2)
    gg = np.ascontiguousarray(np.asarray(g, dtype=np.float64))
    assert gg.shape == tuple(int(s) for s in sz[::-1])
    gval = np.empty(len(val))
    gpts = np.empty((len(val), dim))
    lib().orc_adjoint(int(tail), dim, _d(pts), _d(val), len(val), _d(sg), _d(x0v), _d(st),
                      sz.ctypes.data_as(_i64p), num_iter, mdw, _d(gg), _d(gval), _d(gpts))
    return gval, gpts


# ---- spherical variant (interpolationS2.py, util/lambert_conformal.py) ----------

def lambert_proj(center_lon=11.5, center_lat=34.5, lat1=42.5, lat2=65.5):
    p = np.empty(5)
    lib().orc_lambert_create(center_lon, center_lat, lat1, lat2, _d(p))
    return p


def lambert_scale(lat, proj):
    return lib().orc_lambert_scale(float(lat), _d(np.ascontiguousarray(proj, dtype=np.float64)))


def lambert_to_map(geo, proj):
    g = np.ascontiguousarray(np.asarray(geo, dtype=np.float64).reshape(-1, 2))
    m = np.empty_like(g)
    lib().orc_lambert_to_map(_d(g), _d(m), len(g), _d(np.ascontiguousarray(proj, dtype=np.float64)))
    return m


def lambert_to_geo(mp, proj):
    m = np.ascontiguousarray(np.asarray(mp, dtype=np.float64).reshape(-1, 2))
    g = np.empty_like(m)
    lib().orc_lambert_to_geo(_d(m), _d(g), len(m), _d(np.ascontiguousarray(proj, dtype=np.float64)))
    return g


def lambert_grid(step):
    """(lam_x0, lam_size) of the fixed Lambert grid of barnes_S2 (interpolationS2.py:186-188)."""
    st = _vec(step, 2)
    return np.array([-32.0, -2.0]), (int(64.0 / st[0]), int(44.0 / st[1]))


def barnes_s2_opt(pts, val, sigma, x0, step, size, num_iter=4, max_dist=3.5, resample=True):
    pts, dim, val, sg, x0v, st, sz = _common(pts, val, sigma, x0, step, size)
    assert dim == 2
    mdw = math.exp(-max_dist ** 2 / 2)
    if resample:
        out = np.empty((int(sz[1]), int(sz[0])), dtype=np.float32)
    else:
        lam_size = lambert_grid(st)[1]
        out = np.empty((lam_size[1], lam_size[0]), dtype=np.float32)
    lib().orc_barnes_s2_opt(_d(pts), _d(val), len(val), _d(sg), _d(x0v), _d(st), sz.ctypes.data_as(_i64p),
                            num_iter, mdw, int(bool(resample)), out.ctypes.data_as(_fp))
    return out


def resample(lam_field, step, x0, size, proj=None):
    lam = np.ascontiguousarray(lam_field, dtype=np.float32)
    st, x0v = _vec(step, 2), _vec(x0, 2)
    lam_x0 = np.array([-32.0, -2.0])
    lsz = np.asarray(lam.shape[::-1], dtype=np.int64)
    sz = np.asarray(size, dtype=np.int64)
    out = np.empty((int(sz[1]), int(sz[0])), dtype=np.float32)
    p = lambert_proj() if proj is None else np.ascontiguousarray(proj, dtype=np.float64)
    lib().orc_resample(lam.ctypes.data_as(_fp), _d(lam_x0), lsz.ctypes.data_as(_i64p), _d(x0v), _d(st),
                       sz.ctypes.data_as(_i64p), _d(p), out.ctypes.data_as(_fp))
    return out


def naive_s2(pts, val, sigma, x0, step, size):
    pts, dim, val, sg, x0v, st, sz = _common(pts, val, sigma, x0, step, size)
    out = np.empty((int(sz[1]), int(sz[0])), dtype=np.float64)
    lib().orc_naive_s2(_d(pts), _d(val), len(val), _d(sg), _d(x0v), _d(st), sz.ctypes.data_as(_i64p), _d(out))
    return out
