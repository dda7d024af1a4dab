"""
GPU test of fb_barnes_noalloc, the drop-in entry point that allocates no device memory (its
workspace is the library's fixed per-device scratch array), through the C ABI: its results equal
fb_barnes's on the same device inputs bit for bit, in both precisions and both convolution
methods, on two calls in a row (the second with a smaller field, so stale contents of the
scratch array from the first call must not leak into it), and a call whose workspace exceeds
the array is refused before any kernel runs.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_d3 = ctypes.c_double * 3
_i3 = ctypes.c_int64 * 3


def _args(c):
    dim = len(c['size'])
    sig = np.broadcast_to(np.asarray(c['sigma'], dtype=np.float64), (dim,))
    step = np.broadcast_to(np.asarray(c['step'], dtype=np.float64), (dim,))
    return (dim, _d3(*sig), _d3(*np.asarray(c['x0'], dtype=np.float64)), _d3(*step), _i3(*c['size']),
            float(np.exp(-c['max_dist'] ** 2 / 2)))


def _call(entry, c, method, precision, stream=None):
    """One call of fb_barnes (device pointers) or fb_barnes_noalloc on the field c."""
    from fastbarnes import _lib, hip
    L = _lib.lib()
    dim, sig, x0, step, size, mdw = _args(c)
    d_pts = hip.DeviceArray.from_numpy(np.ascontiguousarray(c['pts'], dtype=np.float64))
    d_val = hip.DeviceArray.from_numpy(np.ascontiguousarray(c['val'], dtype=np.float64))
    shape = tuple(c['size'][::-1])
    d_out = hip.DeviceArray(shape, np.float32)
    d_out.upload(np.full(shape, 3.0, dtype=np.float32))
    s = stream.ptr if stream is not None else None
    n = len(c['val'])
    if entry == 'noalloc':
        rc = L.fb_barnes_noalloc(method, precision, dim, n, d_pts.ptr, d_val.ptr, sig, x0, step, size,
                                 c['num_iter'], mdw, d_out.ptr, s)
    else:
        rc = L.fb_barnes(method, precision, dim, n, d_pts.ptr, d_val.ptr, sig, x0, step, size, c['num_iter'], mdw,
                         0.0, d_out.ptr, _lib.FB_FLAG_DEVICE_PTRS, s)
    _lib.check(rc)
    return d_out.numpy()


def _bitwise(a, b):
    assert a.shape == b.shape and a.dtype == b.dtype
    na, nb = np.isnan(a), np.isnan(b)
    np.testing.assert_array_equal(na, nb)
    assert np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
@pytest.mark.parametrize('method', ['optimized_convolution', 'convolution'])
def test_noalloc_matches_fb_barnes(precision, method):
    from fastbarnes import _lib, hip, workloads
    prec = _lib.PRECISIONS[precision]
    meth = _lib.FB_METHOD_OPTIMIZED_CONVOLUTION if method == 'optimized_convolution' else _lib.FB_METHOD_CONVOLUTION
    big = workloads.config2(seed=31)     # 2400x1200, 3490 stations
    small = workloads.config1(seed=32)   # 300x150: the second call is smaller
    stream = hip.Stream()
    outs = [_call('noalloc', big, meth, prec, stream), _call('noalloc', small, meth, prec)]
    print('noalloc %s %s: two calls done' % (method, precision), flush=True)
    _bitwise(outs[0], _call('fb_barnes', big, meth, prec, stream))
    _bitwise(outs[1], _call('fb_barnes', small, meth, prec))
    assert np.isfinite(outs[1]).mean() > 0.5
    # and once more the large one after the small one
    _bitwise(_call('noalloc', big, meth, prec, stream), outs[0])


def test_noalloc_rejects_oversized_workspace():
    """3D 512x512x128 in FP32C needs 2 x 128 MiB of v/w planes, beyond the fixed array: refused
    with FB_ERR_UNSUPPORTED while the plan is laid out, before any kernel runs."""
    from fastbarnes import _lib, workloads
    c = workloads.config4(n=1000)
    with pytest.raises(RuntimeError, match='exceeds the fixed scratch array'):
        _call('noalloc', c, _lib.FB_METHOD_OPTIMIZED_CONVOLUTION, _lib.FB_PREC_FP32C)


def test_noalloc_rejects_other_methods():
    from fastbarnes import _lib, workloads
    c = workloads.config1(seed=33)
    with pytest.raises(RuntimeError, match='convolution methods only'):
        _call('noalloc', c, _lib.FB_METHOD_NAIVE, _lib.FB_PREC_FP64)
