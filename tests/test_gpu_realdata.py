"""
GPU tests on the real station data (tests/golden/qff, 3490 QFF stations) with the paper's
geometry (fastbarnes.realdata): FP64 bit parity with the oracle at 2400x1200, the FP32C
error, and the RMSE-versus-iterations curve against the oracle's.
"""
import os

import numpy as np
import pytest

from fastbarnes import realdata

pytestmark = pytest.mark.gpu

QFF = os.path.join(os.path.dirname(__file__), 'golden', 'qff', 'PressQFF_202007271200_3490.csv')


def _bitwise(a, b):
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(a)
    np.testing.assert_array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32))


@pytest.mark.parametrize('method', ['optimized_convolution', 'convolution'])
def test_real_fp64_bitwise_paper_grid(orc, method):
    from fastbarnes.interpolation import barnes
    pts, val = realdata.read_csv_array(QFF)
    x0, step, size = realdata.paper_grid(32.0)
    got = barnes(pts, val, 1.0, x0, step, size, method=method, num_iter=4)
    ref = orc.interpolate_convol(pts, val, 1.0, x0, step, size, num_iter=4, tail=method == 'optimized_convolution')
    _bitwise(got, ref)


def test_real_fp32c_error(orc):
    """QFF values are about 1000 hPa, where one float32 ulp is 6.1e-5: the FP32C result
    rounds differently from the FP64 one on a small fraction of nodes.  Bound: RMSE 1e-5
    and max-abs 2 ulp (1.3e-4), no mask mismatch."""
    from fastbarnes.interpolation import barnes
    pts, val = realdata.read_csv_array(QFF)
    x0, step, size = realdata.paper_grid(32.0)
    f32 = barnes(pts, val, 1.0, x0, step, size, precision='fp32c')
    f64 = barnes(pts, val, 1.0, x0, step, size, precision='fp64')
    np.testing.assert_array_equal(np.isnan(f32), np.isnan(f64))
    ok = ~np.isnan(f64)
    d = f32[ok].astype(np.float64) - f64[ok]
    rmse, mx = float(np.sqrt(np.mean(d * d))), float(np.abs(d).max())
    print('real data fp32c vs fp64: rmse %.3e max %.3e' % (rmse, mx))
    assert rmse <= 1e-5 and mx <= 1.3e-4


def test_real_rmse_curve_matches_oracle(orc):
    """The paper's figure-8 curve at resolution 4 (300x150), both methods, n in 1..50:
    GPU (FP64 path, one batch per method, naive on the GPU) against the oracle."""
    pts, val = realdata.read_csv_array(QFF)
    x0, step, size = realdata.paper_grid(4.0)
    ns = [1, 2, 3, 4, 5, 6, 8, 10, 20, 50]
    curve, naive = realdata.rmse_curve(pts, val, resolution=4.0, num_iters=ns, precision='fp64')
    onaive = orc.interpolate_naive(pts, val, 1.0, x0, step, size)
    np.testing.assert_allclose(naive, onaive, rtol=1e-12, atol=1e-9)
    for method, tail in (('convolution', False), ('optimized_convolution', True)):
        ref = [realdata.rmse_west_europe(onaive, orc.interpolate_convol(pts, val, 1.0, x0, step, size, num_iter=k,
                                                                        tail=tail), x0, step) for k in ns]
        np.testing.assert_allclose(curve[method], ref, rtol=1e-9)
    # the optimized convolution converges monotonically, the plain one does not (paper, figure 8)
    oc = curve['optimized_convolution']
    assert all(a > b for a, b in zip(oc, oc[1:]))


def test_real_rmse_curve_fp32c_routes_wrapped_counts_to_fp64():
    """With FP32C, the curve's plain-convolution counts whose integer threshold wraps (K**n >=
    2**63: n >= 23 at resolution 8) run in an FP64 batch and equal the FP64 curve there exactly;
    the other counts stay within FP32C rounding of the FP64 curve."""
    pts, val = realdata.read_csv_array(QFF)
    ns = [1, 4, 10, 22, 23, 50]
    wrapped = realdata.fp64_counts('convolution', 8.0, num_iters=ns)
    assert wrapped == [23, 50] and realdata.fp64_counts('optimized_convolution', 8.0, num_iters=ns) == []
    c32, naive = realdata.rmse_curve(pts, val, resolution=8.0, num_iters=ns, precision='fp32c')
    c64, _ = realdata.rmse_curve(pts, val, resolution=8.0, num_iters=ns, precision='fp64', naive=naive)
    for method in ('convolution', 'optimized_convolution'):
        for k, a, b in zip(ns, c32[method], c64[method]):
            if method == 'convolution' and k in wrapped:
                assert a == b
            else:
                assert abs(a - b) <= 1e-5, (method, k, a, b)


@pytest.mark.parametrize('method,num_iter', [('optimized_convolution', 50), ('optimized_convolution', 30),
                                             ('convolution', 12)])
def test_real_fp32c_many_iterations(method, num_iter):
    """Regression: FP32C stores the fields in float32 between iterations; with the
    normalisation applied only after the last iteration, (K+2 alpha)^(n-1) overflowed float32
    for many iterations (K=15, n=50 at 2400x1200: 15^49 ~ 1e57) and the result became NaN.
    The normalisation is now applied in groups; FP32C must match the FP64 path."""
    from fastbarnes.interpolation import barnes
    pts, val = realdata.read_csv_array(QFF)
    x0, step, size = realdata.paper_grid(32.0)
    f32 = barnes(pts, val, 1.0, x0, step, size, method=method, num_iter=num_iter, precision='fp32c')
    f64 = barnes(pts, val, 1.0, x0, step, size, method=method, num_iter=num_iter, precision='fp64')
    assert np.isnan(f64).mean() < 0.2
    np.testing.assert_array_equal(np.isnan(f32), np.isnan(f64))
    ok = ~np.isnan(f64)
    d = f32[ok].astype(np.float64) - f64[ok]
    assert float(np.sqrt(np.mean(d * d))) <= 1e-5 and float(np.abs(d).max()) <= 1.3e-4


def test_real_fp32c_rejects_wrapped_convolution_threshold():
    """'convolution' computes its threshold as the integer power K**n (interpolation.py:669-670),
    which wraps around in 64 bits (K=17, n=50 at 2400x1200).  The FP64 path reproduces the
    wrapped threshold, which leaves nodes with weights ~1e-100 of the peak unmasked; float32
    cannot carry them, so FP32C refuses instead of returning different values there."""
    from fastbarnes.interpolation import barnes
    pts, val = realdata.read_csv_array(QFF)
    x0, step, size = realdata.paper_grid(32.0)
    with pytest.raises(RuntimeError, match='FP32C does not support method convolution'):
        barnes(pts, val, 1.0, x0, step, size, method='convolution', num_iter=50, precision='fp32c')
    f64 = barnes(pts, val, 1.0, x0, step, size, method='convolution', num_iter=50, precision='fp64')
    assert np.isfinite(f64).mean() > 0.99
