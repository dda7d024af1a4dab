"""
GPU tests of barnes_S2 (marked `gpu`) against the oracle's restatement of
interpolationS2.py (tests/test_oracle_s2.py pins the oracle).  The device's sin, cos, tan,
pow, acos and exp may differ from glibc's in the last ulp, so the comparison is to
rounding: the Lambert-grid field and the resampled field within a few float32 ulp, the
naive_S2 field within 1e-9 relative.
"""
import os

import numpy as np
import pytest

from fastbarnes import realdata

pytestmark = pytest.mark.gpu

QFF = os.path.join(os.path.dirname(__file__), 'golden', 'qff', 'PressQFF_202007271200_3490.csv')


def _close_f32(got, ref, ulps=4):
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    d = np.abs(got[ok].astype(np.float64) - ref[ok])
    tol = ulps * np.spacing(np.abs(ref[ok]).astype(np.float32)).astype(np.float64)
    assert np.all(d <= tol), (d.max(), (d > tol).sum())


@pytest.mark.parametrize('resample', [True, False])
def test_opt_s2_matches_oracle(orc, resample):
    from fastbarnes.interpolationS2 import barnes_S2
    pts, val = realdata.read_csv_array(QFF)
    x0, step, size = np.array([-26.0 + 0.125, 34.5]), 0.125, (600, 300)
    got = barnes_S2(pts, val, 1.0, x0, step, size, method='optimized_convolution_S2', resample=resample)
    ref = orc.barnes_s2_opt(pts, val, 1.0, x0, step, size, resample=resample)
    assert got.dtype == np.float32 and got.shape == ref.shape
    _close_f32(got, ref)


def test_opt_s2_default_method_and_fp32c(orc):
    from fastbarnes.interpolationS2 import barnes_S2
    pts, val = realdata.read_csv_array(QFF)
    x0, step, size = np.array([-26.0 + 1 / 32, 34.5]), 1 / 32, (2400, 1200)
    f64 = barnes_S2(pts, val, 1.0, x0, step, size)
    f32 = barnes_S2(pts, val, 1.0, x0, step, size, precision='fp32c')
    ok = ~np.isnan(f64) & ~np.isnan(f32)
    assert (np.isnan(f64) != np.isnan(f32)).sum() <= 1e-5 * f64.size
    d = f32[ok].astype(np.float64) - f64[ok]
    assert np.sqrt(np.mean(d * d)) <= 1e-5


def test_naive_s2_matches_oracle(orc):
    from fastbarnes.interpolationS2 import barnes_S2
    pts, val = realdata.read_csv_array(QFF)
    x0, step, size = np.array([-26.0 + 0.5, 34.5]), 0.5, (150, 75)
    got = barnes_S2(pts, val, 1.0, x0, step, size, method='naive_S2')
    ref = orc.naive_s2(pts, val, 1.0, x0, step, size)
    assert got.dtype == np.float64
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=0)


def test_s2_errors():
    from fastbarnes.interpolationS2 import barnes_S2
    pts = np.zeros((3, 2))
    val = np.zeros(3)
    with pytest.raises(RuntimeError, match='invalid Barnes interpolation method'):
        barnes_S2(pts, val, 1.0, (0.0, 40.0), 0.5, (20, 20), method='naive')
    with pytest.raises(RuntimeError, match='invalid length'):
        barnes_S2(pts, val, [1.0, 1.0, 1.0], (0.0, 40.0), 0.5, (20, 20))
