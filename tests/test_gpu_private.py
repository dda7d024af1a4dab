"""
GPU tests of the reference's other private entry points that its callers use
(fastbarnes/interpolationS2.py:194 calls _interpolate_opt_convol; demo/quantinterpolation.py:138-145
calls _normalize_values, _inject_data_2d and _convolve_2d), through the C ABI, against the oracle:

* _inject_data_1d/2d/3d add into the caller's existing grids in the reference's per-cell point
  order: bitwise equal to the oracle's restatement, with points outside the grid, on its last
  line, and many points per cell;
* _interpolate_opt_convol / _interpolate_convol return the oracle's float32 field bit for bit
  and leave the caller's val normalised as the reference does;
* the quantised-interpolation sequence of demo/quantinterpolation.py (normalise, inject,
  plain 2D convolution, quotient) gives the oracle's FP64 fields bit for bit.
"""
import numpy as np
import pytest

from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu


def _points(rng, dim, n, size, step, x0, crowd=False):
    ext = (np.asarray(size) - 1) * step
    pts = x0 + rng.uniform(-0.05, 1.05, (n, dim)) * ext   # some outside
    pts[: n // 20, 0] = x0[0] + ext[0]                     # on the last line: skipped
    if crowd:   # 300 points in 5 cells
        cells = rng.integers(0, 5, 300)
        pts[-300:] = x0 + (cells[:, None] * 1.0 + rng.uniform(0, 1, (300, dim))) * step
    return pts


@pytest.mark.parametrize('dim', [1, 2, 3])
@pytest.mark.parametrize('crowd', [False, True])
def test_inject_data_matches_oracle(orc, dim, crowd):
    from fastbarnes import interpolation as ip
    rng = np.random.default_rng(100 * dim + crowd)
    size = {1: (500,), 2: (120, 80), 3: (40, 30, 20)}[dim]
    step = np.array([0.1, 0.2, 0.25][:dim])
    x0 = np.array([-3.0, 2.0, 0.5][:dim])
    n = 4000
    pts = _points(rng, dim, n, size, step, x0, crowd)
    val = rng.normal(0.0, 3.0, n)
    shape = tuple(size[::-1])
    vg0 = rng.normal(0.0, 1.0, shape)          # existing contents are kept
    wg0 = rng.uniform(0.0, 2.0, shape)
    vg, wg = vg0.copy(), wg0.copy()
    {1: ip._inject_data_1d, 2: ip._inject_data_2d, 3: ip._inject_data_3d}[dim](vg, wg, pts, val, x0, step, size)
    rv, rw = vg0.copy(), wg0.copy()
    orc.inject_raw(rv, rw, pts, val, x0, step, size)
    assert np.array_equal(vg.view(np.uint64), rv.view(np.uint64))
    assert np.array_equal(wg.view(np.uint64), rw.view(np.uint64))
    assert (wg != wg0).sum() > 100


@pytest.mark.parametrize('tail', [True, False])
def test_interpolate_private_matches_oracle(orc, tail):
    from fastbarnes import interpolation as ip
    from fastbarnes import workloads
    c = workloads.config1(seed=61)
    mdw = float(np.exp(-c['max_dist'] ** 2 / 2))
    sigma = np.full(2, c['sigma'])
    step = np.full(2, c['step'])
    val = c['val'].copy()
    fn = ip._interpolate_opt_convol if tail else ip._interpolate_convol
    out = fn(c['pts'], val, sigma, c['x0'], step, c['size'], c['num_iter'], mdw)
    ref = orc.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'],
                                 num_iter=c['num_iter'], max_dist=c['max_dist'], tail=tail)
    assert_bitwise(out, ref)
    v = c['val'].copy()
    off = (np.amin(v) + np.amax(v)) / 2.0
    assert np.array_equal(val, v - off)


def test_quantised_interpolation_sequence_matches_oracle(orc):
    """normalise, inject, plain 2D convolution, quotient + offset (demo/quantinterpolation.py:138-150)."""
    from fastbarnes import interpolation as ip
    from fastbarnes import workloads
    c = workloads.config1(seed=62)
    sigma = np.full(2, c['sigma'])
    step = np.full(2, c['step'])
    mdw = float(np.exp(-c['max_dist'] ** 2 / 2))
    val = c['val'].copy()
    offset = ip._normalize_values(val)
    vg = np.zeros(c['size'][::-1])
    wg = np.zeros(c['size'][::-1])
    ip._inject_data_2d(vg, wg, c['pts'], val, c['x0'], step, c['size'])
    T = ip._get_half_kernel_size(sigma, step, c['num_iter'])
    ip._convolve_2d(vg, wg, sigma, step, c['size'], 2 * T + 1, c['num_iter'], mdw)
    got = vg / wg + offset
    ref, rvg, rwg = orc.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'],
                                           num_iter=c['num_iter'], max_dist=c['max_dist'], tail=False,
                                           return_fields=True)
    np.testing.assert_array_equal(np.isnan(wg), np.isnan(rwg))
    ok = ~np.isnan(rwg)
    assert np.array_equal(vg[ok].view(np.uint64), rvg[ok].view(np.uint64))
    assert np.array_equal(got.astype(np.float32)[ok].view(np.uint32), ref[ok].view(np.uint32))
