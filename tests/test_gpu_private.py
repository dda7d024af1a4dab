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


def _row_points(rng, counts, width, step, x0, crowd_rows=()):
    """2D points whose base rows hold counts[r] points each (rows listed in crowd_rows put
    their points into 3 cells), in random point order."""
    ys, xs = [], []
    for r, c in enumerate(counts):
        if c == 0:
            continue
        if r in crowd_rows:
            x = rng.integers(0, 3, c) * 7.0 + rng.uniform(0.01, 0.99, c)
        else:
            x = rng.uniform(0.0, width - 1.0, c)
        xs.append(x)
        ys.append(r + rng.uniform(0.01, 0.99, c))
    g = np.stack([np.concatenate(xs), np.concatenate(ys)], axis=1)
    g = g[rng.permutation(len(g))]
    return x0 + g * step


@pytest.mark.parametrize('dim', [1, 2])
def test_inject_staged_and_global_workgroups(orc, dim):
    """The FP64 injection stages a workgroup's points in LDS when its 64 lines lie in one 1D or
    2D field and their rows (with, in 2D, the row below the first line) hold at most 1024
    points; otherwise it merges from global memory.  Workgroups of both kinds, at the boundary
    (1024 and 1025 points) and with many points per cell, give the oracle's grids bit for bit:
    added into the caller's grids, and through the FP64 pipeline (cleared grids)."""
    from fastbarnes import interpolation as ip
    rng = np.random.default_rng(7 + dim)
    if dim == 2:
        width, height = 50, 256
        counts = np.zeros(height, dtype=np.int64)
        counts[:63] = rng.multinomial(1000, np.full(63, 1 / 63))    # lines 0-63: rows 0-63, 1024
        counts[63] = 24
        counts[64:127] = rng.multinomial(1000, np.full(63, 1 / 63))  # lines 64-127: rows 63-127, 1025
        counts[127] = 1
        counts[128:192] = rng.multinomial(200, np.full(64, 1 / 64))  # lines 128-191: 201, crowded
        counts[200:211] = 200                                       # lines 192-255: 2200
        crowd = set(range(130, 140)) | set(range(200, 205))
        step = np.array([0.5, 0.25])
        x0 = np.array([1.0, -2.0])
        size = (width, height)
        cases = [_row_points(rng, counts, width, step, x0, crowd)]
    else:
        step = np.array([0.5])
        x0 = np.array([3.0])
        size = (3000,)
        cases = []
        for n in (1024, 1025):
            g = rng.uniform(0.0, 2999.0, n)
            g[:200] = np.floor(g[:200] / 500.0) + rng.uniform(0.01, 0.99, 200)   # 6 crowded cells
            cases.append((x0 + g * step)[:, None])
    for pts in cases:
        n = len(pts)
        val = rng.normal(0.0, 3.0, n)
        shape = tuple(size[::-1])
        vg0 = rng.normal(0.0, 1.0, shape)
        wg0 = rng.uniform(0.0, 2.0, shape)
        vg, wg = vg0.copy(), wg0.copy()
        (ip._inject_data_2d if dim == 2 else ip._inject_data_1d)(vg, wg, pts, val, x0, step, size)
        rv, rw = vg0.copy(), wg0.copy()
        orc.inject_raw(rv, rw, pts, val, x0, step, size)
        assert np.array_equal(vg.view(np.uint64), rv.view(np.uint64)), n
        assert np.array_equal(wg.view(np.uint64), rw.view(np.uint64)), n
        out = ip.barnes(pts, val, 1.0, x0, step, size, method='optimized_convolution', num_iter=4,
                        max_dist=3.5, precision='fp64')
        ref = orc.interpolate_convol(pts, val, 1.0, x0, step, size, num_iter=4, max_dist=3.5)
        assert_bitwise(out, ref)
