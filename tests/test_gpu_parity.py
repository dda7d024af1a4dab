"""
GPU parity of the HIP path against the CPU oracle (marked `gpu`; runs through the C ABI).

Bar:
* FP64 path (precision='fp64'): float32 output bit-identical to the oracle, which restates
  the reference's FP64 operation order; NaN masks identical.
* FP32C path (precision='fp32c'): identical NaN mask except at nodes whose oracle weight
  lies within 1e-5 (relative) of the threshold, RMSE <= 1e-5 and max-abs <= 2e-4 against
  the oracle's FP64 result on the non-NaN nodes (tolerance of the fast path).
* the reference's own unit tests (AccumulationTest, GaussianApproximationTest, BasicTest)
  restated against the GPU-backed helpers.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RMSE_TOL = 1e-5
MAXABS_TOL = 2e-4


def _run(c, method='optimized_convolution', precision='fp64', **kw):
    from fastbarnes.interpolation import barnes
    return barnes(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], method=method,
                  num_iter=c['num_iter'], max_dist=c['max_dist'], precision=precision, **kw)


def _oracle(orc, c, tail=True):
    return orc.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'],
                                  num_iter=c['num_iter'], max_dist=c['max_dist'], tail=tail, return_fields=True)


def assert_bitwise(a, b):
    assert a.shape == b.shape and a.dtype == b.dtype
    na, nb = np.isnan(a), np.isnan(b)
    np.testing.assert_array_equal(na, nb)
    diff = (a.view(np.uint32) != b.view(np.uint32)) & ~na
    assert not diff.any(), 'first mismatches at %s' % (np.argwhere(diff)[:5],)


def fast_errors(out, ref, wg, thr):
    """(rmse, maxabs, mask mismatches away from the threshold) of a FP32C result."""
    na, nb = np.isnan(out), np.isnan(ref)
    mism = na != nb
    near = np.zeros_like(mism)
    w = np.nan_to_num(wg, nan=0.0)
    near = np.abs(w - thr) <= 1e-5 * thr
    ok = ~na & ~nb
    d = out[ok].astype(np.float64) - ref[ok].astype(np.float64)
    return (float(np.sqrt(np.mean(d * d))) if d.size else 0.0, float(np.max(np.abs(d))) if d.size else 0.0,
            int((mism & ~near).sum()), int(mism.sum()))


def _thr(orc, c, tail=True):
    dim = len(c['size']) if isinstance(c['size'], (tuple, list)) else 1
    sig = np.full(dim, c['sigma']) if np.isscalar(c['sigma']) else np.asarray(c['sigma'])
    st = np.full(dim, c['step']) if np.isscalar(c['step']) else np.asarray(c['step'])
    n = c['num_iter']
    T = (orc.half_kernel_size_opt if tail else orc.half_kernel_size)(sig, st, n, dim=dim)
    a = orc.tail_value(sig, st, n, dim=dim) if tail else np.zeros(dim)
    return orc.threshold(tail, sig, st, 2 * T + 1, n, a, math.exp(-c['max_dist'] ** 2 / 2), dim)


# ---------------------------------------------------------------------------
# the reference's own unit tests, restated against the GPU helpers

@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
def test_accumulation_known_answers(known_answers, precision):
    from fastbarnes import interpolation
    interpolation._precision_private = precision
    try:
        for case in known_answers['accumulation']:
            inp = np.zeros(case['size'])
            for i, v in case['impulses']:
                inp[i] = v
            exp = np.zeros(case['size'])
            for s, t, v in case['expect']:
                exp[s:t] = v
            h = np.empty(case['size'])
            if case['fn'] == 'tail':
                out = interpolation._accumulate_tail_array(inp, h, case['size'], case['rect_len'], case['num_iter'],
                                                           case['alpha'])
            else:
                out = interpolation._accumulate_array(inp, h, case['size'], case['rect_len'], case['num_iter'])
            np.testing.assert_array_equal(out, exp, err_msg=case['src'])
        v = known_answers['accumulation_versions']
        inp = np.zeros(v['size'])
        for i, x in v['impulses']:
            inp[i] = x
        for p in v['pairs']:
            a = interpolation._accumulate_array(inp, None, v['size'], p['plain_rect_len'], v['num_iter'])
            b = interpolation._accumulate_tail_array(inp, None, v['size'], p['tail_rect_len'], v['num_iter'],
                                                     p['alpha'])
            np.testing.assert_array_equal(a, b, err_msg=v['src'])
    finally:
        interpolation._precision_private = 'fp64'


def test_line_pass_matches_oracle(orc):
    """GPU _accumulate_tail_array == oracle bit for bit (fp64) on random lines, all phases."""
    from fastbarnes import interpolation
    rng = np.random.default_rng(3)
    for T, alpha, n, L in [(3, 0.0, 4, 300), (27, 0.2083333333333333, 4, 2400), (9, 0.890625, 4, 128),
                           (5, 0.5, 1, 40), (12, 0.99, 6, 150), (1, 0.3, 3, 4), (63, 0.7, 6, 1200)]:
        line = rng.normal(size=L) * np.exp(rng.normal(size=L))
        ref = orc.accumulate_tail_array(line, 2 * T + 1, n, alpha)
        out = interpolation._accumulate_tail_array(line, None, L, 2 * T + 1, n, alpha)
        np.testing.assert_array_equal(out.view(np.uint64), ref.view(np.uint64))
        ref0 = orc.accumulate_array(line, 2 * T + 1, n)
        out0 = interpolation._accumulate_array(line, None, L, 2 * T + 1, n)
        np.testing.assert_array_equal(out0.view(np.uint64), ref0.view(np.uint64))


def test_fast_line_pass_close_to_oracle(orc):
    from fastbarnes import interpolation
    rng = np.random.default_rng(4)
    interpolation._precision_private = 'fp32c'
    try:
        for T, alpha, n, L in [(27, 0.2083333333333333, 4, 2400), (9, 0.890625, 4, 128), (63, 0.7, 6, 1200),
                               (2, 0.1, 3, 4000)]:
            line = np.abs(rng.normal(size=L)) * (rng.uniform(size=L) < 0.05)
            ref = orc.accumulate_tail_array(line, 2 * T + 1, n, alpha)
            out = interpolation._accumulate_tail_array(line, None, L, 2 * T + 1, n, alpha)
            # FP32 storage between iterations: relative 2^-24 per stored value and iteration
            np.testing.assert_allclose(out, ref, rtol=2e-6 * n, atol=1e-6 * np.max(ref))
    finally:
        interpolation._precision_private = 'fp64'


@pytest.mark.parametrize('opt', [False, True])
def test_gaussian_approximation(known_answers, opt):
    """tests/GaussianApproximationTest.py:32-100 on the GPU (the values field only)."""
    from fastbarnes import interpolation
    g = known_answers['gaussian']
    for sigma in g['sigmas']:
        for num_iter, max_diff in (g['opt'] if opt else g['plain']):
            delta = 1.0 / g['resol']
            grid = np.arange(-g['width_sigmas'] * sigma, g['width_sigmas'] * sigma + 0.000001, delta)
            np_sigma = interpolation._to_np(sigma)
            np_delta = interpolation._to_np(delta)
            if opt:
                s_ref = sigma
                ks = 2 * interpolation._get_half_kernel_size_opt(np_sigma, np_delta, num_iter) + 1
                tv = interpolation._get_tail_value(np_sigma, np_delta, num_iter)
            else:
                s_ref = interpolation.get_sigma_effective(sigma, delta, num_iter)
                ks = 2 * interpolation._get_half_kernel_size(np_sigma, np_delta, num_iter) + 1
            gauss = 1.0 / math.sqrt(2.0 * math.pi) / s_ref * np.exp(-(grid / s_ref) ** 2 / 2)
            values = np.zeros(len(grid))
            values[len(grid) // 2] = 1.0 / delta
            weights = values.copy()
            if opt:
                interpolation._convolve_tail_1d(values, weights, np_sigma, np_delta, (len(grid),), ks, num_iter, tv,
                                                9999.9)
            else:
                interpolation._convolve_1d(values, weights, np_sigma, np_delta, (len(grid),), ks, num_iter, 9999.9)
            values *= (delta / 2 / math.sqrt(3 / num_iter) / s_ref) ** num_iter
            assert np.max(np.abs(gauss - values)) <= max_diff / s_ref, (sigma, num_iter)


@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
def test_basic_ranges(known_answers, precision):
    """tests/BasicTest.py:42-101: results within the input range [0, 1] or NaN."""
    from fastbarnes.interpolation import barnes
    b = known_answers['basic']
    d1, d2, d3 = b['d1'], b['d2'], b['d3']
    # the fast path's FP32 storage widens the reference's eps to 1e-6 (relative to range 1)
    eps = lambda e: e if precision == 'fp64' else 1e-6
    data = np.array(d2['data'])
    rng = np.random.default_rng(307)
    pts3 = 2.0 * rng.random((d3['num_values'], 3))
    val3 = rng.integers(2, size=d3['num_values']).astype(np.float64)
    for method in ('convolution', 'optimized_convolution'):
        for sigma in d1['sigmas']:
            for n in d1['num_iters']:
                r = barnes(np.array(d1['pts']), np.array(d1['val']), sigma, 0.0, d1['step'], d1['size'],
                           method=method, num_iter=n, precision=precision)
                assert np.all(((r >= -eps(d1['eps'])) & (r <= 1 + eps(d1['eps']))) | np.isnan(r))
                r = barnes(data[:, :2], data[:, 2], sigma, 0.0, d2['step'], tuple(d2['size']), method=method,
                           num_iter=n, precision=precision)
                assert np.all(((r >= -eps(d2['eps'])) & (r <= 1 + eps(d2['eps']))) | np.isnan(r))
            for n in d3['num_iters']:
                r = barnes(pts3, val3, sigma, 0.0, d3['step'], tuple(d3['size']), method=method, num_iter=n,
                           precision=precision)
                assert np.all(((r >= -eps(d3['eps'])) & (r <= 1 + eps(d3['eps']))) | np.isnan(r))


def test_kernel_size_boundary_runs(known_answers):
    """tests/BasicTest.py:117-119,137-139: a grid of kernel_size+1 is accepted."""
    from fastbarnes import interpolation
    kc = known_answers['basic']['kernel_check']
    pts, val = np.array(kc['pts']), np.array(kc['val'])
    for method, fn in (('convolution', interpolation.get_half_kernel_size),
                       ('optimized_convolution', interpolation.get_half_kernel_size_opt)):
        for n in kc['num_iters']:
            ks = 2 * fn(kc['sigma'], kc['step'], n) + 1
            r = interpolation.barnes(pts, val, kc['sigma'], 0.0, kc['step'], ks + 1, method=method, num_iter=n)
            assert r.shape == (ks + 1,) and r.dtype == np.float32


# ---------------------------------------------------------------------------
# end-to-end parity with the oracle

def _configs():
    from fastbarnes import workloads
    return {'cfg1': workloads.config1(), 'cfg2': workloads.config2()}


@pytest.mark.parametrize('name', ['cfg1', 'cfg2'])
@pytest.mark.parametrize('method', ['optimized_convolution', 'convolution'])
def test_fp64_bitwise(orc, name, method):
    c = _configs()[name]
    out = _run(c, method=method, precision='fp64')
    ref, _, _ = _oracle(orc, c, tail=method == 'optimized_convolution')
    assert_bitwise(out, ref)


@pytest.mark.parametrize('name', ['cfg1', 'cfg2'])
def test_fp32c_accuracy(orc, name, capsys):
    c = _configs()[name]
    out = _run(c, precision='fp32c')
    ref, vg, wg = _oracle(orc, c)
    rmse, maxabs, mism_far, mism = fast_errors(out, ref, wg, _thr(orc, c))
    print('%s fp32c: rmse %.3g maxabs %.3g mask mismatches %d (%d near threshold)' % (name, rmse, maxabs, mism,
                                                                                     mism - mism_far))
    assert mism_far == 0
    assert rmse <= RMSE_TOL
    assert maxabs <= MAXABS_TOL


def _small_cases():
    from fastbarnes import workloads
    rng = np.random.default_rng(11)
    cases = []
    # 1D
    p = rng.uniform(-1, 11, 400)
    cases.append(dict(pts=p, val=np.sin(p) + rng.normal(0, 0.1, 400), sigma=0.3, x0=0.0, step=0.02, size=500,
                      num_iter=4, max_dist=3.5))
    # 2D anisotropic sigma/step, uneven grid
    p = workloads.stations_2d(2000, 5)
    cases.append(dict(pts=p, val=workloads.values_sinusoids(p, 5), sigma=[0.8, 1.3], x0=np.array([-26.0, 34.0]),
                      step=[0.11, 0.07], size=(683, 545), num_iter=5, max_dist=3.0))
    # 3D
    c4 = workloads.config4(n=20000, size=(96, 80, 40))
    cases.append(c4)
    # 2D, large sigma (T > segment length), n = 6
    p = workloads.stations_2d(1500, 6)
    cases.append(dict(pts=p, val=workloads.values_cfg(p, 6), sigma=2.0, x0=np.array([-26.0, 34.0]), step=0.0625,
                      size=(1200, 600), num_iter=6, max_dist=3.5))
    return cases


@pytest.mark.parametrize('idx', range(4))
def test_small_cases(orc, idx):
    c = _small_cases()[idx]
    out64 = _run(c, precision='fp64')
    ref, vg, wg = _oracle(orc, c)
    assert_bitwise(out64, ref)
    out32 = _run(c, precision='fp32c')
    rmse, maxabs, mism_far, mism = fast_errors(out32, ref, wg, _thr(orc, c))
    assert mism_far == 0 and rmse <= RMSE_TOL and maxabs <= MAXABS_TOL, (rmse, maxabs, mism)


def test_edge_cases(orc):
    from fastbarnes.interpolation import barnes
    x0, step, size = np.array([0.0, 0.0]), 0.1, (60, 40)
    # points exactly on the last grid line are dropped, at x0 kept
    pts = np.array([[0.0, 0.0], [5.9, 1.0], [2.0, 3.9], [3.3, 2.2]])
    val = np.array([1.0, 5.0, 7.0, 3.0])
    for prec in ('fp64', 'fp32c'):
        out = barnes(pts, val, 0.3, x0, step, size, precision=prec)
        ref = orc.interpolate_convol(pts, val, 0.3, x0, step, size)
        if prec == 'fp64':
            assert_bitwise(out, ref)
        else:
            np.testing.assert_array_equal(np.isnan(out), np.isnan(ref))
            ok = ~np.isnan(ref)
            np.testing.assert_allclose(out[ok], ref[ok], atol=2e-5)
    # all points outside: everything NaN
    out = barnes(np.array([[-1.0, -1.0], [100.0, 1.0]]), np.array([1.0, 2.0]), 0.3, x0, step, size)
    assert np.isnan(out).all()
    # one point, constant output inside the radius
    out = barnes(np.array([[3.02, 1.97]]), np.array([4.25]), 0.3, x0, step, size, precision='fp32c')
    ok = ~np.isnan(out)
    assert ok.sum() > 0 and np.all(out[ok] == np.float32(4.25))
    # NaN in val: the reference's offset becomes NaN, so every node is NaN
    out = barnes(pts, np.array([1.0, np.nan, 7.0, 3.0]), 0.3, x0, step, size)
    assert np.isnan(out).all()


def test_determinism():
    c = _configs()['cfg2']
    for prec in ('fp64', 'fp32c'):
        a = _run(c, precision=prec)
        b = _run(c, precision=prec)
        assert_bitwise(a, b)


# ---------------------------------------------------------------------------
# naive and radius

def test_naive_matches_oracle(orc):
    from fastbarnes.interpolation import barnes
    c = _configs()['cfg1']
    sub = dict(c, size=(60, 30), step=1.25)
    out = barnes(sub['pts'], sub['val'], sub['sigma'], sub['x0'], sub['step'], sub['size'], method='naive')
    ref = orc.interpolate_naive(sub['pts'], sub['val'], sub['sigma'], sub['x0'], sub['step'], sub['size'])
    assert out.dtype == np.float64
    np.testing.assert_array_equal(np.isnan(out), np.isnan(ref))
    np.testing.assert_allclose(out, ref, rtol=1e-12)


def test_radius_matches_oracle(orc):
    from fastbarnes.interpolation import barnes
    c = _configs()['cfg1']
    out = barnes(c['pts'], c['val'], 1.0, c['x0'], 0.5, (150, 75), method='radius')
    ref = orc.interpolate_radius(c['pts'], c['val'], 1.0, c['x0'], 0.5, (150, 75))
    np.testing.assert_array_equal(np.isnan(out), np.isnan(ref))
    np.testing.assert_allclose(out, ref, rtol=1e-12)
