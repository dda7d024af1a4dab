"""
Pins the CPU oracle (oracle/barnes_oracle.c) before anything is checked against it.

Each check is chosen so that a plausible slip in a restatement of the reference
(a dropped tail term, an off-by-one window edge, a wrong corner weight, a swapped
axis) fails at least one of them:

* the literal known answers of the reference's tests/AccumulationTest.py;
* brute-force O(L*K) convolutions with a zero-padded tailed kernel;
* closed forms: impulse mass (K+2a)^n, kernel variance = sigma^2, T and alpha;
* the reference's GaussianApproximationTest tolerance tables and BasicTest ranges;
* invariants of the whole interpolation (constant field, affine map, separability,
  single station, alpha=0 equivalence) and brute-force injection / naive sums.

CPU only (no GPU marker).
"""
import math

import numpy as np
import pytest


# ---------------------------------------------------------------------------
# helpers: independent brute-force statements (numpy, no running sums)

def box_bruteforce(line, T, alpha, num_iter):
    """n-fold convolution with [alpha, 1 x (2T+1), alpha], zero padding every pass."""
    kern = np.ones(2 * T + 3)
    kern[0] = kern[-1] = alpha
    out = np.asarray(line, dtype=np.float64)
    for _ in range(num_iter):
        out = np.convolve(out, kern, mode='same')
    return out


def expand(case):
    a = np.zeros(case['size'])
    for i, v in case['impulses']:
        a[i] = v
    return a


def expected(case):
    e = np.zeros(case['size'])
    for s, t, v in case['expect']:
        e[s:t] = v
    return e


# ---------------------------------------------------------------------------
# known answers of the reference's tests (exact)

def test_accumulation_known_answers(orc, known_answers):
    for case in known_answers['accumulation']:
        inp = expand(case)
        if case['fn'] == 'tail':
            out = orc.accumulate_tail_array(inp, case['rect_len'], case['num_iter'], case['alpha'])
        else:
            out = orc.accumulate_array(inp, case['rect_len'], case['num_iter'])
        np.testing.assert_array_equal(out, expected(case), err_msg=case['src'])


def test_accumulation_versions(orc, known_answers):
    v = known_answers['accumulation_versions']
    inp = expand(v)
    for p in v['pairs']:
        a = orc.accumulate_array(inp, p['plain_rect_len'], v['num_iter'])
        b = orc.accumulate_tail_array(inp, p['tail_rect_len'], v['num_iter'], p['alpha'])
        np.testing.assert_array_equal(a, b, err_msg=v['src'])


def _mutant_tail(in_arr, rect_len, num_iter, alpha, mutation):
    """Python restatement of the tailed running sum with one deliberate slip."""
    a = np.array(in_arr, dtype=np.float64)
    L = len(a)
    h0 = (rect_len - 1) // 2
    h1 = rect_len - h0
    for _ in range(num_iter):
        h = np.zeros(L)
        accu = 0.0
        for k in range(-h0, 0):
            accu += a[k + h0]
        for k in range(0, h1):
            accu += a[k + h0]
            h[k] = accu + (0.0 if mutation == 'drop_b_tail' else alpha * a[k + h0 + 1])
        for k in range(h1, L - h0 - 1):
            accu += a[k + h0] - a[k - h1 + (1 if mutation == 'off_by_one' else 0)]
            h[k] = accu + alpha * (a[k - h1] + a[k + h0 + 1])
        k = L - h0 - 1
        accu += a[k + h0] - a[k - h1]
        h[k] = accu + (alpha * a[k - h1] if mutation != 'drop_c1' else 0.0)
        for k in range(L - h0, L):
            accu -= a[k - h1]
            h[k] = accu + (alpha * a[k - h1] if mutation != 'tail_sign' else -alpha * a[k - h1])
        a = h
    return a


@pytest.mark.parametrize('mutation', ['drop_b_tail', 'off_by_one', 'drop_c1', 'tail_sign'])
def test_pins_catch_mutants(known_answers, mutation):
    """Each planted slip fails at least one pin (known answers or brute force).

    The reference's own known answers never exercise the phase-c' tail term
    (drop_c1 passes them), which is why the brute-force pin is needed as well.
    """
    failed = []
    for case in known_answers['accumulation']:
        if case['fn'] != 'tail':
            continue
        out = _mutant_tail(expand(case), case['rect_len'], case['num_iter'], case['alpha'], mutation)
        if not np.array_equal(out, expected(case)):
            failed.append('known')
    rng = np.random.default_rng(0)
    line = rng.normal(size=64)
    ref = box_bruteforce(line, 4, 0.37, 3)
    out = _mutant_tail(line, 9, 3, 0.37, mutation)
    if np.max(np.abs(out - ref)) > 1e-10:
        failed.append('bruteforce')
    assert failed
    if mutation == 'drop_c1':
        assert failed == ['bruteforce']


def test_unmutated_restatement_passes(known_answers):
    for case in known_answers['accumulation']:
        if case['fn'] == 'tail':
            np.testing.assert_array_equal(
                _mutant_tail(expand(case), case['rect_len'], case['num_iter'], case['alpha'], None), expected(case))


# ---------------------------------------------------------------------------
# brute force and closed forms of the line pass

@pytest.mark.parametrize('T,alpha,n,L', [(3, 0.0, 4, 300), (27, 0.2083333333, 4, 2400), (9, 0.890625, 4, 128),
                                        (5, 0.5, 1, 40), (12, 0.99, 6, 150), (1, 0.3, 3, 7)])
def test_line_pass_bruteforce(orc, T, alpha, n, L):
    rng = np.random.default_rng(T * 1000 + L)
    line = rng.normal(size=L) * np.exp(rng.normal(size=L))
    ref = box_bruteforce(line, T, alpha, n)
    out = orc.accumulate_tail_array(line, 2 * T + 1, n, alpha)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(out - ref)) <= 1e-12 * scale * L
    if alpha == 0.0:
        np.testing.assert_array_equal(out, orc.accumulate_array(line, 2 * T + 1, n))


@pytest.mark.parametrize('sigma,step,n', [(1.0, 1 / 32, 4), (1.5, 0.125, 4), (0.7, 0.05, 3), (2.0, 0.125, 6),
                                          (1.0, 0.25, 4), (0.5, 1 / 64, 5)])
def test_impulse_mass_and_variance(orc, sigma, step, n):
    """Mass (K+2a)^n and second moment n*(K*T(T+1)/3 + 2a(T+1)^2)/(K+2a) * step^2 == sigma^2."""
    T = int(orc.half_kernel_size_opt(sigma, step, n)[0])
    a = float(orc.tail_value(sigma, step, n)[0])
    K = 2 * T + 1
    L = 2 * n * (T + 1) + 41
    line = np.zeros(L)
    c = L // 2
    line[c] = 1.0
    out = orc.accumulate_tail_array(line, K, n, a)
    mass = out.sum()
    assert mass == pytest.approx((K + 2 * a) ** n, rel=1e-12)
    x = (np.arange(L) - c) * step
    var = float((out * x * x).sum() / mass)
    assert var == pytest.approx(sigma ** 2, rel=1e-10)
    var_formula = n * (K * T * (T + 1) / 3 + 2 * a * (T + 1) ** 2) / (K + 2 * a) * step ** 2
    assert var_formula == pytest.approx(sigma ** 2, rel=1e-12)


def test_kernel_parameters_closed_form(orc):
    # the derived table: (sigma, step, n) -> (T, alpha)
    assert orc.half_kernel_size_opt(1.0, 0.25, 4)[0] == 3
    assert orc.tail_value(1.0, 0.25, 4)[0] == 0.0
    assert orc.half_kernel_size_opt(1.0, 1 / 32, 4)[0] == 27
    assert orc.tail_value(1.0, 1 / 32, 4)[0] == pytest.approx(0.2083333333333, rel=1e-9)
    assert orc.half_kernel_size_opt(1.5, 0.125, 4)[0] == 9
    assert orc.tail_value(1.5, 0.125, 4)[0] == pytest.approx(0.890625, rel=1e-12)
    rng = np.random.default_rng(5)
    for _ in range(200):
        s = rng.uniform(0.3, 3.0)
        d = rng.uniform(0.01, 0.3)
        n = int(rng.integers(1, 11))
        T = int(orc.half_kernel_size_opt(s, d, n)[0])
        # T is the largest integer whose plain-box variance n*T(T+1)/3*d^2 stays <= sigma^2
        assert n * T * (T + 1) / 3 * d * d <= s * s * (1 + 1e-12)
        assert n * (T + 1) * (T + 2) / 3 * d * d > s * s * (1 - 1e-12)
        a = float(orc.tail_value(s, d, n)[0])
        assert -1e-9 <= a <= 1.0 + 1e-9
        Tc = int(orc.half_kernel_size(s, d, n)[0])
        assert Tc == int(math.sqrt(3.0 / n) * s / d + 0.5)
        se = float(orc.sigma_effective(s, d, n)[0])
        assert se == pytest.approx(math.sqrt(n / 3.0 * Tc * (Tc + 1)) * d, rel=1e-15)


def test_threshold_closed_form(orc):
    sig, st, n = [1.0, 2.0], [1 / 32, 0.125], 4
    T = orc.half_kernel_size_opt(sig, st, n, dim=2)
    a = orc.tail_value(sig, st, n, dim=2)
    K = 2 * T + 1
    mdw = math.exp(-3.5 ** 2 / 2)
    thr = orc.threshold(True, sig, st, K, n, a, mdw, 2)
    expect = np.prod((K + 2 * a) ** n / math.sqrt(2 * math.pi) / (np.array(sig) / np.array(st))) * mdw
    assert thr == pytest.approx(expect, rel=1e-14)
    thr0 = orc.threshold(False, sig, st, K, n, None, mdw, 2)
    expect0 = np.prod(K.astype(np.float64) ** n / math.sqrt(2 * math.pi) / (np.array(sig) / np.array(st))) * mdw
    assert thr0 == pytest.approx(expect0, rel=1e-14)


def _binary_power(a, e):
    r = 1.0
    while e:
        if e & 1:
            r *= a
        e >>= 1
        a *= a
    return r


def _threshold_params():
    """(sigma, step, n) over the uneven batch's ranges: sigma ~ U[0.5, 2], W in {600, 1200, 2400},
    step = 75/W, n in 3..6 (744 combinations)."""
    rng = np.random.default_rng(2024)
    out = []
    for W in (600, 1200, 2400):
        for n in (3, 4, 5, 6):
            for s in rng.uniform(0.5, 2.0, 62):
                out.append((float(s), 75.0 / W, n))
    return out


def test_threshold_power_is_libm_pow(orc):
    """The tailed threshold's (K + 2 alpha)**n (interpolation.py:424) is NumPy's float64 power loop,
    i.e. the C library's pow.  Three pins over 744 parameter sets of the uneven batch's ranges:
    * the oracle's per-axis factor p / sqrt(2 pi) / (sigma/step) * max_dist_weight is the one formed
      from pow(K + 2 alpha, n) bit for bit;
    * that pow is checked against the exact rational power (fractions), independent of any pow
      implementation: within 0.52 ulp everywhere, and the correctly rounded value wherever the
      exact power is more than 1/50 ulp away from a rounding midpoint (glibc's pow is not
      correctly rounded at near-midpoints);
    * binary exponentiation, which a restatement might use instead, differs from it on some sets."""
    from fractions import Fraction
    params = _threshold_params()
    assert len(params) >= 700
    differ = near_mid = 0
    for s, d, n in params:
        T = int(orc.half_kernel_size_opt(s, d, n)[0])
        a = float(orc.tail_value(s, d, n)[0])
        base = float(2 * T + 1) + 2 * a
        p = math.pow(base, n)
        mdw = 0.5   # a power of two, exact
        expect = p / math.sqrt(2 * math.pi) / (s / d) * mdw
        assert orc.threshold(True, s, d, 2 * T + 1, n, a, mdw, 1) == expect, (s, d, n)
        exact = Fraction(base) ** n
        rounded = float(exact)               # correctly rounded (exact integer division)
        ulp = Fraction(math.ulp(rounded))
        err = abs(Fraction(p) - exact) / ulp
        assert err <= Fraction(52, 100), (s, d, n, float(err))
        dist_mid = abs(abs(Fraction(rounded) - exact) / ulp - Fraction(1, 2))
        if dist_mid > Fraction(1, 50):
            assert p == rounded, (s, d, n)
        else:
            near_mid += 1
        differ += _binary_power(base, n) != p
    assert near_mid < len(params) // 20
    assert differ > 0, 'the parameter set does not separate pow from binary exponentiation'


def test_nan_value_propagates_to_offset(orc):
    """np.amin/np.amax (interpolation.py:209) propagate NaN: a NaN value anywhere makes the
    offset NaN, and the whole result NaN."""
    rng = np.random.default_rng(3)
    pts = rng.uniform(0, 10, (50, 2))
    for idx in (0, 1, 17, 49):
        val = rng.normal(size=50)
        val[idx] = np.nan
        v = val.copy()
        _, _, off = orc.inject(pts, v, (0.0, 0.0), 0.25, (40, 40))
        assert math.isnan(off)
        out = orc.interpolate_convol(pts, val, 1.0, (0.0, 0.0), 0.25, (40, 40))
        assert np.isnan(out).all(), idx


# ---------------------------------------------------------------------------
# the reference's Gaussian convergence tables (tests/GaussianApproximationTest.py)

@pytest.mark.parametrize('opt', [False, True])
def test_gaussian_approximation(orc, known_answers, opt):
    g = known_answers['gaussian']
    table = g['opt'] if opt else g['plain']
    for sigma in g['sigmas']:
        for num_iter, max_diff in table:
            delta = 1.0 / g['resol']
            width = g['width_sigmas'] * sigma
            grid = np.arange(-width, width + 0.000001, delta)
            if opt:
                s_ref = sigma
                T = orc.half_kernel_size_opt(sigma, delta, num_iter)
                tv = orc.tail_value(sigma, delta, num_iter)
            else:
                s_ref = float(orc.sigma_effective(sigma, delta, num_iter)[0])
                T = orc.half_kernel_size(sigma, delta, num_iter)
                tv = np.zeros(1)
            gauss = 1.0 / math.sqrt(2.0 * math.pi) / s_ref * np.exp(-(grid / s_ref) ** 2 / 2)
            values = np.zeros(len(grid))
            values[len(grid) // 2] = 1.0 / delta
            weights = values.copy()
            orc.convolve(opt, values, weights, [sigma], [delta], (len(grid),), 2 * T + 1, num_iter, tv, 9999.9)
            values *= (delta / 2 / math.sqrt(3 / num_iter) / s_ref) ** num_iter
            assert np.max(np.abs(gauss - values)) <= max_diff / s_ref, (sigma, num_iter)


# ---------------------------------------------------------------------------
# injection by brute force

def inject_bruteforce(pts, val, x0, step, size):
    pts = np.atleast_2d(np.asarray(pts, dtype=np.float64))
    if pts.shape[0] == 1 and len(size) == 1:
        pts = pts.reshape(-1, 1)
    dim = len(size)
    v = np.asarray(val, dtype=np.float64).copy()
    off = (v.min() + v.max()) / 2.0
    v -= off
    vg = np.zeros(size[::-1])
    wg = np.zeros(size[::-1])
    for k in range(len(v)):
        c = [(pts[k, m] - x0[m]) / step[m] for m in range(dim)]
        if any(cm < 0.0 or cm >= size[m] - 1 for m, cm in enumerate(c)):
            continue
        i = [int(cm) for cm in c]
        f = [cm - im for cm, im in zip(c, i)]
        for corner in range(1 << dim):
            w = 1.0
            idx = []
            for m in range(dim):
                b = (corner >> m) & 1
                w *= f[m] if b else (1.0 - f[m])
                idx.append(i[m] + b)
            vg[tuple(idx[::-1])] += w * v[k]
            wg[tuple(idx[::-1])] += w
    return vg, wg, off


@pytest.mark.parametrize('dim', [1, 2, 3])
def test_injection_bruteforce(orc, dim):
    rng = np.random.default_rng(dim)
    size = [17, 11, 7][:dim]
    step = [0.5, 0.25, 1.0][:dim]
    x0 = [-1.0, 2.0, 0.5][:dim]
    n = 200
    ext = np.array([s * d for s, d in zip(size, step)])
    pts = rng.uniform(-0.1, 1.1, (n, dim)) * ext + np.array(x0)
    pts[:5] = np.array(x0)  # exactly on the first grid line: kept
    pts[5:8] = np.array(x0) + (np.array(size) - 1) * np.array(step)  # on the last line: dropped
    val = rng.normal(size=n) * 3 + 10
    vg, wg, off = orc.inject(pts if dim > 1 else pts[:, 0], val, x0, step, size)
    vb, wb, ob = inject_bruteforce(pts, val, x0, step, size)
    assert off == ob
    # the corner order only permutes which corner gets which product: sums agree to rounding
    np.testing.assert_allclose(vg, vb, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(wg, wb, rtol=1e-13, atol=1e-13)
    inside = np.all((pts - x0) / step >= 0, axis=1) & np.all((pts - x0) / step < np.array(size) - 1, axis=1)
    assert wg.sum() == pytest.approx(inside.sum(), rel=1e-13)
    # bilinear weights reproduce the point positions (first moment)
    for m in range(dim):
        coord = np.arange(size[m]).reshape([-1 if a == dim - 1 - m else 1 for a in range(dim)])
        c = (pts[inside, m] - x0[m]) / step[m]
        assert (wg * coord).sum() == pytest.approx(c.sum(), rel=1e-12)


# ---------------------------------------------------------------------------
# whole-interpolation invariants

def _small_case(seed, n=300, size=(60, 40), step=0.25):
    rng = np.random.default_rng(seed)
    x0 = np.array([-2.0, 1.0])
    ext = np.array(size) * step
    pts = rng.uniform(0, 1, (n, 2)) * ext + x0
    val = rng.normal(size=n) * 2 + 5
    return pts, val, x0, step, size


def test_constant_field(orc):
    pts, val, x0, step, size = _small_case(1)
    out = orc.interpolate_convol(pts, np.full(len(pts), 7.25), 1.0, x0, step, size)
    ok = ~np.isnan(out)
    assert ok.sum() > 0.5 * out.size
    np.testing.assert_allclose(out[ok], 7.25, rtol=0, atol=2e-6)


def test_affine_invariance(orc):
    pts, val, x0, step, size = _small_case(2)
    a, b = -3.0, 11.0
    f = orc.interpolate_convol(pts, val, 1.0, x0, step, size).astype(np.float64)
    g = orc.interpolate_convol(pts, a * val + b, 1.0, x0, step, size).astype(np.float64)
    np.testing.assert_array_equal(np.isnan(f), np.isnan(g))
    ok = ~np.isnan(f)
    np.testing.assert_allclose(g[ok], a * f[ok] + b, rtol=0, atol=1e-4)


def test_single_station(orc):
    x0, step, size = np.array([0.0, 0.0]), 0.1, (80, 60)
    out = orc.interpolate_convol(np.array([[4.03, 2.97]]), np.array([3.5]), 0.5, x0, step, size, max_dist=3.0)
    ok = ~np.isnan(out)
    assert 0 < ok.sum() < out.size
    np.testing.assert_array_equal(out[ok], np.float32(3.5))
    # mask radius ~ max_dist*sigma around the station
    yy, xx = np.mgrid[0:size[1], 0:size[0]] * step
    d = np.hypot(xx - 4.03, yy - 2.97)
    assert d[ok].max() < 3.0 * 0.5 * 1.3 and d[~ok].min() > 3.0 * 0.5 * 0.7


def test_separability(orc):
    """2D impulse response = outer product of the two 1D responses (grid-aligned station)."""
    W, H, step = 90, 70, 0.1
    pts = np.array([[4.0, 3.0]])
    size = (W, H)
    out, vg, wg = orc.interpolate_convol(pts, np.array([1.0]), [0.6, 0.9], [0.0, 0.0], step, size,
                                         max_dist=1e9, return_fields=True)
    _, vx, wx = orc.interpolate_convol(pts[:, 0], np.array([1.0]), 0.6, 0.0, step, W, max_dist=1e9,
                                       return_fields=True)
    _, vy, wy = orc.interpolate_convol(pts[:, 1], np.array([1.0]), 0.9, 0.0, step, H, max_dist=1e9,
                                       return_fields=True)
    outer = np.outer(np.nan_to_num(wy), np.nan_to_num(wx))
    # nodes far from the station hold +-1e-17 running-sum residues; negative ones are masked
    # (wg < 0 = thr), exactly as in the reference
    ok = ~np.isnan(wg)
    assert np.all(np.abs(outer[~ok]) < 1e-12 * outer.max())
    np.testing.assert_allclose(wg[ok], outer[ok], rtol=1e-12, atol=1e-12 * outer.max())
    vok = ~np.isnan(vg)
    vouter = np.outer(np.nan_to_num(vy), np.nan_to_num(vx))
    np.testing.assert_allclose(vg[vok], vouter[vok], rtol=1e-12, atol=1e-12 * outer.max())


def test_alpha_zero_equals_convolution(orc):
    """Config-1 geometry has alpha == 0: optimized == plain convolution bit for bit."""
    from fastbarnes import workloads
    c = workloads.config1()
    a = orc.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], tail=True)
    b = orc.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], tail=False)
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_converges_to_naive(orc):
    """Fast Barnes approximates the exact Gaussian Barnes in the interior, better with larger n."""
    pts, val, x0, step, size = _small_case(3, n=400, size=(80, 60), step=0.1)
    g = orc.interpolate_naive(pts, val, 0.5, x0, step, size)[15:-15, 15:-15]
    errs = []
    for n in (4, 6, 10):
        f = orc.interpolate_convol(pts, val, 0.5, x0, step, size, num_iter=n, max_dist=2.0)
        f = f.astype(np.float64)[15:-15, 15:-15]
        assert not np.isnan(f).any()
        errs.append(np.sqrt(np.mean((f - g) ** 2)))
    assert errs[0] < 0.015 * np.std(val)
    assert errs[0] > errs[1] > errs[2]


def test_naive_bruteforce(orc):
    rng = np.random.default_rng(9)
    pts = rng.uniform(0, 3, (25, 2))
    val = rng.normal(size=25)
    x0, step, size = np.array([0.2, 0.1]), 0.3, (9, 7)
    out = orc.interpolate_naive(pts, val, [0.4, 0.6], x0, step, size)
    off = (val.min() + val.max()) / 2
    ref = np.empty((7, 9))
    for j in range(7):
        for i in range(9):
            w = np.exp(-((pts[:, 0] - (x0[0] + i * step)) ** 2 / (2 * 0.16)
                         + (pts[:, 1] - (x0[1] + j * step)) ** 2 / (2 * 0.36)))
            ref[j, i] = (w * (val - off)).sum() / w.sum() + off
    np.testing.assert_allclose(out, ref, rtol=1e-13)


def test_radius_bruteforce(orc):
    rng = np.random.default_rng(10)
    pts = rng.uniform(0, 3, (40, 2))
    val = rng.normal(size=40)
    x0, step, size = np.array([0.0, 0.0]), 0.25, (13, 11)
    out = orc.interpolate_radius(pts, val, 0.3, x0, step, size, max_dist=3.5, min_weight=0.001)
    r2 = -2.0 * math.log(0.001) * 0.3 ** 2
    off = (val.min() + val.max()) / 2
    for j in range(11):
        for i in range(13):
            d2 = (pts[:, 0] - i * 0.25) ** 2 + (pts[:, 1] - j * 0.25) ** 2
            sel = d2 <= r2 * (1 + 1e-15)
            w = np.exp(-d2[sel] / (2 * 0.09))
            if w.sum() >= math.exp(-3.5 ** 2 / 2):
                assert out[j, i] == pytest.approx((w * (val[sel] - off)).sum() / w.sum() + off, rel=1e-12)
            else:
                assert np.isnan(out[j, i])


# ---------------------------------------------------------------------------
# the reference's BasicTest ranges

def test_basic_ranges_1d_2d(orc, known_answers):
    b = known_answers['basic']
    d1 = b['d1']
    for tail in (False, True):
        for sigma in d1['sigmas']:
            for n in d1['num_iters']:
                r = orc.interpolate_convol(np.array(d1['pts']), np.array(d1['val']), sigma, 0.0, d1['step'],
                                           d1['size'], num_iter=n, tail=tail)
                assert np.all(((r >= -d1['eps']) & (r <= 1 + d1['eps'])) | np.isnan(r))
    d2 = b['d2']
    data = np.array(d2['data'])
    for tail in (False, True):
        for sigma in d2['sigmas']:
            for n in (1, 4, 10):
                r = orc.interpolate_convol(data[:, :2], data[:, 2], sigma, 0.0, d2['step'], tuple(d2['size']),
                                           num_iter=n, tail=tail)
                assert np.all(((r >= -d2['eps']) & (r <= 1 + d2['eps'])) | np.isnan(r))


# ---------------------------------------------------------------------------
# the OpenMP build timed as the CPU baseline

@pytest.mark.parametrize('threads', [1, 3, 8])
def test_omp_build_matches_checker(orc, threads):
    """liboracle_omp.so (-O3, OpenMP over the lines of each pass) gives the checker's bits, in 2D
    (config 1, both methods) and 3D, on any thread count."""
    from fastbarnes import workloads
    c = workloads.config1(seed=5)
    c3 = workloads.config4(n=3000, size=(40, 36, 28))   # plain kernel 21 < 28
    for cc in (c, c3):
        for tail in (True, False):
            a, va, wa = orc.interpolate_convol(cc['pts'], cc['val'], cc['sigma'], cc['x0'], cc['step'], cc['size'],
                                               num_iter=cc['num_iter'], max_dist=cc['max_dist'], tail=tail,
                                               return_fields=True)
            b, vb, wb = orc.interpolate_convol(cc['pts'], cc['val'], cc['sigma'], cc['x0'], cc['step'], cc['size'],
                                               num_iter=cc['num_iter'], max_dist=cc['max_dist'], tail=tail,
                                               return_fields=True, threads=threads)
            assert np.array_equal(va.view(np.uint64), vb.view(np.uint64))
            assert np.array_equal(np.isnan(wa), np.isnan(wb))
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
