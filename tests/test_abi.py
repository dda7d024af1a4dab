"""
CPU checks of the C ABI and of the host logic of the drop-in (no GPU needed):

* libfastbarnes.so loads and exports every function declared in include/fastbarnes.h;
* the host-side kernel-parameter math equals the oracle's bit for bit;
* barnes() raises the reference's RuntimeErrors (tests/BasicTest.py:105-148 and
  interpolation.py:105-199) before anything reaches the GPU.
"""
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, 'include', 'fastbarnes.h')) as f:
        text = f.read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(fb_[a-z_0-9]+)\s*\(', text)))


def test_library_exports_header_symbols():
    from fastbarnes import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    assert L.fb_version() >= 10000


def test_field_struct_layout():
    import ctypes
    from fastbarnes import _lib
    # int32 x4, int64 x3, int64[3], double[3] x3, double
    assert ctypes.sizeof(_lib.FbField) == 16 + 24 + 24 + 72 + 8


@pytest.mark.parametrize('seed', range(5))
def test_parameters_match_oracle(orc, seed):
    from fastbarnes import interpolation as I
    rng = np.random.default_rng(seed)
    for _ in range(100):
        s = float(rng.uniform(0.05, 4.0))
        d = float(rng.choice([1 / 512, 1 / 64, 1 / 32, 0.05, 0.125, 0.25, rng.uniform(0.01, 0.5)]))
        n = int(rng.integers(1, 11))
        assert I.get_half_kernel_size_opt(s, d, n) == orc.half_kernel_size_opt(s, d, n)[0]
        assert I.get_half_kernel_size(s, d, n) == orc.half_kernel_size(s, d, n)[0]
        a = I.get_tail_value(s, d, n)
        b = orc.tail_value(s, d, n)[0]
        assert a == b or (math.isnan(a) and math.isnan(b))
        assert I.get_sigma_effective(s, d, n) == orc.sigma_effective(s, d, n)[0]


def test_threshold_matches_oracle(orc):
    import ctypes
    from fastbarnes import _lib
    rng = np.random.default_rng(77)
    extra = [([float(rng.uniform(0.5, 2.0))] * 2, [75.0 / W] * 2, int(n)) for W in (600, 1200, 2400)
             for n in (3, 4, 5, 6) for _ in range(8)]
    for sig, st, n in [([1.0, 1.0], [1 / 32, 1 / 32], 4), ([1.5] * 3, [0.125] * 3, 4), ([0.7], [0.01], 6),
                       ([2.0, 0.5], [0.1, 0.05], 3)] + extra:
        dim = len(sig)
        sig_a, st_a = np.array(sig), np.array(st)
        for tail in (True, False):
            T = (orc.half_kernel_size_opt if tail else orc.half_kernel_size)(sig, st, n, dim=dim)
            K = (2 * T + 1).astype(np.int32)
            a = orc.tail_value(sig, st, n, dim=dim) if tail else np.zeros(dim)
            mdw = math.exp(-3.5 ** 2 / 2)
            ref = orc.threshold(tail, sig, st, K, n, a, mdw, dim)
            out = ctypes.c_double()
            dp = ctypes.POINTER(ctypes.c_double)
            _lib.check(_lib.lib().fb_threshold(0 if tail else 1, dim, sig_a.ctypes.data_as(dp),
                                               st_a.ctypes.data_as(dp),
                                               K.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n,
                                               a.ctypes.data_as(dp), mdw, ctypes.byref(out)))
            assert out.value == ref


def test_validation_errors():
    from fastbarnes.interpolation import barnes
    pts = np.array([[0.1, 0.2], [0.5, 0.5]])
    val = np.array([1.0, 2.0])
    with pytest.raises(RuntimeError, match='specified pts is not a numpy ndarray'):
        barnes([[0.1, 0.2]], val, 1.0, 0.0, 0.1, (20, 20))
    with pytest.raises(RuntimeError, match=r'expected pts array of shape \(N, M\)'):
        barnes(np.zeros((2, 2, 2)), val, 1.0, 0.0, 0.1, (20, 20))
    with pytest.raises(RuntimeError, match='dimensions 1, 2 or 3'):
        barnes(np.zeros((2, 4)), val, 1.0, 0.0, 0.1, (20, 20, 20, 20))
    with pytest.raises(RuntimeError, match='specified val is not a numpy ndarray'):
        barnes(pts, [1.0, 2.0], 1.0, 0.0, 0.1, (20, 20))
    with pytest.raises(RuntimeError, match=r'expected val array of shape \(N\)'):
        barnes(pts, np.zeros((2, 1)), 1.0, 0.0, 0.1, (20, 20))
    with pytest.raises(RuntimeError, match='inconsistent shapes'):
        barnes(pts, np.zeros(3), 1.0, 0.0, 0.1, (20, 20))
    with pytest.raises(RuntimeError, match='specified sigma with invalid length: 3'):
        barnes(pts, val, [1.0, 1.0, 1.0], 0.0, 0.1, (20, 20))
    with pytest.raises(RuntimeError, match='specified x0 with invalid length: 1'):
        barnes(pts, val, 1.0, [0.0], 0.1, (20, 20))
    with pytest.raises(RuntimeError, match='specified step with invalid length: 3'):
        barnes(pts, val, 1.0, 0.0, (0.1, 0.1, 0.1), (20, 20))
    with pytest.raises(RuntimeError, match='specified size with invalid length: 1'):
        barnes(pts, val, 1.0, 0.0, 0.1, (20,))
    with pytest.raises(RuntimeError, match='array size should be array-like of length: 2'):
        barnes(pts, val, 1.0, 0.0, 0.1, 20)
    with pytest.raises(RuntimeError, match='encountered invalid Barnes interpolation method: foo'):
        barnes(pts, val, 0.05, 0.0, 0.1, (20, 20), method='foo')
    with pytest.raises(RuntimeError, match='radius algorithm works only in 2D but data is: 1D'):
        barnes(np.array([0.1, 0.2]), val, 0.05, 0.0, 0.1, 20, method='radius')
    with pytest.raises(RuntimeError, match='only for scalar sigma value'):
        barnes(pts, val, [0.05, 0.1], 0.0, 0.1, (20, 20), method='radius')


def test_kernel_size_check(known_answers):
    """tests/BasicTest.py:105-148: a grid of kernel_size raises, for both convolution methods."""
    from fastbarnes import interpolation
    kc = known_answers['basic']['kernel_check']
    pts, val = np.array(kc['pts']), np.array(kc['val'])
    for method, fn in (('convolution', interpolation.get_half_kernel_size),
                       ('optimized_convolution', interpolation.get_half_kernel_size_opt)):
        for n in kc['num_iters']:
            ks = 2 * fn(kc['sigma'], kc['step'], n) + 1
            with pytest.raises(RuntimeError, match='resulting rectangular kernel size should be smaller'):
                interpolation.barnes(pts, val, kc['sigma'], 0.0, kc['step'], ks, method=method, num_iter=n)


def test_precision_and_extra_validation():
    from fastbarnes.interpolation import barnes
    pts = np.array([[0.1, 0.2], [0.5, 0.5]])
    val = np.array([1.0, 2.0])
    with pytest.raises(RuntimeError, match='unknown precision'):
        barnes(pts, val, 0.05, 0.0, 0.01, (200, 200), precision='fp16')
    with pytest.raises(RuntimeError, match='sigma and step must be finite and positive'):
        barnes(pts, val, 0.05, 0.0, -0.01, (200, 200))
    with pytest.raises(RuntimeError, match='NaN'):
        barnes(np.array([[np.nan, 0.2], [0.5, 0.5]]), val, 0.05, 0.0, 0.01, (200, 200))
    with pytest.raises(RuntimeError, match='at least one sample point'):
        barnes(np.zeros((0, 2)), np.zeros(0), 0.05, 0.0, 0.01, (200, 200))
