"""
CPU restatement of the FP32C line-pass arithmetic (the sliding pass of k_fast_pass, see
fb_kernels.hip slide1), checked against the FP64 oracle.  It documents and pins the
precision choices of the fast path:

* every window sum is exact in FP64 (float32 terms converted, re-anchored at each lane
  segment, slid by the FP64 difference of the entering and leaving terms);
* each output rounds the window sum to float32 and adds the two tail terms with one float32
  FMA (times the float32 scale on a scaling iteration);
* the quotient is formed in float32, the value offset added in FP64.

A variant that sums the window re-anchoring in float32 quads (rejected) loses an order of
magnitude in the maximum error where the weight field falls steeply; the test shows it.
The oracle is the checker; nothing here runs on the GPU.
"""
import math

import numpy as np

f32, f64 = np.float32, np.float64
SEG = 41


def _scale_groups(g, n, total):
    m = n
    if g > 1.0:
        m = max(1, min(n, int(math.floor(60.0 * math.log(2.0) / math.log(g)))))
    ngroups = (n + m - 1) // m
    last = n - m * (ngroups - 1)
    sgscale = (1.0 / g) ** m if ngroups > 1 else 1.0
    slast = (1.0 / g) ** last if ngroups > 1 else total
    return m, sgscale, slast


def _line_pass(X, T, alpha, n, total, quad_init=False):
    """n tailed-box iterations of the float32 lines X (nlines, L), as the kernel does them."""
    nl, L = X.shape
    K = 2 * T + 1
    m, sgscale, slast = _scale_groups(K + 2 * alpha, n, total)
    pad = T + 1
    nlanes = -(-L // SEG)
    W = pad + nlanes * SEG + T + 2
    s = pad + np.arange(nlanes) * SEG
    a32 = f32(alpha)
    for it in range(n):
        last = it + 1 == n
        scale = slast if last else (sgscale if (it + 1) % m == 0 else 1.0)
        Xp = np.zeros((nl, W), f32)
        Xp[:, pad:pad + L] = X
        a = np.zeros((nl, nlanes), f64)
        if quad_init:
            for q in range(0, K, 4):
                acc = np.zeros((nl, nlanes), f32)
                for u in range(min(4, K - q)):
                    acc = (acc + Xp[:, s - T + q + u]).astype(f32)
                a += acc.astype(f64)
        else:
            for k in range(K):
                a += Xp[:, s - T + k].astype(f64)
        out = np.zeros((nl, nlanes * SEG), f32)
        pf = Xp[:, s - T - 1]
        for j in range(SEG):
            ad = Xp[:, s + j + T + 1]
            sv = Xp[:, s + j - T]
            t = (pf + ad).astype(f32)
            r = (f64(a32) * t.astype(f64) + a.astype(f32).astype(f64)).astype(f32)   # float32 FMA
            if scale != 1.0:
                r = (r * f32(scale)).astype(f32)
            a += ad.astype(f64) - sv.astype(f64)
            pf = sv
            out[:, np.arange(nlanes) * SEG + j] = r
        X = out[:, :L]
    return X


def _fp32c(orc, pts, val, sigma, x0, step, size, n, quad_init=False, max_dist=3.5):
    vg, wg, off = orc.inject(pts, val, x0, step, size)
    sg = np.array([sigma, sigma], f64)
    st = np.array([step, step], f64)
    T = orc.half_kernel_size_opt(sg, st, n, dim=2)
    al = orc.tail_value(sg, st, n, dim=2)
    K = 2 * T + 1
    thr = orc.threshold(True, sg, st, K.astype(np.int32), n, al, math.exp(-max_dist ** 2 / 2), 2)
    ps = [(1.0 / (K[m] + 2 * al[m])) ** n for m in range(2)]
    V, Wt = vg.astype(f32), wg.astype(f32)
    V = _line_pass(V, int(T[0]), al[0], n, ps[0], quad_init)
    Wt = _line_pass(Wt, int(T[0]), al[0], n, ps[0], quad_init)
    V = _line_pass(np.ascontiguousarray(V.T), int(T[1]), al[1], n, ps[1], quad_init).T
    Wt = _line_pass(np.ascontiguousarray(Wt.T), int(T[1]), al[1], n, ps[1], quad_init).T
    q = (V / Wt).astype(f32)
    r = (q.astype(f64) + off).astype(f32)
    r[Wt.astype(f64) < thr * ps[0] * ps[1]] = np.nan
    return r


def _errors(out, ref):
    np.testing.assert_array_equal(np.isnan(out), np.isnan(ref))
    ok = ~np.isnan(ref)
    d = out[ok].astype(f64) - ref[ok]
    return float(np.sqrt(np.mean(d * d))), float(np.abs(d).max())


def test_fp32c_arithmetic_matches_fp64(orc):
    """Config 1 (300x150, alpha = 0) and a 600x300 field with alpha != 0 (T = 6): the
    restated FP32C arithmetic against the oracle, within float32 rounding of the output."""
    from fastbarnes import workloads
    c = workloads.config1()
    cases = [(c['pts'], c['val'], 1.0, c['x0'], 0.25, (300, 150), 4)]
    pts = workloads.stations_2d(2000, 11)
    cases.append((pts, workloads.values_sinusoids(pts, 11), 1.0, np.array([-26.0, 34.0]), 0.125, (600, 300), 4))
    for pts, val, sigma, x0, step, size, n in cases:
        ref = orc.interpolate_convol(pts, val, sigma, x0, step, size, num_iter=n)
        rmse, mx = _errors(_fp32c(orc, pts, val, sigma, x0, step, size, n), ref)
        assert rmse <= 1e-6 and mx <= 4e-6, (size, rmse, mx)


def test_float32_reanchoring_is_worse(orc):
    """Summing the window re-anchoring in float32 quads (then FP64) raises the maximum error
    by an order of magnitude on config 1: the reason the kernel converts every term."""
    from fastbarnes import workloads
    c = workloads.config1()
    ref = orc.interpolate_convol(c['pts'], c['val'], 1.0, c['x0'], 0.25, (300, 150), num_iter=4)
    _, mx = _errors(_fp32c(orc, c['pts'], c['val'], 1.0, c['x0'], 0.25, (300, 150), 4), ref)
    _, mxq = _errors(_fp32c(orc, c['pts'], c['val'], 1.0, c['x0'], 0.25, (300, 150), 4, quad_init=True), ref)
    assert mxq > 5 * mx, (mx, mxq)
