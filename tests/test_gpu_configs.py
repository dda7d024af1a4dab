"""
GPU tests at the benchmark configurations' full sizes and at the limits of the FP32C line
passes (marked `gpu`; through the C ABI, against the oracle).

* FP32C lines longer than one wave's lanes (several waves per line): 2D fields with lines of
  4000-12000 elements and stations up to the line ends, and a dense 1D line.
* NaN values anywhere: the offset (np.amin/np.amax, interpolation.py:209) is NaN, every node NaN.
* the uneven batch of 256 fields (N~logU[500,50000], grids 600x300..2400x1200, sigma~U[0.5,2],
  n in 3..6) in one plan: every field equals its single-field run bit for bit, graph replays of
  the autotuned plan equal its eager run, and 8 fields meet the FP32C bar against the oracle;
* 3D 512x512x128 with N=200000: FP64 bit-identical to the oracle, FP32C within the bar;
* the batch of 32 3D fields of 256x256x64 with N~logU[1e4,2e5]: FP64 and FP32C against the
  oracle on a sample of fields;
* a plan destroyed right after an asynchronous execute: its buffers return to the pool only
  after its kernels finish;
* FP32C lines longer than one line pass takes (x lines of 16384-40000 nodes in 1D/2D/3D, y lines
  of 16384, a half kernel of 172), run in chunks with num_iter*(T+1) halos, against the oracle;
  chunks along the middle axis of a 3D grid are refused.
"""
import math

import numpy as np
import pytest

from test_gpu_parity import MAXABS_TOL, RMSE_TOL, _oracle, _run, _thr, assert_bitwise, fast_errors

pytestmark = pytest.mark.gpu


def _desc(c):
    return dict(sigma=c['sigma'], x0=c['x0'], step=c['step'], size=c['size'], n=len(c['val']),
                num_iter=c['num_iter'], max_dist=c['max_dist'])


def _check_fast(orc, c, out, label):
    ref, vg, wg = _oracle(orc, c)
    rmse, maxabs, mism_far, mism = fast_errors(out, ref, wg, _thr(orc, c))
    print('%s: rmse %.3g maxabs %.3g mask mismatches %d (%d near the threshold)' % (label, rmse, maxabs, mism,
                                                                                    mism - mism_far), flush=True)
    assert mism_far == 0 and rmse <= RMSE_TOL and maxabs <= MAXABS_TOL, (label, rmse, maxabs, mism)
    return ref


# ---------------------------------------------------------------------------
# FP32C lines of several waves

def _long_line_case(W, H, seed, transpose=False):
    """Stations uniform over the whole grid (the right end of every line included), a small
    sigma (half kernel 3), values with structure on the kernel's scale."""
    rng = np.random.default_rng(seed)
    step = 0.01
    ext = np.array([W - 1, H - 1]) * step
    n = int(W * H // 40)
    pts = rng.uniform(0.0, 1.0, (n, 2)) * ext
    pts[: H, 0] = ext[0] * (1.0 - rng.uniform(0.0, 1e-3, H))   # stations right at the line ends
    pts[: H, 1] = rng.uniform(0.0, ext[1], H)
    val = 5.0 + np.sin(pts[:, 0] * 7.0) * np.cos(pts[:, 1] * 3.0) + rng.normal(0.0, 0.3, n)
    c = dict(pts=pts, val=val, sigma=0.05, x0=np.zeros(2), step=step, size=(W, H), num_iter=4, max_dist=3.5)
    if transpose:
        c = dict(c, pts=np.ascontiguousarray(pts[:, ::-1]), size=(H, W))
    return c


@pytest.mark.parametrize('W,H,transpose', [(4000, 48, False), (6000, 24, False), (12000, 16, False),
                                           (4000, 40, True)])
def test_fp32c_long_lines(orc, W, H, transpose):
    c = _long_line_case(W, H, W + H, transpose)
    out = _run(c, precision='fp32c')
    ref = _check_fast(orc, c, out, 'long lines %s' % (c['size'],))
    # the line ends carry data: the last columns (rows) are not all masked
    edge = ref[:, -3:] if not transpose else ref[-3:, :]
    assert np.isfinite(edge).mean() > 0.5
    assert_bitwise(_run(c, precision='fp64'), ref)


def test_fp32c_dense_long_line(orc):
    """A dense 1D line of 4000 elements (several waves, data up to the end) through the FP32C line
    pass, to float32 rounding of the FP64 result."""
    from fastbarnes import interpolation
    rng = np.random.default_rng(12)
    interpolation._precision_private = 'fp32c'
    try:
        for T, alpha, n, L in [(2, 0.1, 4, 4000), (3, 0.4, 3, 6001), (5, 0.7, 4, 12000)]:
            line = 1.0 + rng.uniform(0.0, 1.0, L)
            ref = orc.accumulate_tail_array(line, 2 * T + 1, n, alpha)
            out = interpolation._accumulate_tail_array(line, None, L, 2 * T + 1, n, alpha)
            np.testing.assert_allclose(out, ref, rtol=4e-7 * n, atol=0)
    finally:
        interpolation._precision_private = 'fp64'


# ---------------------------------------------------------------------------
# NaN values

@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
def test_nan_value_matches_oracle(orc, precision):
    from fastbarnes import workloads
    c = workloads.config1()
    for idx in (1, 17, 2000, len(c['val']) - 1):
        val = c['val'].copy()
        val[idx] = np.nan
        cc = dict(c, val=val)
        out = _run(cc, precision=precision)
        ref, _, _ = _oracle(orc, cc)
        assert np.isnan(ref).all()
        assert np.isnan(out).all(), idx


# ---------------------------------------------------------------------------
# the uneven batch of 256 fields

@pytest.fixture(scope='module')
def uneven_batch():
    from fastbarnes import workloads
    return [workloads.config3_field(i) for i in range(256)]


def test_uneven_batch_256(orc, uneven_batch):
    from fastbarnes import hip
    from fastbarnes.batched import BarnesBatch
    from fastbarnes.interpolation import barnes
    fs = uneven_batch
    batch = BarnesBatch([_desc(c) for c in fs], precision='fp32c')
    pts = np.concatenate([c['pts'] for c in fs])
    val = np.concatenate([c['val'] for c in fs])
    d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
    d_out = hip.DeviceArray((batch.total_out,), np.float32)
    s = hip.Stream()
    # untuned: the cost model's passes, the same as each single-field plan's
    batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
    s.synchronize()
    flat = d_out.numpy()
    outs = [flat[o:o + int(np.prod(sh))].reshape(sh) for o, sh in zip(batch.out_offsets, batch.shapes)]
    for i, (c, o) in enumerate(zip(fs, outs)):
        single = barnes(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], num_iter=c['num_iter'],
                        max_dist=c['max_dist'], precision='fp32c')
        assert_bitwise(o, single)
    print('256 fields equal their single-field runs', flush=True)
    # oracle sample: the largest N, the first 2400x1200 field, sigma near 2 with n = 6, and spread
    ns = [len(c['val']) for c in fs]
    big = [i for i, c in enumerate(fs) if c['size'][0] == 2400]
    s2n6 = [i for i, c in enumerate(fs) if c['sigma'] > 1.9 and c['num_iter'] == 6 and c['size'][0] == 2400]
    sample = sorted({int(np.argmax(ns)), int(np.argmin(ns)), big[0], s2n6[0], 3, 64, 128, 200})
    assert len(sample) >= 8
    for i in sample:
        _check_fast(orc, fs[i], outs[i], 'uneven field %d %s sigma %.3f n %d N %d' % (
            i, fs[i]['size'], fs[i]['sigma'], fs[i]['num_iter'], ns[i]))
    # autotuned, graph replays: equal to the tuned plan's eager run, within the bar of the untuned run
    batch.autotune(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
    batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
    s.synchronize()
    eager = d_out.numpy()
    for _ in range(3):
        d_out.upload(np.full(batch.total_out, 7.0, dtype=np.float32))
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=True)
        s.synchronize()
        assert_bitwise(d_out.numpy(), eager)
    for i in sample:
        o, sh = batch.out_offsets[i], batch.shapes[i]
        tuned = eager[o:o + int(np.prod(sh))].reshape(sh)
        _check_fast(orc, fs[i], tuned, 'tuned uneven field %d' % i)
    batch.close()


# ---------------------------------------------------------------------------
# 3D

def test_3d_512x512x128(orc):
    from fastbarnes import workloads
    c = workloads.config4()
    out64 = _run(c, precision='fp64')
    ref, vg, wg = _oracle(orc, c)
    assert_bitwise(out64, ref)
    out32 = _run(c, precision='fp32c')
    rmse, maxabs, mism_far, mism = fast_errors(out32, ref, wg, _thr(orc, c))
    print('3D 512x512x128 fp32c: rmse %.3g maxabs %.3g mask %d' % (rmse, maxabs, mism), flush=True)
    assert mism_far == 0 and rmse <= RMSE_TOL and maxabs <= MAXABS_TOL


def test_3d_batch_32(orc):
    from fastbarnes import workloads
    from fastbarnes.batched import BarnesBatch
    rng = np.random.default_rng(44)
    fs = [workloads.config4(seed=400 + i, n=int(round(np.exp(rng.uniform(np.log(1e4), np.log(2e5))))),
                            size=(256, 256, 64)) for i in range(32)]
    ns = [len(c['val']) for c in fs]
    sample = sorted({int(np.argmax(ns)), int(np.argmin(ns)), 13})
    outs = {}
    for prec in ('fp64', 'fp32c'):
        batch = BarnesBatch([_desc(c) for c in fs], precision=prec)
        outs[prec] = batch.interpolate([c['pts'] for c in fs], [c['val'] for c in fs])
        batch.close()
    for i in sample:
        ref, vg, wg = _oracle(orc, fs[i])
        assert_bitwise(outs['fp64'][i], ref)
        rmse, maxabs, mism_far, mism = fast_errors(outs['fp32c'][i], ref, wg, _thr(orc, fs[i]))
        print('3D batch field %d N %d: rmse %.3g maxabs %.3g mask %d' % (i, ns[i], rmse, maxabs, mism), flush=True)
        assert mism_far == 0 and rmse <= RMSE_TOL and maxabs <= MAXABS_TOL


# ---------------------------------------------------------------------------
# plan lifetime

def test_destroy_after_async_execute():
    """A plan destroyed right after an asynchronous execute waits for its kernels before its
    buffers go back to the pool: a second plan of the same shape, which takes those buffers at
    once and runs other inputs on another stream, cannot change the first plan's result."""
    from fastbarnes import hip, workloads
    from fastbarnes.batched import BarnesBatch
    fs = [workloads.config2(seed=900 + i) for i in range(8)]
    descs = [_desc(c) for c in fs]
    pts = np.concatenate([c['pts'] for c in fs])
    val = np.concatenate([c['val'] for c in fs])
    d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
    d_val2 = hip.DeviceArray.from_numpy(val * -3.0 + 100.0)
    ref_batch = BarnesBatch(descs, precision='fp32c')
    d_ref = hip.DeviceArray((ref_batch.total_out,), np.float32)
    s0 = hip.Stream()
    ref_batch.execute(d_pts.ptr, d_val.ptr, d_ref.ptr, s0.ptr)
    s0.synchronize()
    ref_batch.close()
    ref = d_ref.numpy()
    s1, s2 = hip.Stream(), hip.Stream()
    for _ in range(3):
        a = BarnesBatch(descs, precision='fp32c')
        d_a = hip.DeviceArray((a.total_out,), np.float32)
        a.execute(d_pts.ptr, d_val.ptr, d_a.ptr, s1.ptr)
        a.close()   # returns immediately only after a's kernels end
        b = BarnesBatch(descs, precision='fp32c')
        d_b = hip.DeviceArray((b.total_out,), np.float32)
        b.execute(d_pts.ptr, d_val2.ptr, d_b.ptr, s2.ptr)
        s1.synchronize()
        s2.synchronize()
        b.close()
        out = d_a.numpy()
        np.testing.assert_array_equal(np.isnan(out), np.isnan(ref))
        ok = ~np.isnan(ref)
        np.testing.assert_array_equal(out[ok].view(np.uint32), ref[ok].view(np.uint32))


# ---------------------------------------------------------------------------
# FP32C lines longer than one line pass takes: chunks with num_iter*(T+1) halos

@pytest.mark.parametrize('W,H,transpose', [(20000, 24, False), (30011, 12, False), (16384, 20, True)])
def test_fp32c_chunked_lines(orc, W, H, transpose):
    """x lines (first pass, injecting) or y lines (last pass) beyond the 256-lane x 49-element
    limit run as chunks; stored nodes depend only on inputs within the halo, so the result
    meets the FP32C bar against the oracle everywhere, chunk seams included."""
    c = _long_line_case(W, H, W + H, transpose)
    out = _run(c, precision='fp32c')
    ref = _check_fast(orc, c, out, 'chunked lines %s' % (c['size'],))
    assert np.isfinite(ref).mean() > 0.5


def test_fp32c_chunked_lines_wide_kernel(orc):
    """A large sigma (half kernel 200, halo 4*201 = 804 per side) on 14000-node lines."""
    rng = np.random.default_rng(5)
    W, H = 14000, 12
    step = 0.01
    ext = np.array([W - 1, H - 1]) * step
    n = 3000
    pts = rng.uniform(0.0, 1.0, (n, 2)) * ext
    val = 3.0 + np.sin(pts[:, 0]) + rng.normal(0.0, 0.2, n)
    c = dict(pts=pts, val=val, sigma=[2.0, 0.04], x0=np.zeros(2), step=step, size=(W, H), num_iter=4,
             max_dist=3.5)
    out = _run(c, precision='fp32c')
    _check_fast(orc, c, out, 'chunked wide kernel')


def test_fp32c_chunked_1d_and_3d(orc):
    rng = np.random.default_rng(6)
    # 1D: one 40000-node line (first and last pass at once)
    L = 40000
    pts = rng.uniform(0.0, (L - 1) * 0.01, (6000, 1))
    val = np.cos(pts[:, 0]) + rng.normal(0.0, 0.1, len(pts))
    c1 = dict(pts=pts, val=val, sigma=0.05, x0=np.zeros(1), step=0.01, size=(L,), num_iter=4, max_dist=3.5)
    _check_fast(orc, c1, _run(c1, precision='fp32c'), 'chunked 1D')
    # 3D: long x lines (the injecting pass)
    size = (13000, 7, 8)
    ext = (np.array(size) - 1) * 0.01
    pts = rng.uniform(0.0, 1.0, (20000, 3)) * ext
    val = np.sin(pts[:, 0] * 3) + pts[:, 1] + rng.normal(0.0, 0.1, len(pts))
    c3 = dict(pts=pts, val=val, sigma=0.03, x0=np.zeros(3), step=0.01, size=size, num_iter=3, max_dist=3.5)
    _check_fast(orc, c3, _run(c3, precision='fp32c'), 'chunked 3D x')


def test_fp32c_chunked_limits():
    """Chunked lines along the middle axis of a 3D grid would be filtered in place and are refused
    with a clear message (precision fp64 takes them)."""
    c = dict(pts=np.zeros((4, 3)) + 0.05, val=np.ones(4), sigma=0.03, x0=np.zeros(3), step=0.01,
             size=(8, 13000, 6), num_iter=3, max_dist=3.5)
    with pytest.raises(RuntimeError, match='fp64'):
        _run(c, precision='fp32c')
