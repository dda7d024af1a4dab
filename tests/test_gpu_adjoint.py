"""
GPU tests of the adjoint (marked `gpu`): fb_plan_forward / fb_plan_backward through the
C ABI and the torch.autograd.Function on top of them, against the FP64 oracle adjoint
(tests/test_oracle_adjoint.py pins that one with central finite differences).

Tolerances: the FP32C path stores the fields between passes in float32.  Value gradients
sum_c w_c A[c] agree with the FP64 adjoint to a relative L2 error of 1e-6 and a maximum
error of 1e-5 of the largest component (measured: ~3e-8 and ~7e-8).  Position gradients use (val_k - offset) A[c] +
Bb[c], two terms that cancel to the station's departure from the interpolated field (a
factor ~20 for the smooth test fields), so their bounds are 5e-5 (relative L2) and 2e-4
(maximum); measured at 2400x1200, N=100000: 1.2e-5 and 2.2e-5.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_VAL = (1e-6, 1e-5)    # (relative L2, maximum relative to the largest component)
TOL_PTS = (5e-5, 2e-4)


def _desc(c):
    return dict(sigma=c['sigma'], x0=c['x0'], step=c['step'], size=c['size'], n=len(c['val']),
                num_iter=c['num_iter'], max_dist=c['max_dist'])


def _run(batch, pts, val, gfun):
    """forward + backward on device buffers; gfun(out) -> dL/dout (float32)."""
    from fastbarnes import hip
    d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
    d_out = hip.DeviceArray((batch.total_out,), np.float32)
    d_ws = hip.DeviceArray((batch.total_out,), np.float32)
    s = hip.Stream()
    batch.forward(d_pts.ptr, d_val.ptr, d_out.ptr, d_ws.ptr, s.ptr)
    s.synchronize()
    out = d_out.numpy()
    d_g = hip.DeviceArray.from_numpy(gfun(out).astype(np.float32))
    d_gval = hip.DeviceArray((len(val),), np.float64)
    d_gpts = hip.DeviceArray(pts.shape, np.float64)
    batch.backward(d_pts.ptr, d_val.ptr, d_out.ptr, d_ws.ptr, d_g.ptr, d_gval.ptr, d_gpts.ptr, s.ptr)
    s.synchronize()
    return out, d_gval.numpy(), d_gpts.numpy()


def _close(got, ref, tol):
    l2 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    mx = np.abs(got - ref).max() / np.abs(ref).max()
    print('relative L2 %.3e, max %.3e' % (l2, mx))
    assert l2 <= tol[0] and mx <= tol[1], (l2, mx)


def _check_against_oracle(c):
    from fastbarnes.batched import BarnesBatch
    from oracle import oracle as orc
    batch = BarnesBatch([_desc(c)], precision='fp32c')
    shape = tuple(c['size'][::-1])
    target = c.get('target')
    if target is None:
        target = np.full(shape, float(np.mean(c['val'])))

    def gfun(out):
        return np.where(np.isnan(out), 0.0, out - target.ravel())

    out, gval, gpts = _run(batch, c['pts'], c['val'], gfun)
    g = np.where(np.isnan(out), np.nan, out - target.ravel()).reshape(shape)
    rval, rpts = orc.adjoint(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], g,
                             num_iter=c['num_iter'], max_dist=c['max_dist'])
    _close(gval, rval, TOL_VAL)
    _close(gpts, rpts.reshape(gpts.shape), TOL_PTS)


def test_adjoint_config1_matches_oracle():
    from fastbarnes import workloads
    _check_against_oracle(workloads.config1())


def test_adjoint_config5_single_field_matches_oracle():
    """2400x1200, N=100000, sigma=1, n=4 against the FP64 reference gradient."""
    from fastbarnes import workloads
    _check_against_oracle(workloads.config5())


def test_adjoint_small_1d_3d():
    from fastbarnes import workloads
    rng = np.random.default_rng(3)
    p1 = rng.uniform(0.0, 20.0, (300, 1))
    c1 = dict(pts=p1, val=np.sin(p1[:, 0]) + 3.0, sigma=0.7, x0=np.array([0.0]), step=0.05, size=(401,),
              num_iter=4, max_dist=3.5)
    _check_against_oracle(c1)
    _check_against_oracle(workloads.config4(seed=9, n=4000, size=(48, 40, 24)))


def test_batch_gradients_equal_single_plans():
    """Gradients of each field of an uneven batch equal those of a one-field plan, bit for bit."""
    from fastbarnes import workloads
    from fastbarnes.batched import BarnesBatch
    fs = [workloads.config3_field(i) for i in range(3)]
    for c in fs:   # keep the test small
        W = min(c['size'][0], 600)
        c['size'] = (W, W // 2)
        c['step'] = 75.0 / W
    batch = BarnesBatch([_desc(c) for c in fs], precision='fp32c')
    pts = np.concatenate([c['pts'] for c in fs])
    val = np.concatenate([c['val'] for c in fs])

    def gfun(out):
        return np.where(np.isnan(out), 0.0, np.cos(np.arange(out.size) * 1e-3))

    _, gval, gpts = _run(batch, pts, val, gfun)
    for f, c in enumerate(fs):
        one = BarnesBatch([_desc(c)], precision='fp32c')
        o0, o1 = batch.out_offsets[f], batch.out_offsets[f] + c['size'][0] * c['size'][1]

        def gf(out, o0=o0, o1=o1):
            return np.where(np.isnan(out), 0.0, np.cos(np.arange(o0, o1) * 1e-3))

        _, gv1, gp1 = _run(one, c['pts'], c['val'], gf)
        p0, p1 = batch.pt_offsets[f], batch.pt_offsets[f] + len(c['val'])
        np.testing.assert_array_equal(gval[p0:p1], gv1)
        np.testing.assert_array_equal(gpts[p0:p1], gp1)


@pytest.mark.parametrize('n', [5000, 30000])
def test_backward_after_another_forward(n):
    """The backward depends only on its own arguments, not on the plan's last run: after a
    forward of other points on the same plan (its sorted lists, value offsets and planes then
    belong to those) it gives the same bits as right after its own forward, for plans that
    sort with the field sort (n = 5000) and with row buckets (n = 30000)."""
    from fastbarnes import hip, workloads
    from fastbarnes.batched import BarnesBatch
    rng = np.random.default_rng(n)
    c = workloads.config5(seed=71, n=n, size=(600, 300))
    c['step'] = 75.0 / 600
    other = c['pts'] + rng.normal(0.0, 3.0, c['pts'].shape)   # other positions, some outside
    batch = BarnesBatch([_desc(c)], precision='fp32c')

    def gfun(out):
        return np.where(np.isnan(out), 0.0, np.sin(np.arange(out.size) * 7e-4))

    out, gval, gpts = _run(batch, c['pts'], c['val'], gfun)
    d_pts, d_val = hip.DeviceArray.from_numpy(c['pts']), hip.DeviceArray.from_numpy(c['val'])
    d_oth = hip.DeviceArray.from_numpy(other)
    d_out = hip.DeviceArray((batch.total_out,), np.float32)
    d_ws = hip.DeviceArray((batch.total_out,), np.float32)
    d_o2 = hip.DeviceArray((batch.total_out,), np.float32)
    d_w2 = hip.DeviceArray((batch.total_out,), np.float32)
    s = hip.Stream()
    batch.forward(d_pts.ptr, d_val.ptr, d_out.ptr, d_ws.ptr, s.ptr)
    batch.forward(d_oth.ptr, d_val.ptr, d_o2.ptr, d_w2.ptr, s.ptr)   # the plan's last sort: other points
    s.synchronize()
    np.testing.assert_array_equal(d_out.numpy().view(np.uint32), out.view(np.uint32))
    d_g = hip.DeviceArray.from_numpy(gfun(out).astype(np.float32))
    d_gval = hip.DeviceArray((n,), np.float64)
    d_gpts = hip.DeviceArray(c['pts'].shape, np.float64)
    batch.backward(d_pts.ptr, d_val.ptr, d_out.ptr, d_ws.ptr, d_g.ptr, d_gval.ptr, d_gpts.ptr, s.ptr)
    s.synchronize()
    np.testing.assert_array_equal(d_gval.numpy(), gval)
    np.testing.assert_array_equal(d_gpts.numpy(), gpts)
    assert np.count_nonzero(gval) > n // 2


def test_torch_autograd_function():
    torch = pytest.importorskip('torch')
    from fastbarnes import workloads
    from fastbarnes.autograd import barnes_torch
    from fastbarnes.batched import BarnesBatch
    from oracle import oracle as orc
    c = workloads.config5(n=20000, size=(600, 300))
    batch = BarnesBatch([_desc(c)], precision='fp32c')
    dev = torch.device('cuda:0')
    pts = torch.tensor(c['pts'], device=dev, requires_grad=True)
    val = torch.tensor(c['val'], device=dev, requires_grad=True)
    target = torch.tensor(c['target'].ravel(), dtype=torch.float32, device=dev)
    out = barnes_torch(pts, val, batch)
    ok = ~torch.isnan(out)
    loss = 0.5 * ((out - target)[ok].double() ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    o = out.detach().cpu().numpy()
    g = np.where(np.isnan(o), np.nan, o - c['target'].ravel().astype(np.float32)).reshape(300, 600)
    rval, rpts = orc.adjoint(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], g)
    _close(val.grad.cpu().numpy(), rval, TOL_VAL)
    _close(pts.grad.cpu().numpy(), rpts, TOL_PTS)
    # the forward value equals the plain batch execution
    ref = batch.interpolate([c['pts']], [c['val']])[0]
    np.testing.assert_array_equal(np.isnan(o.reshape(ref.shape)), np.isnan(ref))
    np.testing.assert_array_equal(np.nan_to_num(o.reshape(ref.shape)), np.nan_to_num(ref))


def test_backward_rejects_fp64_plan():
    from fastbarnes.batched import BarnesBatch
    from fastbarnes import workloads
    c = workloads.config1()
    batch = BarnesBatch([_desc(c)], precision='fp64')
    with pytest.raises(RuntimeError, match='FP32C'):
        _run(batch, c['pts'], c['val'], lambda o: np.zeros_like(o))
