"""
GPU tests of batched plans (marked `gpu`): uneven fields in one plan, eager and hipGraph
replays.  Each field of a batch must equal the single-field drop-in result bit for bit
(both precisions use the same kernels and per-field work decomposition), and the
single-field FP64 results equal the oracle (tests/test_gpu_parity.py).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _fields(seed=0):
    from fastbarnes import workloads
    rng = np.random.default_rng(seed)
    out = []
    for i in range(5):
        W = int(rng.choice([150, 300, 600]))
        H = W // 2
        n = int(rng.integers(200, 4000))
        pts = workloads.stations_2d(n, seed * 100 + i)
        out.append(dict(pts=pts, val=workloads.values_sinusoids(pts, i), sigma=float(rng.uniform(0.5, 2.0)),
                        x0=np.array([-26.0, 34.0]), step=75.0 / W, size=(W, H), num_iter=int(rng.integers(3, 7)),
                        max_dist=3.5))
    return out


def _desc(c):
    return dict(sigma=c['sigma'], x0=c['x0'], step=c['step'], size=c['size'], n=len(c['val']),
                num_iter=c['num_iter'], max_dist=c['max_dist'])


def _bitwise(a, b):
    np.testing.assert_array_equal(np.isnan(a), np.isnan(b))
    ok = ~np.isnan(a)
    np.testing.assert_array_equal(a[ok].view(np.uint32), b[ok].view(np.uint32))


@pytest.mark.parametrize('precision', ['fp64', 'fp32c'])
def test_batch_equals_single_calls(precision):
    from fastbarnes.batched import BarnesBatch
    from fastbarnes.interpolation import barnes
    fs = _fields()
    batch = BarnesBatch([_desc(c) for c in fs], precision=precision)
    outs = batch.interpolate([c['pts'] for c in fs], [c['val'] for c in fs])
    for c, o in zip(fs, outs):
        ref = barnes(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], num_iter=c['num_iter'],
                     max_dist=c['max_dist'], precision=precision)
        _bitwise(o, ref)
    info = batch.info()
    assert info['grid_points'] == sum(c['size'][0] * c['size'][1] for c in fs)


def test_fp64_x_line_workgroup_sizes_agree(orc):
    """The FP64 x lines run through an LDS ring, one wave per 16, 32 or 64 (line, field) slots:
    64 for plans of at least 16384 slots, 32 from 8192, 16 below.  A batch of 7 fields of
    1200x1200 (16800 slots: 64-slot workgroups, the last of each field partial) and one of 4 of
    them (9600 slots: 32) must equal the single-field runs (2400 slots: 16) bit for bit, with odd
    and even iteration counts and both methods, and a field must equal the oracle."""
    from fastbarnes import workloads
    from fastbarnes.batched import BarnesBatch
    from fastbarnes.interpolation import barnes
    rng = np.random.default_rng(5)
    fs = []
    for i in range(7):
        pts = workloads.stations_2d(int(rng.integers(1000, 3000)), 700 + i)
        fs.append(dict(pts=pts, val=workloads.values_sinusoids(pts, i), sigma=float(rng.uniform(0.6, 1.6)),
                       x0=np.array([-26.0, 34.0]), step=75.0 / 1200, size=(1200, 1200), num_iter=3 + i % 4,
                       max_dist=3.5, method='convolution' if i == 3 else 'optimized_convolution'))
    descs = [dict(_desc(c), method=c['method']) for c in fs]
    outs = BarnesBatch(descs, precision='fp64').interpolate([c['pts'] for c in fs], [c['val'] for c in fs])
    outs4 = BarnesBatch(descs[:4], precision='fp64').interpolate([c['pts'] for c in fs[:4]],
                                                                  [c['val'] for c in fs[:4]])
    for i, (c, o) in enumerate(zip(fs, outs)):
        ref = barnes(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], method=c['method'],
                     num_iter=c['num_iter'], max_dist=c['max_dist'], precision='fp64')
        _bitwise(o, ref)
        if i < 4:
            _bitwise(outs4[i], ref)
    c = fs[1]
    ref = orc.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], num_iter=c['num_iter'],
                                 max_dist=c['max_dist'])
    _bitwise(outs[1], ref)


def test_graph_replays_match_eager():
    from fastbarnes import hip
    from fastbarnes.batched import BarnesBatch
    fs = _fields(1)
    batch = BarnesBatch([_desc(c) for c in fs], precision='fp32c')
    pts = np.concatenate([c['pts'] for c in fs])
    val = np.concatenate([c['val'] for c in fs])
    d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
    d_out = hip.DeviceArray((batch.total_out,), np.float32)
    s = hip.Stream()
    batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=False)
    s.synchronize()
    eager = d_out.numpy()
    for _ in range(4):
        d_out.upload(np.full(batch.total_out, 12345.0, dtype=np.float32))   # stale data must be overwritten
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=True)
        s.synchronize()
        _bitwise(d_out.numpy(), eager)
    # new values through the same captured graph (buffers reused in place)
    d_val.upload(val * 2.0 - 1.0)
    batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=True)
    s.synchronize()
    g2 = d_out.numpy()
    batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=False)
    s.synchronize()
    _bitwise(g2, d_out.numpy())


def test_stage_timing():
    from fastbarnes import hip
    from fastbarnes.batched import BarnesBatch
    fs = _fields(2)[:2]
    batch = BarnesBatch([_desc(c) for c in fs], precision='fp32c')
    pts = np.concatenate([c['pts'] for c in fs])
    val = np.concatenate([c['val'] for c in fs])
    d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
    d_out = hip.DeviceArray((batch.total_out,), np.float32)
    s = hip.Stream()
    batch.set_timing(True)
    batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
    t = batch.stage_times()
    assert list(t) == ['sort', 'pass_x', 'pass_y'] and all(v > 0 for v in t.values())


def test_batch_3d_fp64_equals_single():
    from fastbarnes import workloads
    from fastbarnes.batched import BarnesBatch
    from fastbarnes.interpolation import barnes
    fs = [workloads.config4(seed=s, n=5000 * (s + 1), size=(64, 48, 24)) for s in range(3)]
    for prec in ('fp64', 'fp32c'):
        batch = BarnesBatch([_desc(c) for c in fs], precision=prec)
        outs = batch.interpolate([c['pts'] for c in fs], [c['val'] for c in fs])
        for c, o in zip(fs, outs):
            ref = barnes(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], num_iter=c['num_iter'],
                         max_dist=c['max_dist'], precision=prec)
            _bitwise(o, ref)


def test_autotune_keeps_results_and_reports_groups():
    """The autotuner times the segment-length variants of every line pass on the plan's own
    buffers and rebuilds the passes; results stay within FP32C rounding of the untuned plan
    (different segment anchors change only the last bits), execution stays deterministic and
    graph replays are rebuilt."""
    from fastbarnes import hip, workloads
    from fastbarnes.batched import BarnesBatch
    fs = [workloads.config3_field(i) for i in range(6)] + [workloads.config4(seed=7, n=3000, size=(64, 48, 24))]
    for dim_group in ([c for c in fs if len(c['size']) == 2], [c for c in fs if len(c['size']) == 3]):
        batch = BarnesBatch([_desc(c) for c in dim_group], precision='fp32c')
        pts = np.concatenate([c['pts'] for c in dim_group])
        val = np.concatenate([c['val'] for c in dim_group])
        d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
        d_out = hip.DeviceArray((batch.total_out,), np.float32)
        s = hip.Stream()
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
        s.synchronize()
        before = d_out.numpy()
        batch.autotune(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
        for axis in range(len(dim_group[0]['size'])):
            groups = batch.pass_groups(axis)
            segs = {'slide': (19, 21, 23, 25, 29, 33, 37, 39, 41, 45, 49), 'prefix': (15, 17, 19, 21, 23)}
            assert groups and all(seg in segs[kind] and thr in (128, 256, 512) and n > 0
                                  for kind, seg, thr, n in groups)
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=True)
        s.synchronize()
        after = d_out.numpy()
        np.testing.assert_array_equal(np.isnan(before), np.isnan(after))
        ok = ~np.isnan(before)
        assert np.max(np.abs(after[ok] - before[ok]) / (np.abs(before[ok]) + 1.0)) < 1e-5
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=False)
        s.synchronize()
        _bitwise(d_out.numpy(), after)


def test_autotune_cache_round_trip(tmp_path):
    """A plan tuned with a cache file stores one entry per line-length class; a new plan of the
    same batch takes every choice from the file without timing ('cache'), and runs the same
    passes to the same bits.  Other stations on the same grids hit the same entries; an entry
    that no longer fits (a segment length with no instantiation) makes the plan time afresh."""
    import json
    from fastbarnes import hip, workloads
    from fastbarnes.batched import BarnesBatch
    fs = [workloads.config3_field(i) for i in range(4)]
    path = tmp_path / 'tune.json'

    def run(fields, cache):
        batch = BarnesBatch([_desc(c) for c in fields], precision='fp32c')
        pts = np.concatenate([c['pts'] for c in fields])
        val = np.concatenate([c['val'] for c in fields])
        d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
        d_out = hip.DeviceArray((batch.total_out,), np.float32)
        s = hip.Stream()
        how = batch.autotune(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, cache=cache)
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=True)
        s.synchronize()
        return how, [batch.pass_choices(a)[0] for a in range(2)], d_out.numpy()

    how1, ch1, out1 = run(fs, str(path))
    assert how1 == 'tuned'
    entries = json.loads(path.read_text())['entries']
    assert len(entries) == sum(len(c) for c in ch1)
    how2, ch2, out2 = run(fs, str(path))
    assert how2 == 'cache' and ch2 == ch1
    _bitwise(out2, out1)
    # the same grids with other stations: the key has no point counts
    fs3 = [dict(c, pts=c['pts'][: len(c['pts']) // 2], val=c['val'][: len(c['val']) // 2]) for c in fs]
    how3, ch3, _ = run(fs3, str(path))
    assert how3 == 'cache' and ch3 == ch1
    # a stale entry: time afresh and rewrite it
    d = json.loads(path.read_text())
    k0 = sorted(d['entries'])[0]
    d['entries'][k0] = dict(kind='slide', seg=7, threads=256)
    path.write_text(json.dumps(d))
    how4, ch4, _ = run(fs, str(path))
    assert how4 == 'tuned'
    assert json.loads(path.read_text())['entries'][k0]['seg'] != 7


def test_full_size_batch_matches_single_fields():
    """Full-size fields (2400x1200 in 2D, 256x256x64 in 3D) in one batch: every stage runs once
    over all fields with a work list far larger than the GPU, and every field equals its
    single-field run bit for bit."""
    from fastbarnes import workloads
    from fastbarnes.batched import BarnesBatch
    from fastbarnes.interpolation import barnes
    fs = [workloads.config2(seed=90 + i) for i in range(3)]
    fs += [workloads.config4(seed=95 + i, n=20000, size=(256, 256, 64)) for i in range(4)]
    for group in (fs[:3], fs[3:]):
        batch = BarnesBatch([_desc(c) for c in group], precision='fp32c')
        outs = batch.interpolate([c['pts'] for c in group], [c['val'] for c in group])
        for c, o in zip(group, outs):
            ref = barnes(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], num_iter=c['num_iter'],
                         max_dist=c['max_dist'], precision='fp32c')
            _bitwise(o, ref)


@pytest.mark.parametrize('seg', [15, 19, 23])
def test_prefix_pass_matches_fp64(seg):
    """The prefix pass (FP64 prefix sums of the lines in LDS; an autotuner candidate), forced
    on every axis through choose_pass: 2D fields with lines of one wave (L <= 64*seg), of
    several waves and with a partial last lane, a 3D field and a 1D field, against the FP64
    path (same NaN mask, RMSE <= 1e-5 of the value range) and the sliding pass (within
    float32 rounding of each other); deterministic across runs and graph replays."""
    from fastbarnes import hip, workloads
    from fastbarnes.batched import BarnesBatch
    rng = np.random.default_rng(seg)
    cases = []
    for W, n, sigma, it in ((2400, 3490, 1.0, 4), (600, 2000, 1.7, 6), (150, 700, 0.9, 3)):
        pts = workloads.stations_2d(n, W + seg)
        cases.append(dict(pts=pts, val=workloads.values_sinusoids(pts, W), sigma=sigma, x0=np.array([-26.0, 34.0]),
                          step=75.0 / W, size=(W, W // 2), num_iter=it, max_dist=3.5))
    c3 = workloads.config4(seed=seg, n=3000, size=(64, 48, 24))
    cases.append(dict(pts=c3['pts'], val=c3['val'], sigma=c3['sigma'], x0=c3['x0'], step=c3['step'],
                      size=c3['size'], num_iter=4, max_dist=3.5))
    p1 = np.sort(rng.uniform(0.0, 30.0, 300))
    cases.append(dict(pts=p1, val=np.sin(p1 / 3.0) + rng.normal(0, 0.1, 300), sigma=0.8, x0=0.0, step=0.01,
                      size=(3000,), num_iter=5, max_dist=3.5))
    for c in cases:
        dim = len(c['size'])
        pts2 = c['pts'].reshape(len(c['val']), -1)
        ref64 = BarnesBatch([_desc(c)], precision='fp64').interpolate([pts2], [c['val']])[0]
        slide = BarnesBatch([_desc(c)], precision='fp32c').interpolate([pts2], [c['val']])[0]
        batch = BarnesBatch([_desc(c)], precision='fp32c')
        for a in range(dim):
            try:
                batch.choose_pass(a, c['size'][a], 'prefix', seg, 256)
            except RuntimeError:   # lines of more than 128 lanes: 512-thread workgroups
                batch.choose_pass(a, c['size'][a], 'prefix', seg, 512)
            assert all(k == 'prefix' and sg == seg for k, sg, _, _ in batch.pass_groups(a))
        d_pts, d_val = hip.DeviceArray.from_numpy(np.ascontiguousarray(pts2)), hip.DeviceArray.from_numpy(c['val'])
        d_out = hip.DeviceArray((batch.total_out,), np.float32)
        s = hip.Stream()
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
        s.synchronize()
        r = d_out.numpy().reshape(ref64.shape)
        for graph in (True, True, False):
            batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=graph)
            s.synchronize()
            _bitwise(d_out.numpy().reshape(ref64.shape), r)
        np.testing.assert_array_equal(np.isnan(r), np.isnan(ref64))
        ok = ~np.isnan(ref64)
        assert ok.mean() > 0.2
        rng_v = float(np.ptp(c['val']))
        d = r[ok].astype(np.float64) - ref64[ok]
        assert float(np.sqrt(np.mean(d * d))) <= 1e-5 * rng_v, (c['size'], seg)
        assert float(np.max(np.abs(r[ok] - slide[ok]) / (np.abs(slide[ok]) + 1.0))) < 1e-5


def test_blocked_intermediate_matches_row_major():
    """2D fields with W a multiple of 4 keep the intermediate between the x and y passes in
    blocks of 8 rows x 4 columns when their y tiles are 4k columns wide; set_blocked(False)
    keeps every field row-major with the same pass choices, so the same arithmetic: forward
    results, graph replays, the training forward's saved weights and the adjoint's gradients
    must be the same bits."""
    from fastbarnes import hip, workloads
    from fastbarnes.batched import BarnesBatch
    group = [workloads.config2(seed=150), workloads.config3_field(0), workloads.config2(seed=151)]
    group += [c for c in (workloads.config3_field(i) for i in range(1, 12)) if c['size'][1] in (300, 600)][:2]
    shapes = [tuple(c['size']) for c in group]
    assert (1200, 600) in shapes and (600, 300) in shapes, shapes
    # H not a multiple of 8 (the cell space is padded to whole 8-row blocks), and 8-column y
    # tiles (two float4 column groups per row: a small kernel fits 8 lines of 600 in LDS)
    for W, H, sigma, seed in ((2400, 1196, 1.0, 152), (1200, 600, 0.6, 153), (602, 300, 1.0, 154)):
        p = workloads.stations_2d(2500, seed)
        group.append(dict(pts=p, val=workloads.values_sinusoids(p, seed), sigma=sigma, x0=np.array([-26.0, 34.0]),
                          step=75.0 / W, size=(W, H), num_iter=4, max_dist=3.5))
    pts = np.concatenate([c['pts'] for c in group])
    val = np.concatenate([c['val'] for c in group])
    d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
    s = hip.Stream()
    res = {}
    for allow in (True, False):
        batch = BarnesBatch([_desc(c) for c in group], precision='fp32c')
        for H in sorted({c['size'][1] for c in group}):
            batch.choose_pass(1, H, 'slide', 39, 256)
        batch.set_blocked(allow)
        blocked = [batch.intermediate_blocked(f) for f in range(len(group))]
        if allow:
            # the 2400x1200 and 2400x1196 fields (4-column y tiles) and the small-kernel 1200x600
            # field (8-column tiles) are blocked; a field whose y tiles hold a number of columns
            # that is not a multiple of 4 (7 lines of a large kernel fit the LDS) stays
            # row-major, as does the 602-wide field (W not a multiple of 4)
            assert blocked[0] and blocked[2] and blocked[5] and blocked[6] and not blocked[7], blocked
        else:
            assert not any(blocked)
        d_out = hip.DeviceArray((batch.total_out,), np.float32)
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
        s.synchronize()
        out = d_out.numpy()
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=True)
        s.synchronize()
        _bitwise(d_out.numpy(), out)
        d_ws = hip.DeviceArray((batch.total_out,), np.float32)
        batch.forward(d_pts.ptr, d_val.ptr, d_out.ptr, d_ws.ptr, s.ptr)
        s.synchronize()
        fwd = d_out.numpy()
        g = np.nan_to_num(fwd - np.float32(0.5), nan=0.0).astype(np.float32)
        d_g = hip.DeviceArray.from_numpy(g)
        d_gval = hip.DeviceArray((len(val),), np.float64)
        d_gpts = hip.DeviceArray(pts.shape, np.float64)
        batch.backward(d_pts.ptr, d_val.ptr, d_out.ptr, d_ws.ptr, d_g.ptr, d_gval.ptr, d_gpts.ptr, s.ptr)
        s.synchronize()
        res[allow] = (out, fwd, d_ws.numpy(), d_gval.numpy(), d_gpts.numpy())
    assert np.mean(~np.isnan(res[True][0])) > 0.2
    for a, b in zip(res[True], res[False]):
        _bitwise(a, b)


def test_row_buckets_with_crowded_rows():
    """Fields of more than 8192 points are sorted by row buckets: a slot per point in its base
    row, then each row sorted by (cell, point index) -- by one thread (up to 16 entries), by a
    workgroup in LDS (up to 8192) or by rank (more).  Crowded rows exercise the last two: a
    field with 9000 points on one row and 1500 points on each of 3 rows, besides scattered
    ones, one with every point on one row, and a tall sparse field.  Rows of 17 to 256 entries
    are ranked by one wave in groups of 64 entries: a field whose rows hold 17 to 256 points,
    at every group count, odd and even.  The FP64 results must equal the oracle's bit for
    bit (the reference's per-cell point order), and the FP32C path must agree with FP64."""
    from fastbarnes.batched import BarnesBatch
    from oracle import oracle as orc
    rng = np.random.default_rng(77)
    W, H, step = 300, 150, 0.25
    x0 = np.array([-26.0, 34.0])

    def on_rows(rows, n_per_row):
        xs = rng.uniform(x0[0], x0[0] + (W - 1.2) * step, len(rows) * n_per_row)
        ys = np.repeat([x0[1] + (r + 0.5) * step for r in rows], n_per_row)
        return np.stack([xs, ys], axis=1)

    p1 = np.concatenate([on_rows([70], 9000), on_rows([10, 11, 140], 1500),
                         np.stack([rng.uniform(-26, 48, 2000), rng.uniform(34, 71, 2000)], axis=1)])
    p1 = p1[rng.permutation(len(p1))]
    p2 = on_rows([33], 12000)
    counts = [17, 63, 64, 65, 127, 128, 129, 191, 192, 193, 255, 256]
    p4 = np.concatenate([on_rows([20 + 9 * k], c) for k, c in enumerate(counts)] +
                        [np.stack([rng.uniform(-26, 48, 9000), rng.uniform(65.25, 71, 9000)], axis=1)])
    p4 = p4[rng.permutation(len(p4))]
    cases = []
    for p in (p1, p2, p4):
        cases.append(dict(pts=p, val=np.sin(p[:, 0] / 3.0) * 10 + p[:, 1], sigma=1.2, x0=x0, step=step,
                          size=(W, H), num_iter=4, max_dist=3.5))
    # a tall sparse field: most rows of the batch hold 0-1 points (the single-thread class)
    p3 = np.stack([rng.uniform(x0[0], x0[0] + 70, 300), rng.uniform(x0[1], x0[1] + 990, 300)], axis=1)
    cases.append(dict(pts=p3, val=np.cos(p3[:, 1] / 7.0), sigma=1.2, x0=x0, step=step, size=(W, 4000), num_iter=4,
                      max_dist=3.5))
    for prec in ('fp64', 'fp32c'):
        batch = BarnesBatch([_desc(c) for c in cases], precision=prec)
        outs = batch.interpolate([c['pts'] for c in cases], [c['val'] for c in cases])
        for c, o in zip(cases, outs):
            ref = orc.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'],
                                         num_iter=c['num_iter'], max_dist=c['max_dist'])
            if prec == 'fp64':
                _bitwise(o, ref)
            else:
                np.testing.assert_array_equal(np.isnan(o), np.isnan(ref))
                ok = ~np.isnan(ref)
                d = o[ok].astype(np.float64) - ref[ok]
                assert float(np.sqrt(np.mean(d * d))) <= 1e-5 * float(np.ptp(c['val']))


def test_field_sort_rank_and_bitonic_paths():
    """Fields of at most 8192 points are sorted in LDS: by rank within their rows when no row
    holds more than 64 points and the field has at most 4096 rows, else by a bitonic network.
    Fields at both sides of each limit: a row of exactly 64 points and one of 65 (with
    duplicate cells, so the point-index order of equal keys matters), 4096 and 4097 rows, and a
    1D field of 64 and 65 points (one row).  The FP64 results must equal the oracle's bit for
    bit (the per-cell point order)."""
    from fastbarnes.batched import BarnesBatch
    from oracle import oracle as orc
    rng = np.random.default_rng(91)
    step = 0.25
    x0 = np.array([-26.0, 34.0])
    cases = []
    for crowd in (64, 65):
        W, H = 120, 60
        xs = rng.uniform(x0[0], x0[0] + (W - 1.2) * step, crowd)
        xs[: crowd // 4] = xs[0]                      # equal cells on the crowded row
        crowded = np.stack([xs, np.full(crowd, x0[1] + 20.5 * step)], axis=1)
        sparse = np.stack([rng.uniform(x0[0], x0[0] + (W - 1) * step, 700),
                           rng.uniform(x0[1], x0[1] + (H - 1) * step, 700)], axis=1)
        sparse = sparse[np.floor((sparse[:, 1] - x0[1]) / step) != 20]   # the crowded row holds `crowd`
        p = np.concatenate([crowded, sparse])
        p = p[rng.permutation(len(p))]
        cases.append(dict(pts=p, val=np.sin(p[:, 0] / 3.0) * 10 + p[:, 1], sigma=1.1, x0=x0, step=step,
                          size=(W, H), num_iter=4, max_dist=3.5))
    for H in (4096, 4097):
        W = 40
        p = np.stack([rng.uniform(x0[0], x0[0] + (W - 1) * step, 900),
                      rng.uniform(x0[1], x0[1] + (H - 1) * step, 900)], axis=1)
        cases.append(dict(pts=p, val=np.cos(p[:, 1] / 7.0) + p[:, 0], sigma=1.1, x0=x0, step=step, size=(W, H),
                          num_iter=4, max_dist=3.5))
    cases1 = []
    for n in (64, 65):
        p = rng.uniform(-3.0, 3.0, (n, 1))
        p[:8, 0] = p[0, 0]
        cases1.append(dict(pts=p, val=np.sin(p[:, 0]) + rng.normal(0, 0.1, n), sigma=0.4, x0=-3.5, step=0.05,
                           size=(150,), num_iter=4, max_dist=3.5))
    for group in (cases, cases1):
        batch = BarnesBatch([_desc(c) for c in group], precision='fp64')
        outs = batch.interpolate([c['pts'] for c in group], [c['val'] for c in group])
        for c, o in zip(group, outs):
            pts = c['pts'] if len(c['size']) > 1 else c['pts'][:, 0]
            size = c['size'] if len(c['size']) > 1 else c['size'][0]
            ref = orc.interpolate_convol(pts, c['val'], c['sigma'], c['x0'], c['step'], size,
                                         num_iter=c['num_iter'], max_dist=c['max_dist'])
            _bitwise(o.reshape(ref.shape), ref)


def test_choose_pass_rejects_bad_choices():
    from fastbarnes.batched import BarnesBatch
    c = _fields(3)[0]
    batch = BarnesBatch([_desc(c)], precision='fp32c')
    with pytest.raises(RuntimeError, match='segment length'):
        batch.choose_pass(0, c['size'][0], 'prefix', 41, 256)
    with pytest.raises(RuntimeError, match='no field'):
        batch.choose_pass(0, c['size'][0] + 1, 'slide', 41, 256)
    with pytest.raises(RuntimeError, match='threads'):
        batch.choose_pass(0, c['size'][0], 'slide', 41, 64)
    with pytest.raises(RuntimeError, match='FP32C'):
        BarnesBatch([_desc(c)], precision='fp64').choose_pass(0, c['size'][0], 'slide', 41, 256)


def test_sort_paths_agree_bitwise():
    """Plans whose fields all have at most 8192 points sort each field in LDS (one
    workgroup per field); a plan with a larger field sorts by row buckets.  Both must give
    the same per-cell point order (the reference's injection order), hence the same FP64
    bits: a batch with a 9000-point field (row buckets) against single-field calls of its
    small fields (LDS sort), and the large field against the oracle."""
    from fastbarnes import workloads
    from fastbarnes.batched import BarnesBatch
    from fastbarnes.interpolation import barnes
    fs = _fields(5)[:3]
    pts = workloads.stations_2d(9000, 77)
    fs.append(dict(pts=pts, val=workloads.values_sinusoids(pts, 77), sigma=1.1, x0=np.array([-26.0, 34.0]),
                   step=0.125, size=(600, 300), num_iter=4, max_dist=3.5))
    batch = BarnesBatch([_desc(c) for c in fs], precision='fp64')
    outs = batch.interpolate([c['pts'] for c in fs], [c['val'] for c in fs])
    for c, o in zip(fs, outs):
        ref = barnes(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], num_iter=c['num_iter'],
                     max_dist=c['max_dist'], precision='fp64')
        _bitwise(o, ref)
    from oracle import oracle as orc
    c = fs[-1]
    _bitwise(outs[-1], orc.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'],
                                              num_iter=c['num_iter'], max_dist=c['max_dist']))
