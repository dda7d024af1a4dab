#!/usr/bin/env python3
"""
Throughput of batched FP32C fast Barnes interpolation on one MI355X.

Workload (one step): a batch of --fields fields (default 64), each a 2400x1200 grid
(step 1/32, x0=(-26,34)) with 3490 seeded stations (40 % in 6 Gaussian clusters, the
rest uniform over [-26,49]x[34,72]), values 15+10 sin(x/8) cos(y/6)+N(0,0.5), sigma=1,
num_iter=4, max_dist=3.5, optimized_convolution, FP32C.  Inputs are resident in HBM
before the timed region; each step replays the batch's captured hipGraph.

Printed JSON line: value = grid points per second of the whole batch; roofline = the
dominant kernel's algorithmic bytes over its average duration (HIP events on its
stream), with frac_moved at the bytes this build's kernels move (w kept in float32) and
frac_of_copy against this GPU's own copy bandwidth (a 1 GiB device-to-device copy, read +
write bytes, timed in the same run; copy_gbs);
cpu_baseline = the CPU oracle (FP64 C restatement of the reference path, -O3, OpenMP
over lines) on a bounded sample of the same fields, on all host cores and on one.  'workloads' adds the other
workloads, each with its throughput, pipeline roofline fraction and error against the
FP64 path: cfg3 (256 uneven 2D fields), cfg4 (3D 512x512x128) and cfg4_batch (32 3D
fields of 256x256x64), cfg5 (16 fields of 2400x1200, N=100000, forward + adjoint), real
(the RMSE-versus-iterations curve on the 3490 real stations of tests/golden/qff), rows
(method='convolution', barnes_S2, naive and radius, each beside the CPU oracle).

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--fields 64] [--cpu-seconds 10]
                    [--legs cfg3,cfg4,cfg5,real,rows|none] [--leg-steps 10]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, 'fast-barnes-py_amd')
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = 'grid-pts/s/GPU batched FP32 vs plain FP64 at RMSE≤1e-5; % HBM peak'
HBM_PEAK_GBS = 8000.0
# algorithmic bytes of the 2D FP32C pipeline (inputs f64, v f32, w f64, out f32):
#   x-pass: writes v, w (12 B per grid point) and reads pts, val (24 B per sample point)
#   y-pass: reads v, w (12 B) and writes the float32 result (4 B) per grid point
BYTES_PER_POINT = {'pass_x': 12.0, 'pass_y': 16.0}
BYTES_PER_SAMPLE = {'pass_x': 24.0, 'pass_y': 0.0}
# the bytes this build's passes move: w stays float32 between the passes, so the x pass writes
# 8 B per grid point and the y pass reads 8 and writes 4
MOVED_PER_POINT = {'pass_x': 8.0, 'pass_y': 12.0}


def log(msg):
    print('[bench] ' + msg, file=sys.stderr, flush=True)


def make_fields(nfields, seed0=100):
    from fastbarnes import workloads
    cfgs = [workloads.config2(seed=seed0 + i) for i in range(nfields)]
    descs = [dict(sigma=c['sigma'], x0=c['x0'], step=c['step'], size=c['size'], n=len(c['val']),
                  num_iter=c['num_iter'], max_dist=c['max_dist']) for c in cfgs]
    return cfgs, descs


def _cpu_model():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def _host_threads():
    """Cores this process may use: its CPU affinity, capped by OMP_NUM_THREADS when set."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get('OMP_NUM_THREADS', '')
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(n, 1)


def _time_oracle(cfgs, seconds, threads):
    """Whole fields of the batch through the oracle's OpenMP build on `threads` threads, one after
    another until `seconds` have passed (at least one field, at most the whole batch)."""
    from oracle import oracle
    done = pts_total = 0
    t0 = time.perf_counter()
    for c in cfgs:
        oracle.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'],
                                  num_iter=c['num_iter'], max_dist=c['max_dist'], threads=threads)
        done += 1
        pts_total += c['size'][0] * c['size'][1]
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return pts_total / dt, done, dt


def cpu_baseline(cfgs, seconds):
    """The oracle (the FP64 C restatement of the reference path, in its loop and operation order;
    built at -O3 for x86-64-v3 with OpenMP over the independent lines of each pass) on a bounded
    sample of the batch's fields: on all host cores (the value) and on one core (the
    reference's own sequential design)."""
    from oracle import oracle
    oracle.omp_lib()
    nthr = _host_threads()
    flags = 'gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off'
    one, n1, t1 = _time_oracle(cfgs, seconds / 2, 1)
    allc, na, ta = _time_oracle(cfgs, seconds / 2, nthr)
    model = _cpu_model()
    return dict(value=round(allc, 1), unit='grid-pts/s', cores=nthr, kind='port', cpu=model, flags=flags,
                sample='%d field(s) of 2400x1200, N=3490, n=4 (FP64 oracle, OpenMP over lines, %d threads), %.2f s'
                       % (na, nthr, ta),
                single_core=dict(value=round(one, 1), cores=1,
                                 sample='%d field(s), 1 thread, %.2f s' % (n1, t1)))


# algorithmic bytes per grid point / per sample point of the other workloads' pipelines
PIPE_BYTES = {2: (28.0, 24.0), 3: (52.0, 32.0)}
AUTOTUNE = True   # --no-autotune: keep the cost model's segment lengths
# training (2D): forward also keeps the weight field (+8 B/pt); backward reads g, f, W (16),
# round-trips two float32 fields (8 + 8), reads pts/val and writes the gradients (24 + 24)
TRAIN_BYTES = (36.0 + 32.0, 24.0 + 48.0)


def _descs(cfgs):
    return [dict(sigma=c['sigma'], x0=c['x0'], step=c['step'], size=c['size'], n=len(c['val']),
                 num_iter=c['num_iter'], max_dist=c['max_dist']) for c in cfgs]


def _time_steps(fn, stream, steps, warmup):
    from fastbarnes import hip
    for _ in range(warmup):
        fn()
    stream.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1) / steps


def _copy_bandwidth(stream, mib=1024, reps=10):
    """Achievable HBM bandwidth of this GPU: a device-to-device copy of `mib` MiB (the runtime's
    copy kernel), timed with HIP events on `stream`; bytes read + bytes written per second."""
    from fastbarnes import hip
    nb = mib << 20
    a = hip.DeviceArray((nb // 4,), np.float32)
    b = hip.DeviceArray((nb // 4,), np.float32)
    cp = hip._load().hipMemcpyAsync
    ms = _time_steps(lambda: hip.check(cp(b.handle, a.handle, nb, hip.D2D, stream.ptr)), stream, reps, 2)
    a.free()
    b.free()
    return 2.0 * nb / (ms * 1e-3) / 1e9


def _errors(out, ref):
    ok = ~np.isnan(ref) & ~np.isnan(out)
    d = out[ok].astype(np.float64) - ref[ok].astype(np.float64)
    return dict(rmse=float(np.sqrt(np.mean(d * d))), maxabs=float(np.max(np.abs(d))),
                mask_mismatch=int((np.isnan(ref) != np.isnan(out)).sum()))


def leg_batch(name, cfgs, steps, warmup, check_fields, note):
    """FP32C batch of uneven fields (graph replays) and its error against the FP64 path
    (the reference's operation order) on the first check_fields fields."""
    from fastbarnes import hip
    from fastbarnes.batched import BarnesBatch
    log('%s: %d fields' % (name, len(cfgs)))
    descs = _descs(cfgs)
    batch = BarnesBatch(descs, precision='fp32c')
    pts = np.concatenate([c['pts'] for c in cfgs])
    val = np.concatenate([c['val'] for c in cfgs])
    d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
    d_out = hip.DeviceArray((batch.total_out,), np.float32)
    s = hip.Stream()
    tuned = None
    if AUTOTUNE:
        batch.autotune(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
        tuned = {'axis%d' % a: batch.pass_groups(a) for a in range(batch.dim)}
    ms = _time_steps(lambda: batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=True), s, steps, warmup)
    batch.set_timing(True)
    acc = {}
    for _ in range(max(3, steps // 4)):
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
        for k, v in batch.stage_times().items():
            acc.setdefault(k, []).append(v)
    batch.set_timing(False)
    npts, nsamp = batch.total_out, batch.total_points
    bp, bs = PIPE_BYTES[batch.dim]
    pipe = bp * npts + bs * nsamp
    res = dict(workload=note, fields=len(cfgs), grid_points=npts, sample_points=nsamp,
               value=round(npts / (ms * 1e-3), 1), unit='grid-pts/s', ms_per_step=round(ms, 4),
               pipeline_alg_bytes=pipe, pipeline_frac=round(pipe / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
               stage_ms={k: round(float(np.mean(v)), 5) for k, v in acc.items()},
               autotuned_segments=tuned)
    out = d_out.numpy()
    b64 = BarnesBatch(descs[:check_fields], precision='fp64')
    d_o64 = hip.DeviceArray((b64.total_out,), np.float32)
    b64.execute(d_pts.ptr, d_val.ptr, d_o64.ptr, s.ptr)
    s.synchronize()
    n = b64.total_out
    res['check_fields'] = check_fields
    res.update({k + '_vs_fp64': v for k, v in _errors(out[:n], d_o64.numpy()).items()})
    return res


def leg_train(cfgs, steps, warmup):
    """Config-5 training step: FP32C forward (keeping the weight field) + adjoint of a batch."""
    from fastbarnes import hip
    from fastbarnes.batched import BarnesBatch
    from oracle import oracle
    log('train: %d fields forward+backward' % len(cfgs))
    descs = _descs(cfgs)
    batch = BarnesBatch(descs, precision='fp32c')
    pts = np.concatenate([c['pts'] for c in cfgs])
    val = np.concatenate([c['val'] for c in cfgs])
    tgt = np.concatenate([c['target'].ravel() for c in cfgs]).astype(np.float32)
    d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
    d_out = hip.DeviceArray((batch.total_out,), np.float32)
    d_ws = hip.DeviceArray((batch.total_out,), np.float32)
    d_gval = hip.DeviceArray((batch.total_points,), np.float64)
    d_gpts = hip.DeviceArray(pts.shape, np.float64)
    s = hip.Stream()
    if AUTOTUNE:
        batch.autotune(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
    batch.forward(d_pts.ptr, d_val.ptr, d_out.ptr, d_ws.ptr, s.ptr)
    s.synchronize()
    out = d_out.numpy()
    g = np.where(np.isnan(out), 0.0, out - tgt).astype(np.float32)   # dL/df of 0.5*||f - target||^2
    d_g = hip.DeviceArray.from_numpy(g)

    def step():
        batch.forward(d_pts.ptr, d_val.ptr, d_out.ptr, d_ws.ptr, s.ptr)
        batch.backward(d_pts.ptr, d_val.ptr, d_out.ptr, d_ws.ptr, d_g.ptr, d_gval.ptr, d_gpts.ptr, s.ptr)

    ms = _time_steps(step, s, steps, warmup)
    npts, nsamp = batch.total_out, batch.total_points
    pipe = TRAIN_BYTES[0] * npts + TRAIN_BYTES[1] * nsamp
    res = dict(workload='%d fields 2400x1200, N=%d each, sigma=1, n=4: fp32c forward + adjoint (dL/dval, dL/dpts) '
                        'of 0.5*||f - target||^2' % (len(cfgs), len(cfgs[0]['val'])),
               fields=len(cfgs), grid_points=npts, value=round(npts / (ms * 1e-3), 1), unit='grid-pts/s (fwd+bwd)',
               ms_per_step=round(ms, 4), pipeline_alg_bytes=pipe,
               pipeline_frac=round(pipe / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
    # gradient of field 0 against the FP64 adjoint
    c = cfgs[0]
    n0 = len(c['val'])
    sh = tuple(c['size'][::-1])
    o0 = out[:sh[0] * sh[1]]
    g0 = np.where(np.isnan(o0), np.nan, o0 - tgt[:o0.size]).reshape(sh)
    rval, rpts = oracle.adjoint(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'], g0)
    gv, gp = d_gval.numpy()[:n0], d_gpts.numpy()[:n0]
    res['grad_val_rel_l2_vs_fp64'] = float(np.linalg.norm(gv - rval) / np.linalg.norm(rval))
    res['grad_pts_rel_l2_vs_fp64'] = float(np.linalg.norm(gp - rpts) / np.linalg.norm(rpts))
    return res


def leg_real(steps):
    """The paper's RMSE-versus-iterations curve (figure 8) on the 3490 real QFF stations at
    2400x1200: num_iter 1..50 for both convolution methods, each method one FP32C batch, against
    naive Barnes (GPU, FP64); the FP64 path's curve beside it.  The plain-convolution counts
    whose integer threshold wraps (FP32C refuses them) run in an FP64 batch."""
    from fastbarnes import hip, realdata
    log('real: figure-8 curve, 3490 stations, 2400x1200')
    pts, val = realdata.read_csv_array(os.path.join(ROOT, 'tests', 'golden', 'qff', 'PressQFF_202007271200_3490.csv'))
    ns = list(range(1, 51))
    t0 = time.perf_counter()
    c32, naive = realdata.rmse_curve(pts, val, 32.0, num_iters=ns, precision='fp32c')
    hip.synchronize()
    t_curve = time.perf_counter() - t0
    c64, _ = realdata.rmse_curve(pts, val, 32.0, num_iters=ns, precision='fp64', naive=naive)
    pick = [1, 2, 3, 4, 5, 10, 20, 50]
    wr = realdata.fp64_counts('convolution', 32.0, num_iters=ns)
    diff = max(abs(a - b) for m in c32 for a, b in zip(c32[m], c64[m]))
    return dict(workload='3490 real QFF stations, paper grid 2400x1200 (x0=(-26+1/32, 34.5)), sigma=1; '
                         'RMSE over [-7,5]x[36,56] against naive Barnes',
                seconds_whole_curve_fp32c_incl_naive_and_host_io=round(t_curve, 3),
                rmse_conv={n: round(c32['convolution'][n - 1], 5) for n in pick},
                rmse_opt_conv={n: round(c32['optimized_convolution'][n - 1], 5) for n in pick},
                convolution_counts_in_fp64='%d..%d' % (wr[0], wr[-1]) if wr else 'none',
                max_abs_rmse_diff_fp32c_vs_fp64=diff)


def _best_of(fn, reps=3):
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        r = fn()
        dt = time.perf_counter() - t0
        best = dt if best is None or dt < best else best
    return best, r


def _err_vs(out, ref):
    ok = ~np.isnan(ref) & ~np.isnan(out)
    d = out[ok].astype(np.float64) - ref[ok].astype(np.float64)
    return dict(rmse=float(np.sqrt(np.mean(d * d))) if d.size else 0.0,
                maxabs=float(np.max(np.abs(d))) if d.size else 0.0,
                mask_mismatch=int((np.isnan(ref) != np.isnan(out)).sum()))


def leg_rows(cfgs, steps):
    """The other entry points of the path, each timed on the GPU beside the CPU oracle on a
    bounded sample, with its error against the oracle there:
    * method='convolution' (alpha = 0): FP32C batch of the headline fields (graph replay) and
      the FP64 drop-in of one field (numpy in and out);
    * barnes_S2 (Lambert projection, convolution on the Lambert grid, resampling) on the 3490
      real QFF stations onto the paper grid 2400x1200: FP32C and FP64 drop-in;
    * method='naive' and method='radius' (FP64 brute force) on one 2400x1200 field."""
    from fastbarnes import hip, realdata
    from fastbarnes.batched import BarnesBatch
    from fastbarnes.interpolation import barnes
    from fastbarnes.interpolationS2 import barnes_S2
    from oracle import oracle
    log('rows: convolution, barnes_S2, naive, radius')
    res = {}
    c0 = cfgs[0]
    args0 = (c0['pts'], c0['val'], c0['sigma'], c0['x0'], c0['step'], c0['size'])
    npf = c0['size'][0] * c0['size'][1]
    # method='convolution', FP32C batch
    descs = [dict(d, method='convolution') for d in _descs(cfgs)]
    batch = BarnesBatch(descs, precision='fp32c')
    pts = np.concatenate([c['pts'] for c in cfgs])
    val = np.concatenate([c['val'] for c in cfgs])
    d_pts, d_val = hip.DeviceArray.from_numpy(pts), hip.DeviceArray.from_numpy(val)
    d_out = hip.DeviceArray((batch.total_out,), np.float32)
    s = hip.Stream()
    if AUTOTUNE:
        batch.autotune(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
    ms = _time_steps(lambda: batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr, graph=True), s, steps, 2)
    f0 = d_out.numpy()[:npf].reshape(c0['size'][::-1])
    t64, ref64 = _best_of(lambda: barnes(*args0, method='convolution', precision='fp64'))
    res['convolution'] = dict(workload='%d fields 2400x1200, N=3490, n=4, method=convolution' % len(cfgs),
                              fp32c_batch_grid_pts_per_s=round(batch.total_out / (ms * 1e-3), 1),
                              fp32c_batch_ms=round(ms, 4), fp32c_vs_fp64=_err_vs(f0, ref64),
                              fp64_dropin_ms_per_field=round(t64 * 1e3, 3))
    # barnes_S2 on the real stations
    rp, rv = realdata.read_csv_array(os.path.join(ROOT, 'tests', 'golden', 'qff', 'PressQFF_202007271200_3490.csv'))
    x0, step, size = realdata.paper_grid(32.0)
    t32, s32 = _best_of(lambda: barnes_S2(rp, rv, 1.0, x0, step, size, precision='fp32c'))
    ts64, s64 = _best_of(lambda: barnes_S2(rp, rv, 1.0, x0, step, size, precision='fp64'))
    xs, ss, sz = realdata.paper_grid(4.0)
    tc = time.perf_counter()
    o_s2 = oracle.barnes_s2_opt(rp, rv, 1.0, xs, ss, sz)
    tc = time.perf_counter() - tc
    g_s2 = barnes_S2(rp, rv, 1.0, xs, ss, sz, precision='fp64')
    res['barnes_S2'] = dict(workload='3490 QFF stations (lon, lat) onto the paper grid 2400x1200, sigma=1, n=4',
                            fp32c_dropin_ms=round(t32 * 1e3, 3), fp64_dropin_ms=round(ts64 * 1e3, 3),
                            fp32c_vs_fp64=_err_vs(s32, s64),
                            cpu_oracle_s_at_300x150=round(tc, 3), fp64_vs_oracle_at_300x150=_err_vs(g_s2, o_s2))
    # naive and radius: FP64 brute force over all stations (radius: within max_dist only)
    for m in ('naive', 'radius'):
        tg, og = _best_of(lambda: barnes(*args0, method=m), reps=2)
        rows = 12   # CPU oracle on the first rows of a coarser copy of the grid
        sub = (c0['size'][0] // 10, rows)
        t0 = time.perf_counter()
        if m == 'naive':
            oc = oracle.interpolate_naive(c0['pts'], c0['val'], c0['sigma'], c0['x0'], c0['step'] * 10, sub)
        else:
            oc = oracle.interpolate_radius(c0['pts'], c0['val'], c0['sigma'], c0['x0'], c0['step'] * 10, sub)
        tcpu = time.perf_counter() - t0
        gs = barnes(c0['pts'], c0['val'], c0['sigma'], c0['x0'], c0['step'] * 10, sub, method=m)
        res[m] = dict(workload='one 2400x1200 field, N=3490 (drop-in, numpy in and out)',
                      gpu_ms=round(tg * 1e3, 3), gpu_grid_pts_per_s=round(npf / tg, 1),
                      cpu_oracle_grid_pts_per_s=round(sub[0] * sub[1] / tcpu, 1), cpu_sample='%dx%d nodes' % sub,
                      gpu_vs_oracle_on_sample=_err_vs(gs, oc))
    return res


def run_legs(names, args):
    from fastbarnes import workloads
    legs = {}
    if 'cfg3' in names:
        cfgs = [workloads.config3_field(i) for i in range(256)]
        legs['cfg3'] = leg_batch('cfg3', cfgs, args.leg_steps, 2, 4,
                                 '256 uneven fields: N~logU[500,50000], grid in {600x300,1200x600,2400x1200}, '
                                 'step 75/W, sigma~U[0.5,2], num_iter~U{3..6}; fp32c, hipGraph replay')
    if 'cfg4' in names:
        legs['cfg4'] = leg_batch('cfg4', [workloads.config4()], args.leg_steps, 2, 1,
                                 '3D 512x512x128, N=200000, step 0.125, sigma=1.5, n=4, max_dist=5; fp32c')
        rng = np.random.default_rng(44)
        cfgs = [workloads.config4(seed=400 + i, n=int(round(np.exp(rng.uniform(np.log(1e4), np.log(2e5))))),
                                  size=(256, 256, 64)) for i in range(32)]
        legs['cfg4_batch'] = leg_batch('cfg4_batch', cfgs, args.leg_steps, 2, 2,
                                       '32 fields 3D 256x256x64, N~logU[1e4,2e5], step 0.125, sigma=1.5, n=4; '
                                       'fp32c, hipGraph replay')
    if 'real' in names:
        legs['real'] = leg_real(args.leg_steps)
    if 'rows' in names:
        legs['rows'] = leg_rows(make_fields(args.fields)[0], args.leg_steps)
    if 'cfg5' in names:
        cfgs = [workloads.config5(seed=500 + i) for i in range(16)]
        legs['cfg5'] = leg_train(cfgs, args.leg_steps, 2)
    return legs


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--fields', type=int, default=64)
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-extras', action='store_true', help='skip accuracy and FP64-path legs')
    ap.add_argument('--legs', default='cfg3,cfg4,cfg5,real,rows',
                    help="comma list of extra workloads reported under 'workloads' (cfg3, cfg4, cfg5, real) "
                         "or 'none'")
    ap.add_argument('--leg-steps', type=int, default=10)
    ap.add_argument('--no-autotune', action='store_true',
                    help="keep the cost model's segment lengths instead of timing the variants")
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='after the run, one device-to-device copy of MIB MiB (a plain copy kernel of known '
                         'size that scales HBM byte counters)')
    args = ap.parse_args()
    global AUTOTUNE
    AUTOTUNE = not args.no_autotune
    if args.gpus != 1:
        print('bench.py measures one GPU only: there is no multi-GPU path (a batch would shard by fields '
              'across GPUs with no communication). Run with --gpus 1.', file=sys.stderr)
        sys.exit(2)

    from fastbarnes import hip, _lib
    from fastbarnes.batched import BarnesBatch
    if _lib.device_count() < 1:
        raise RuntimeError('no GPU visible')

    log('generating %d fields' % args.fields)
    cfgs, descs = make_fields(args.fields)
    batch = BarnesBatch(descs, precision='fp32c')
    pts = np.concatenate([c['pts'] for c in cfgs])
    val = np.concatenate([c['val'] for c in cfgs])
    d_pts = hip.DeviceArray.from_numpy(pts)
    d_val = hip.DeviceArray.from_numpy(val)
    d_out = hip.DeviceArray((batch.total_out,), np.float32)
    stream = hip.Stream()
    npts = batch.total_out
    nsamples = batch.total_points
    tuned = None
    if AUTOTUNE:
        log('autotune')
        batch.autotune(d_pts.ptr, d_val.ptr, d_out.ptr, stream.ptr)
        tuned = {'axis%d' % a: batch.pass_groups(a) for a in range(batch.dim)}

    log('warmup %d' % args.warmup)
    for _ in range(args.warmup):
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, stream.ptr, graph=True)
    stream.synchronize()

    # ---- timed region: K graph replays back to back ----
    log('timed %d steps' % args.steps)
    e0, e1 = hip.Event(), hip.Event()
    hip.synchronize()
    w0 = time.perf_counter()
    e0.record(stream)
    for _ in range(args.steps):
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, stream.ptr, graph=True)
    e1.record(stream)
    e1.synchronize()
    hip.synchronize()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_ms(e1)
    ms_step = ms / args.steps
    value = npts * args.steps / (ms / 1e3)

    # ---- per-kernel durations: the same steps with HIP events between the stages ----
    log('stage timing %d steps' % args.steps)
    batch.set_timing(True)
    acc = {}
    for _ in range(args.steps):
        batch.execute(d_pts.ptr, d_val.ptr, d_out.ptr, stream.ptr, graph=False)
        for k, v in batch.stage_times().items():
            acc.setdefault(k, []).append(v)
    batch.set_timing(False)
    stage_ms = {k: float(np.mean(v)) for k, v in acc.items()}
    kern = max(BYTES_PER_POINT, key=lambda k: stage_ms.get(k, 0.0))
    alg_bytes = BYTES_PER_POINT[kern] * npts + BYTES_PER_SAMPLE[kern] * nsamples
    achieved = alg_bytes / (stage_ms[kern] * 1e-3) / 1e9
    moved_bytes = MOVED_PER_POINT[kern] * npts + BYTES_PER_SAMPLE[kern] * nsamples
    # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counters of a profiled run of this
    # workload (profiles/traffic.json names the commit and command it came from); not measured
    # in this run
    traffic = traffic_src = None
    tfile = os.path.join(ROOT, 'profiles', 'traffic.json')
    if os.path.exists(tfile):
        with open(tfile) as f:
            tj = json.load(f)
        t = tj.get(kern)
        if t and t.get('fields') == args.fields:
            traffic = t.get('bytes_per_launch')
            traffic_src = tj.get('source')
    # the GPU's achievable bandwidth beside the published peak
    copy_gbs = _copy_bandwidth(stream)
    roof = dict(bound='hbm', kernel=kern, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=traffic_src,
                alg_bytes_per_launch=alg_bytes, avg_launch_ms=round(stage_ms[kern], 5),
                moved_bytes_per_launch=moved_bytes,
                frac_moved=round(moved_bytes / (stage_ms[kern] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                copy_gbs=round(copy_gbs, 1), frac_of_copy=round(achieved / copy_gbs, 4))
    pipe_bytes = 28.0 * npts + 24.0 * nsamples
    pipe_moved = 20.0 * npts + 24.0 * nsamples
    pipeline = dict(alg_bytes_per_step=pipe_bytes, achieved=round(pipe_bytes / (ms_step * 1e-3) / 1e9, 1),
                    frac=round(pipe_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    moved_bytes_per_step=pipe_moved,
                    frac_moved=round(pipe_moved / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                    frac_of_copy=round(pipe_bytes / (ms_step * 1e-3) / 1e9 / copy_gbs, 4),
                    stage_ms={k: round(v, 5) for k, v in stage_ms.items()})

    extras = {}
    if not args.no_extras:
        from fastbarnes.interpolation import barnes
        out = d_out.numpy()
        f0 = out[:descs[0]['size'][0] * descs[0]['size'][1]].reshape(descs[0]['size'][::-1])
        c0 = cfgs[0]
        # FP64 path (the reference's operation order; bit-identical to the oracle in tests)
        log('fp64 drop-in')
        ref = barnes(c0['pts'], c0['val'], c0['sigma'], c0['x0'], c0['step'], c0['size'], precision='fp64')
        t = []
        for _ in range(3):
            t0 = time.perf_counter()
            barnes(c0['pts'], c0['val'], c0['sigma'], c0['x0'], c0['step'], c0['size'], precision='fp64')
            t.append(time.perf_counter() - t0)
        fp64_dropin_s = min(t)
        b64 = BarnesBatch(descs[:1], precision='fp64')
        d_o64 = hip.DeviceArray((b64.total_out,), np.float32)
        n0 = descs[0]['n']
        b64.execute(d_pts.ptr, d_val.ptr, d_o64.ptr, stream.ptr)
        stream.synchronize()
        ea, eb = hip.Event(), hip.Event()
        ea.record(stream)
        for _ in range(5):
            b64.execute(d_pts.ptr, d_val.ptr, d_o64.ptr, stream.ptr)
        eb.record(stream)
        eb.synchronize()
        fp64_dev_ms = ea.elapsed_ms(eb) / 5
        ok = ~np.isnan(ref) & ~np.isnan(f0)
        d = f0[ok].astype(np.float64) - ref[ok].astype(np.float64)
        # naive FP64 Barnes on the 240x120 subgrid of every 10th node
        log('naive subgrid')
        nv = barnes(c0['pts'], c0['val'], c0['sigma'], c0['x0'], c0['step'] * 10, (240, 120), method='naive')
        sub = f0[::10, ::10].astype(np.float64)
        sub64 = ref[::10, ::10].astype(np.float64)
        okn = ~np.isnan(sub) & ~np.isnan(nv)
        # interior: nodes at least max_dist*sigma from every border, where no station outside
        # the grid (which naive Barnes uses and the convolution drops) has weight above the
        # max_dist cut-off
        W0, H0 = c0['size']
        xs = np.arange(0, W0, 10) * c0['step']
        ys = np.arange(0, H0, 10) * c0['step']
        reach = 3.5 * c0['sigma']
        inner = ((ys >= reach) & (ys <= (H0 - 1) * c0['step'] - reach))[:, None] & \
                ((xs >= reach) & (xs <= (W0 - 1) * c0['step'] - reach))[None, :]
        oki = okn & inner
        okn64 = ~np.isnan(sub64) & ~np.isnan(nv)
        field_rate = args.fields / (ms_step * 1e-3)
        extras = dict(
            rmse_vs_fp64=float(np.sqrt(np.mean(d * d))), maxabs_vs_fp64=float(np.max(np.abs(d))),
            mask_mismatch_vs_fp64=int((np.isnan(ref) != np.isnan(f0)).sum()),
            rmse_vs_naive_subgrid=float(np.sqrt(np.mean((sub[okn] - nv[okn]) ** 2))),
            rmse_vs_naive_subgrid_interior=float(np.sqrt(np.mean((sub[oki] - nv[oki]) ** 2))),
            maxabs_vs_naive_subgrid_interior=float(np.max(np.abs(sub[oki] - nv[oki]))),
            rmse_fp64_vs_naive_subgrid=float(np.sqrt(np.mean((sub64[okn64] - nv[okn64]) ** 2))),
            rmse_fp64_vs_naive_subgrid_interior=float(np.sqrt(np.mean((sub64[okn64 & inner] - nv[okn64 & inner]) ** 2))),
            naive_subgrid_interior_nodes=int(oki.sum()),
            fp64_dropin_ms_per_field=round(fp64_dropin_s * 1e3, 3),
            fp64_device_ms_per_field=round(fp64_dev_ms, 4),
            fp32c_batched_fields_per_s=round(field_rate, 1),
            speedup_vs_fp64_dropin=round(field_rate * fp64_dropin_s, 1),
            speedup_vs_fp64_device=round(field_rate * fp64_dev_ms * 1e-3, 1))

    legs = run_legs([x for x in args.legs.split(',') if x and x != 'none'], args)

    cpu = None
    if not args.no_cpu:
        log('cpu baseline (%.0f s budget)' % args.cpu_seconds)
        cpu = cpu_baseline(cfgs, args.cpu_seconds)

    res = {
        'metric': METRIC, 'value': round(value, 1), 'unit': 'grid-pts/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(ms_step, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32 storage, f64 window sums (fp32c)', 'data': 'synthetic',
        'config': {'workload': 'batch of %d fields, 2400x1200 grid, N=3490 stations each, sigma=1, num_iter=4, '
                               'max_dist=3.5, optimized_convolution, fp32c, hipGraph replay' % args.fields,
                   'fields': args.fields, 'grid': [2400, 1200], 'stations_per_field': 3490, 'num_iter': 4,
                   'grid_points_per_step': npts, 'autotuned_segments': tuned},
        'roofline': roof, 'pipeline_roofline': pipeline, 'cpu_baseline': cpu,
        'wall_s_timed': round(wall, 4),
    }
    res.update(extras)
    if legs:
        res['workloads'] = legs
    if args.copy_calibration:
        nb = args.copy_calibration << 20
        a = hip.DeviceArray((nb // 4,), np.float32)
        b = hip.DeviceArray((nb // 4,), np.float32)
        hip.check(hip._load().hipMemcpy(b.handle, a.handle, nb, hip.D2D))
        hip.synchronize()
        res['copy_calibration_bytes'] = nb
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
