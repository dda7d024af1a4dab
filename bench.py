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
stream); cpu_baseline = the CPU oracle (FP64 C restatement of the reference path,
one core) on a bounded sample of the same fields.

    python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--fields 64] [--cpu-seconds 10]
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


def log(msg):
    print('[bench] ' + msg, file=sys.stderr, flush=True)


def make_fields(nfields, seed0=100):
    from fastbarnes import workloads
    cfgs = [workloads.config2(seed=seed0 + i) for i in range(nfields)]
    descs = [dict(sigma=c['sigma'], x0=c['x0'], step=c['step'], size=c['size'], n=len(c['val']),
                  num_iter=c['num_iter'], max_dist=c['max_dist']) for c in cfgs]
    return cfgs, descs


def cpu_baseline(cfgs, seconds):
    """The oracle (FP64 C restatement of the reference, one core) on a bounded sample of the
    batch's fields: as many whole fields as fit in `seconds` (at least one)."""
    from oracle import oracle
    oracle.lib()
    done = 0
    pts_total = 0
    t0 = time.perf_counter()
    for c in cfgs:
        oracle.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'],
                                  num_iter=c['num_iter'], max_dist=c['max_dist'])
        done += 1
        pts_total += c['size'][0] * c['size'][1]
        if time.perf_counter() - t0 >= seconds:
            break
    dt = time.perf_counter() - t0
    return dict(value=pts_total / dt, unit='grid-pts/s', cores=1, kind='port',
                sample='%d field(s) of 2400x1200, N=3490, n=4 (FP64 oracle, -O2, 1 thread), %.2f s' % (done, dt))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=50)
    ap.add_argument('--warmup', type=int, default=5)
    ap.add_argument('--fields', type=int, default=64)
    ap.add_argument('--cpu-seconds', type=float, default=10.0)
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--no-extras', action='store_true', help='skip accuracy and FP64-path legs')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MIB',
                    help='after the run, one device-to-device copy of MIB MiB (a plain copy kernel of known '
                         'size that scales HBM byte counters)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build has no multi-GPU path (fields shard across GPUs '
              'with no communication; see DESIGN.md). Run with --gpus 1.', file=sys.stderr)
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
    traffic = None
    tfile = os.path.join(ROOT, 'profiles', 'traffic.json')
    if os.path.exists(tfile):
        with open(tfile) as f:
            t = json.load(f).get(kern)
        if t and t.get('fields') == args.fields:
            traffic = t.get('bytes_per_launch')
    roof = dict(bound='hbm', kernel=kern, achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                alg_bytes_per_launch=alg_bytes, avg_launch_ms=round(stage_ms[kern], 5))
    pipe_bytes = 28.0 * npts + 24.0 * nsamples
    pipeline = dict(alg_bytes_per_step=pipe_bytes, achieved=round(pipe_bytes / (ms_step * 1e-3) / 1e9, 1),
                    frac=round(pipe_bytes / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
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
        okn = ~np.isnan(sub) & ~np.isnan(nv)
        field_rate = args.fields / (ms_step * 1e-3)
        extras = dict(
            rmse_vs_fp64=float(np.sqrt(np.mean(d * d))), maxabs_vs_fp64=float(np.max(np.abs(d))),
            mask_mismatch_vs_fp64=int((np.isnan(ref) != np.isnan(f0)).sum()),
            rmse_vs_naive_subgrid=float(np.sqrt(np.mean((sub[okn] - nv[okn]) ** 2))),
            fp64_dropin_ms_per_field=round(fp64_dropin_s * 1e3, 3),
            fp64_device_ms_per_field=round(fp64_dev_ms, 4),
            fp32c_batched_fields_per_s=round(field_rate, 1),
            speedup_vs_fp64_dropin=round(field_rate * fp64_dropin_s, 1),
            speedup_vs_fp64_device=round(field_rate * fp64_dev_ms * 1e-3, 1))

    cpu = None
    if not args.no_cpu:
        log('cpu baseline (%.0f s budget)' % args.cpu_seconds)
        cpu = cpu_baseline(cfgs, args.cpu_seconds)
        cpu['value'] = round(cpu['value'], 1)

    res = {
        'metric': METRIC, 'value': round(value, 1), 'unit': 'grid-pts/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': round(ms_step, 4), 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32 storage, f64 window sums (fp32c)', 'data': 'synthetic',
        'config': {'workload': 'batch of %d fields, 2400x1200 grid, N=3490 stations each, sigma=1, num_iter=4, '
                               'max_dist=3.5, optimized_convolution, fp32c, hipGraph replay' % args.fields,
                   'fields': args.fields, 'grid': [2400, 1200], 'stations_per_field': 3490, 'num_iter': 4,
                   'grid_points_per_step': npts},
        'roofline': roof, 'pipeline_roofline': pipeline, 'cpu_baseline': cpu,
        'wall_s_timed': round(wall, 4),
    }
    res.update(extras)
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
