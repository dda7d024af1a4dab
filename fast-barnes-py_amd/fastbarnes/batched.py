"""
Many fields in one batch: fields with uneven station counts, grid sizes, sigmas and
iteration counts (time steps, variables) interpolated by one plan.

The reference interpolates one field per ``barnes()`` call; a batch runs the same
algorithm over all fields with every stage launched once for the whole batch, and
optionally replays the launches as one captured hipGraph.

    batch = BarnesBatch([dict(sigma=1.0, x0=(-26, 34), step=1/32, size=(2400, 1200), n=3490), ...],
                        precision='fp32c')
    outs = batch.interpolate(pts_list, val_list)            # numpy in, numpy out
    batch.execute(d_pts, d_val, d_out, stream, graph=True)  # device pointers (fastbarnes.hip / torch)
"""
import ctypes
from math import exp

import numpy as np

from fastbarnes import _lib
from fastbarnes._lib import check, lib

_METHOD = {'optimized_convolution': _lib.FB_METHOD_OPTIMIZED_CONVOLUTION, 'convolution': _lib.FB_METHOD_CONVOLUTION}


def _axis3(x, dim, fill):
    if isinstance(x, (list, tuple, np.ndarray)):
        v = [float(a) for a in x]
        if len(v) != dim:
            raise RuntimeError('expected %d values per field, got %d' % (dim, len(v)))
    else:
        v = [float(x)] * dim
    return v + [fill] * (3 - dim)


class BarnesBatch:
    """A plan for a fixed list of field descriptions.

    Each field is a dict with keys sigma, x0, step, size (x, y[, z] order), n (number of
    sample points) and optionally num_iter (4), max_dist (3.5) and method
    ('optimized_convolution' or 'convolution').  All fields share one dimension.
    Points of field f are rows [pt_offsets[f], pt_offsets[f] + n) of the batch's pts
    array; its result occupies out_sizes[f] float32 elements at out_offsets[f].
    """

    def __init__(self, fields, precision='fp32c'):
        if not fields:
            raise RuntimeError('empty batch')
        self.precision = precision
        self.prec = _lib.PRECISIONS[precision]
        arr = (_lib.FbField * len(fields))()
        self.shapes = []
        self.pt_offsets = []
        self.out_offsets = []
        pt = 0
        out = 0
        dim = None
        for i, f in enumerate(fields):
            size = f['size'] if isinstance(f['size'], (list, tuple, np.ndarray)) else (f['size'],)
            d = len(size)
            if dim is None:
                dim = d
            elif d != dim:
                raise RuntimeError('all fields of a batch must have the same dimension')
            a = arr[i]
            a.dim = d
            a.method = _METHOD[f.get('method', 'optimized_convolution')]
            a.num_iter = int(f.get('num_iter', 4))
            a.n = int(f['n'])
            a.pt_offset = pt
            a.out_offset = out
            for m, s in enumerate(list(size) + [1] * (3 - d)):
                a.size[m] = int(s)
            for m, s in enumerate(_axis3(f['sigma'], d, 1.0)):
                a.sigma[m] = s
            for m, s in enumerate(_axis3(f['x0'], d, 0.0)):
                a.x0[m] = s
            for m, s in enumerate(_axis3(f['step'], d, 1.0)):
                a.step[m] = s
            a.max_dist_weight = exp(-float(f.get('max_dist', 3.5)) ** 2 / 2)
            self.shapes.append(tuple(int(s) for s in size[::-1]))
            self.pt_offsets.append(pt)
            self.out_offsets.append(out)
            pt += a.n
            out += int(np.prod(size))
        self.dim = dim
        self.total_points = pt
        self.total_out = out
        self._fields = arr
        self._blocked = True
        self._plan = ctypes.c_void_p()
        check(lib().fb_plan_create(ctypes.byref(self._plan), arr, len(fields), self.prec, 0))

    @property
    def nfields(self):
        return len(self.shapes)

    def info(self):
        cells = ctypes.c_int64()
        ws = ctypes.c_int64()
        check(lib().fb_plan_info(self._plan, ctypes.byref(cells), ctypes.byref(ws)))
        return dict(grid_points=cells.value, workspace_bytes=ws.value)

    def execute(self, d_pts, d_val, d_out, stream, graph=False):
        """Runs the batch on device buffers (integer device addresses) on `stream` (a hipStream_t
        address); asynchronous.  graph=True captures the launches once per buffer/stream tuple
        and replays the captured hipGraph afterwards."""
        fn = lib().fb_plan_execute_graph if graph else lib().fb_plan_execute
        check(fn(self._plan, d_pts, d_val, d_out, stream))

    def forward(self, d_pts, d_val, d_out, d_wsave, stream):
        """Training forward (FP32C plans): execute() that also writes the normalised weight
        field (float32, total_out elements) to d_wsave for backward()."""
        check(lib().fb_plan_forward(self._plan, d_pts, d_val, d_out, d_wsave, stream))

    def backward(self, d_pts, d_val, d_out, d_wsave, d_gout, d_gval, d_gpts, stream):
        """Adjoint of forward(): dL/dout (float32, total_out) -> dL/dval (float64, one per
        point) and, if d_gpts is not None/0, dL/dpts (float64, points x dim)."""
        check(lib().fb_plan_backward(self._plan, d_pts, d_val, d_out, d_wsave, d_gout, d_gval, d_gpts or None,
                                     stream))

    def autotune(self, d_pts, d_val, d_out, stream, cache=None):
        """Times the pass kinds, segment lengths and workgroup sizes of every line pass on these
        buffers and keeps the fastest per axis and line length (FP32C plans).  d_out is left
        undefined.

        cache: a tunecache.TuneCache or the path of its JSON file.  When it holds every
        line-length class of this plan, the stored choices are applied without timing and
        'cache' is returned; otherwise the plan is timed, its choices are stored (the file is
        rewritten) and 'tuned' is returned.  FP64 plans have nothing to tune ('none')."""
        if self.prec != _lib.FB_PREC_FP32C:
            check(lib().fb_plan_autotune(self._plan, d_pts, d_val, d_out, stream))
            return 'none'
        if cache is None:
            check(lib().fb_plan_autotune(self._plan, d_pts, d_val, d_out, stream))
            return 'tuned'
        from fastbarnes import tunecache
        tc = cache if isinstance(cache, tunecache.TuneCache) else tunecache.TuneCache(cache)
        keys = self._class_keys()
        hits = [(axis, length, tc.get(key)) for axis, length, key in keys]
        if all(e is not None for _, _, e in hits):
            try:
                for axis, length, e in hits:
                    self.choose_pass(axis, length, e['kind'], e['seg'], e['threads'])
                return 'cache'
            except RuntimeError:
                pass   # an entry that no longer fits: time the plan afresh (overrides every class)
        check(lib().fb_plan_autotune(self._plan, d_pts, d_val, d_out, stream))
        chosen = {}
        for axis in range(self.dim):
            for length, kind, seg, threads in self.pass_choices(axis)[0]:
                chosen[(axis, length)] = (kind, seg, threads)
        for axis, length, key in keys:
            tc.put(key, *chosen[(axis, length)])
        tc.save()
        return 'tuned'

    def pass_choices(self, axis):
        """([(line length, kind, segment length, workgroup threads)] per line-length class of the
        pass along `axis`, [class index of each field]) with the choices in force (autotuned,
        forced, or the cost model's); FP32C plans."""
        cap = 16
        while True:
            ln = (ctypes.c_int64 * cap)()
            kd = (ctypes.c_int32 * cap)()
            sg = (ctypes.c_int32 * cap)()
            th = (ctypes.c_int32 * cap)()
            fc = (ctypes.c_int32 * self.nfields)()
            n = ctypes.c_int()
            check(lib().fb_plan_pass_choices(self._plan, int(axis), ln, kd, sg, th, cap, ctypes.byref(n), fc))
            if n.value <= cap:
                break
            cap = n.value
        return ([(ln[i], ('slide', 'prefix')[kd[i]], sg[i], th[i]) for i in range(n.value)], list(fc))

    def _class_keys(self):
        """[(axis, line length, cache key)] of every line-length class of the plan."""
        from fastbarnes import tunecache
        arch = ctypes.create_string_buffer(256)
        cus = ctypes.c_int()
        check(lib().fb_device_arch(arch, 256, ctypes.byref(cus)))
        dim = self.dim
        halves = []
        for a in self._fields:
            sig = (ctypes.c_double * 3)(*a.sigma)
            stp = (ctypes.c_double * 3)(*a.step)
            T = (ctypes.c_int32 * 3)()
            check(lib().fb_half_kernel_size(a.method, dim, sig, stp, a.num_iter, T))
            halves.append(list(T))
        out = []
        for axis in range(dim):
            classes, fclass = self.pass_choices(axis)
            for ci, (length, _, _, _) in enumerate(classes):
                members = [(tuple(int(s) for s in a.size[:dim]), halves[f][axis], a.num_iter, a.method)
                           for f, a in enumerate(self._fields) if fclass[f] == ci]
                out.append((axis, length, tunecache.class_key(arch.value.decode(), cus.value, self.precision, axis,
                                                              length, self._blocked, members)))
        return out

    def choose_pass(self, axis, length, kind, seg, threads=256):
        """Forces the pass along `axis` for lines of `length`: kind 'slide' or 'prefix', the
        segment length (an instantiated one) and the workgroup size (FP32C plans)."""
        k = {'slide': 0, 'prefix': 1}[kind]
        check(lib().fb_plan_choose_pass(self._plan, int(axis), int(length), k, int(seg), int(threads)))

    def pass_groups(self, axis):
        """[(kind, segment length, workgroup threads, work items)] of the launches of the pass
        along `axis`; kind is 'slide' (float32 lines, FP64 window sums) or 'prefix' (FP64
        prefix sums)."""
        seg = (ctypes.c_int32 * 32)()
        thr = (ctypes.c_int32 * 32)()
        items = (ctypes.c_int32 * 32)()
        kind = (ctypes.c_int32 * 32)()
        n = ctypes.c_int()
        check(lib().fb_plan_pass_groups(self._plan, axis, seg, thr, items, kind, 32, ctypes.byref(n)))
        return [(('slide', 'prefix')[kind[i]], seg[i], thr[i], items[i]) for i in range(min(n.value, 32))]

    def intermediate_blocked(self, field):
        """True when field `field` keeps its intermediate between the x and y passes in blocks
        of 8 rows x 4 columns (FP32C 2D fields that allow it; same results either way)."""
        b = ctypes.c_int()
        check(lib().fb_plan_field_layout(self._plan, int(field), ctypes.byref(b)))
        return bool(b.value)

    def set_blocked(self, enable):
        """Allows (default) or forbids the blocked intermediate; the same results either way."""
        check(lib().fb_plan_set_blocked(self._plan, int(bool(enable))))
        self._blocked = bool(enable)

    def set_timing(self, enable):
        check(lib().fb_plan_set_timing(self._plan, int(bool(enable))))

    def stage_times(self):
        """{stage: ms} of the last execute() run with timing enabled (HIP events on its stream)."""
        ms = (ctypes.c_float * 16)()
        n = ctypes.c_int()
        names = ctypes.create_string_buffer(512)
        check(lib().fb_plan_stage_times(self._plan, ms, 16, ctypes.byref(n), names, 512))
        keys = names.value.decode().split(',') if n.value else []
        return dict(zip(keys, [ms[i] for i in range(n.value)]))

    def interpolate(self, pts_list, val_list, stream=None):
        """Host convenience: uploads the fields' points and values, runs the batch and
        returns one float32 array per field (shape size[::-1])."""
        from fastbarnes import hip
        if len(pts_list) != self.nfields or len(val_list) != self.nfields:
            raise RuntimeError('expected %d fields' % self.nfields)
        pts = np.concatenate([np.asarray(p, dtype=np.float64).reshape(len(p), -1) for p in pts_list])
        val = np.concatenate([np.asarray(v, dtype=np.float64) for v in val_list])
        if pts.shape[0] != self.total_points or val.shape[0] != self.total_points:
            raise RuntimeError('point counts differ from the plan')
        d_pts = hip.DeviceArray.from_numpy(pts)
        d_val = hip.DeviceArray.from_numpy(val)
        d_out = hip.DeviceArray((self.total_out,), np.float32)
        s = stream or hip.Stream()
        self.execute(d_pts.ptr, d_val.ptr, d_out.ptr, s.ptr)
        s.synchronize()
        flat = d_out.numpy()
        return [flat[o:o + int(np.prod(sh))].reshape(sh) for o, sh in zip(self.out_offsets, self.shapes)]

    def close(self):
        if self._plan:
            lib().fb_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
