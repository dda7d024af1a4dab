"""
Real-data harness: station CSV files, the paper's grid geometry, the subdomain RMSE
convention and the RMSE-versus-iterations curve of the reference's paper (figure 8).

* ``read_csv_array(filename)``: header ``N,3``; rows ``lat,lon,value``; returns
  (pts (N, 2) as (lon, lat), values (N,)) -- demo/reader.py:20-52 (the column swap included).
* ``paper_grid(resolution)``: x0 = (-26 + step, 34.5), step = 1/resolution, size =
  (int(75/step), int(37.5/step)) -- demo/figure8.py:71-73.
* ``rmse_west_europe(naive, other, x0, step)``: RMSE with ``nanmean`` over the Western
  Europe subdomain [-7, 5] x [36, 56] -- demo/rmse.py:20-33.
* ``rmse_curve(...)``: RMSE against naive Barnes for num_iter = 1 .. N and the two
  convolution methods.  On the GPU each method's curve is one batch (one field per
  iteration count) and the naive baseline one call; with FP32C, the plain-convolution
  counts whose integer threshold wraps run in a second, FP64 batch.
"""
import numpy as np

from fastbarnes.batched import BarnesBatch
from fastbarnes.interpolation import barnes

__all__ = ['read_csv_array', 'paper_grid', 'rmse_west_europe', 'rmse_curve', 'fp32c_supported',
           'fp64_counts']


def read_csv_array(filename):
    """Station file -> (pts (N, 2) float64 as (lon, lat), values (N,) float64)."""
    with open(filename, 'r') as f:
        n, m = (int(t) for t in f.readline().split(','))
        if m != 3:
            raise RuntimeError('expected 3 values per line, header says %d' % m)
        rows = np.loadtxt(f, delimiter=',', dtype=np.float64, ndmin=2)
    if rows.shape != (n, 3):
        raise RuntimeError('expected %d rows of 3 values, found %s' % (n, rows.shape))
    pts = np.ascontiguousarray(rows[:, [1, 0]])
    return pts, np.ascontiguousarray(rows[:, 2])


def paper_grid(resolution):
    """(x0, step, size) of the paper's European domain at `resolution` nodes per degree."""
    step = 1.0 / resolution
    x0 = np.asarray([-26.0 + step, 34.5], dtype=np.float64)
    size = (int(75.0 / step), int(37.5 / step))
    return x0, step, size


def rmse_west_europe(naive_field, other_field, x0, step):
    """RMSE over [-7, 5] x [36, 56] (nodes ll <= idx < ur), NaN nodes ignored."""
    ll = [int((-7.0 - x0[0] + step) / step), int((36.0 - x0[1]) / step)]
    ur = [int((5.0 - x0[0] + step) / step), int((56.0 - x0[1]) / step)]
    d = other_field[ll[1]:ur[1], ll[0]:ur[0]].astype(np.float64) - naive_field[ll[1]:ur[1], ll[0]:ur[0]]
    return float(np.sqrt(np.nanmean(d ** 2)))


def fp32c_supported(method, sigma, step, num_iter, dim=2):
    """False where FP32C refuses the field: plain 'convolution' with kernel_size**num_iter
    >= 2**63, where the reference's integer threshold (interpolation.py:669-670) wraps."""
    if method != 'convolution':
        return True
    from fastbarnes.interpolation import _get_half_kernel_size
    ks = 2 * _get_half_kernel_size([sigma] * dim, [step] * dim, num_iter).astype(np.float64) + 1
    return bool(np.all(ks ** num_iter < 2.0 ** 63))


def rmse_curve(pts, val, resolution=32.0, sigma=1.0, num_iters=range(1, 51),
               methods=('convolution', 'optimized_convolution'), precision='fp64', naive=None):
    """{method: [rmse for n in num_iters]} against naive Barnes on the paper grid, plus the
    naive field used (computed once on the GPU unless given).  With precision 'fp32c' the
    counts that FP32C refuses (see fp32c_supported) run in an FP64 batch of their own;
    ``fp64_counts(method, ...)`` lists them."""
    x0, step, size = paper_grid(resolution)
    if naive is None:
        naive = barnes(pts, val, sigma, x0, step, size, method='naive')
    num_iters = list(num_iters)
    out = {}
    for method in methods:
        if precision == 'fp32c':
            wrapped = [k for k in num_iters if not fp32c_supported(method, sigma, step, k)]
            runs = [('fp32c', [k for k in num_iters if k not in wrapped]), ('fp64', wrapped)]
        else:
            runs = [(precision, num_iters)]
        rmse = {}
        for prec, ks in runs:
            if not ks:
                continue
            desc = [dict(sigma=sigma, x0=x0, step=step, size=size, n=len(val), num_iter=k, method=method)
                    for k in ks]
            fields = BarnesBatch(desc, precision=prec).interpolate([pts] * len(ks), [val] * len(ks))
            rmse.update((k, rmse_west_europe(naive, f, x0, step)) for k, f in zip(ks, fields))
        out[method] = [rmse[k] for k in num_iters]
    return out, naive


def fp64_counts(method, resolution=32.0, sigma=1.0, num_iters=range(1, 51)):
    """The iteration counts of rmse_curve(..., precision='fp32c') computed in FP64."""
    step = paper_grid(resolution)[1]
    return [k for k in num_iters if not fp32c_supported(method, sigma, step, k)]
