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
  iteration count) and the naive baseline one call.
"""
import numpy as np

from fastbarnes.batched import BarnesBatch
from fastbarnes.interpolation import barnes

__all__ = ['read_csv_array', 'paper_grid', 'rmse_west_europe', 'rmse_curve']


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


def rmse_curve(pts, val, resolution=32.0, sigma=1.0, num_iters=range(1, 51),
               methods=('convolution', 'optimized_convolution'), precision='fp64', naive=None):
    """{method: [rmse for n in num_iters]} against naive Barnes on the paper grid, plus the
    naive field used (computed once on the GPU unless given)."""
    x0, step, size = paper_grid(resolution)
    if naive is None:
        naive = barnes(pts, val, sigma, x0, step, size, method='naive')
    num_iters = list(num_iters)
    out = {}
    for method in methods:
        desc = [dict(sigma=sigma, x0=x0, step=step, size=size, n=len(val), num_iter=k, method=method)
                for k in num_iters]
        batch = BarnesBatch(desc, precision=precision)
        fields = batch.interpolate([pts] * len(num_iters), [val] * len(num_iters))
        out[method] = [rmse_west_europe(naive, f, x0, step) for f in fields]
    return out, naive
