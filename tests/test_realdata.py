"""
CPU tests of the real-data harness (fastbarnes.realdata) and of the oracle on the real
station file (tests/golden/qff, MeteoSwiss QFF pressure, BSD-3).
"""
import os

import numpy as np
import pytest

from fastbarnes import realdata

QFF = os.path.join(os.path.dirname(__file__), 'golden', 'qff')


@pytest.mark.parametrize('n', [54, 218, 872, 3490])
def test_read_csv_array(n):
    pts, val = realdata.read_csv_array(os.path.join(QFF, 'PressQFF_202007271200_%d.csv' % n))
    assert pts.shape == (n, 2) and val.shape == (n,)
    assert pts.flags.c_contiguous and pts.dtype == np.float64
    # columns are (lon, lat): Europe spans lon -26..49, lat 34..72 in the paper domain
    assert np.nanmin(val) > 950 and np.nanmax(val) < 1050


def test_read_csv_column_swap():
    pts, val = realdata.read_csv_array(os.path.join(QFF, 'PressQFF_202007271200_3490.csv'))
    # first data line: 48.2167,46.7333,1014.6 (lat, lon, value)
    assert tuple(pts[0]) == (46.7333, 48.2167) and val[0] == 1014.6
    pts, val = realdata.read_csv_array(os.path.join(QFF, 'Iceland_PressOFF_constellation.csv'))
    assert pts.shape == (15, 2) and tuple(pts[0]) == (-13.576, 65.268)


def test_read_csv_rejects_bad_header(tmp_path):
    f = tmp_path / 'bad.csv'
    f.write_text('3,3\n1,2,3\n4,5,6\n')
    with pytest.raises(RuntimeError):
        realdata.read_csv_array(str(f))


def test_paper_grid():
    x0, step, size = realdata.paper_grid(32.0)
    assert size == (2400, 1200) and step == 1 / 32 and tuple(x0) == (-26 + 1 / 32, 34.5)
    assert realdata.paper_grid(4.0)[2] == (300, 150)


@pytest.mark.parametrize('res', [4.0, 8.0, 32.0])
def test_rmse_subdomain_extent(res):
    """The subdomain is the nodes with lon in (-7, 5] and lat in [36, 56) of the paper grid:
    a unit difference there gives RMSE 1; differences outside, and NaN, are ignored."""
    x0, step, (W, H) = realdata.paper_grid(res)
    lon = x0[0] + step * np.arange(W)
    lat = x0[1] + step * np.arange(H)
    inside = ((lon[None, :] > -7 + step / 2) & (lon[None, :] < 5 + step / 2)
              & (lat[:, None] > 36 - step / 2) & (lat[:, None] < 56 - step / 2))
    naive = np.zeros((H, W))
    other = np.where(inside, 1.0, 1000.0).astype(np.float32)
    assert realdata.rmse_west_europe(naive, other, x0, step) == 1.0
    other[np.argwhere(inside)[0][0], np.argwhere(inside)[0][1]] = np.nan
    assert realdata.rmse_west_europe(naive, other, x0, step) == 1.0
    # the box has 12 x 20 degrees of nodes
    assert inside.sum() == int(round(12 * res)) * int(round(20 * res))


def test_oracle_real_data_curve_decreases(orc):
    """Paper harness at resolution 4: the optimized convolution approaches naive Barnes
    as the number of iterations grows (the trend of the paper's figure 8)."""
    pts, val = realdata.read_csv_array(os.path.join(QFF, 'PressQFF_202007271200_3490.csv'))
    x0, step, size = realdata.paper_grid(4.0)
    naive = orc.interpolate_naive(pts, val, 1.0, x0, step, size)
    r = [realdata.rmse_west_europe(naive, orc.interpolate_convol(pts, val, 1.0, x0, step, size, num_iter=k),
                                   x0, step) for k in (1, 2, 4, 8)]
    assert r[0] > r[1] > r[2] > r[3]
