"""
CPU pins of the oracle's spherical variant (oracle/barnes_oracle.c: orc_lambert_*,
orc_resample, orc_barnes_s2_opt, orc_naive_s2), restating fastbarnes/interpolationS2.py
and fastbarnes/util/lambert_conformal.py.  The reference has no tests for them; the
mathematics pins them:
  * Lambert conformal conic: the projection origin maps to (0, 0); the scale factor is 1
    on both standard parallels; to_geo inverts to_map; the map is conformal with local
    scale get_scale(lat) along meridians and parallels (finite differences);
  * resample: bilinear resampling reproduces an affine function of the map coordinates
    at every node (a swapped axis or an off-by-one index breaks this);
  * naive_S2: two-station fields against great-circle distances from the haversine
    formula (an independent formula for the spherical distance), invariance under a
    rotation in longitude;
  * optimized_convolution_S2 approaches naive_S2 as num_iter grows.
"""
import os

import numpy as np
import pytest

from oracle import oracle as orc

QFF = os.path.join(os.path.dirname(__file__), 'golden', 'qff', 'PressQFF_202007271200_3490.csv')


def test_lambert_origin_and_standard_parallels():
    p = orc.lambert_proj()
    np.testing.assert_allclose(orc.lambert_to_map([[11.5, 34.5]], p), [[0.0, 0.0]], atol=1e-12)
    for lat in (42.5, 65.5):
        assert abs(orc.lambert_scale(lat, p) - 1.0) < 1e-12
    assert orc.lambert_scale(54.0, p) < 1.0 < orc.lambert_scale(34.5, p)


def test_lambert_round_trip():
    p = orc.lambert_proj()
    rng = np.random.default_rng(0)
    geo = np.column_stack([rng.uniform(-30, 50, 500), rng.uniform(30, 75, 500)])
    np.testing.assert_allclose(orc.lambert_to_geo(orc.lambert_to_map(geo, p), p), geo, atol=1e-9)


def test_lambert_is_conformal_with_analytic_scale():
    p = orc.lambert_proj()
    rng = np.random.default_rng(1)
    h = 1e-5
    for lon, lat in zip(rng.uniform(-25, 45, 20), rng.uniform(35, 70, 20)):
        m = lambda a, b: orc.lambert_to_map([[a, b]], p)[0]
        dlon = (m(lon + h, lat) - m(lon - h, lat)) / (2 * h)
        dlat = (m(lon, lat + h) - m(lon, lat - h)) / (2 * h)
        k = orc.lambert_scale(lat, p)
        assert abs(np.linalg.norm(dlat) - k) < 1e-7
        assert abs(np.linalg.norm(dlon) - k * np.cos(np.radians(lat))) < 1e-7
        assert abs(np.dot(dlon, dlat)) < 1e-7          # meridians and parallels stay orthogonal


def test_resample_reproduces_affine_functions():
    p = orc.lambert_proj()
    step = np.array([0.25, 0.25])
    lam_x0, (lw, lh) = orc.lambert_grid(step)
    mx = lam_x0[0] + step[0] * np.arange(lw)
    my = lam_x0[1] + step[1] * np.arange(lh)
    a, b, c = 0.37, -0.81, 5.0
    lam = (a * mx[None, :] + b * my[:, None] + c).astype(np.float32)
    x0, size = np.array([-26.0 + 0.25, 34.5]), (300, 150)
    out = orc.resample(lam, step, x0, size, p)
    gx = x0[0] + step[0] * np.arange(size[0])
    gy = x0[1] + step[1] * np.arange(size[1])
    geo = np.column_stack([np.tile(gx, size[1]), np.repeat(gy, size[0])])
    mp = orc.lambert_to_map(geo, p)
    expect = (a * mp[:, 0] + b * mp[:, 1] + c).reshape(size[1], size[0])
    assert not np.isnan(out).any()
    np.testing.assert_allclose(out, expect, atol=2e-5)


def _haversine_deg(lon0, lat0, lon1, lat1):
    f0, f1, dl = np.radians(lat0), np.radians(lat1), np.radians(lon1 - lon0)
    s = np.sin((f1 - f0) / 2) ** 2 + np.cos(f0) * np.cos(f1) * np.sin(dl / 2) ** 2
    return np.degrees(2 * np.arcsin(np.sqrt(s)))


def test_naive_s2_two_stations_haversine():
    pts = np.array([[2.0, 47.0], [9.5, 51.0]])
    val = np.array([3.0, 11.0])
    x0, step, size, sigma = np.array([0.0, 45.0]), 0.5, (24, 14), 2.0
    out = orc.naive_s2(pts, val, sigma, x0, step, size)
    gx = x0[0] + step * np.arange(size[0])
    gy = x0[1] + step * np.arange(size[1])
    X, Y = np.meshgrid(gx, gy)
    w = [np.exp(-_haversine_deg(X, Y, px, py) ** 2 / (2 * sigma ** 2)) for px, py in pts]
    expect = (w[0] * val[0] + w[1] * val[1]) / (w[0] + w[1])
    np.testing.assert_allclose(out, expect, rtol=0, atol=1e-9)


def test_naive_s2_longitude_rotation():
    rng = np.random.default_rng(2)
    pts = np.column_stack([rng.uniform(0, 20, 40), rng.uniform(40, 60, 40)])
    val = rng.uniform(990, 1030, 40)
    a = orc.naive_s2(pts, val, 1.5, (0.0, 40.0), 0.5, (40, 40))
    pts2 = pts + np.array([37.0, 0.0])
    b = orc.naive_s2(pts2, val, 1.5, (37.0, 40.0), 0.5, (40, 40))
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-9)


def test_opt_s2_approaches_naive_s2():
    from fastbarnes import realdata
    pts, val = realdata.read_csv_array(QFF)
    x0, step, size = np.array([-26.0 + 0.25, 34.5]), 0.25, (300, 150)
    naive = orc.naive_s2(pts, val, 1.0, x0, step, size)
    r = [realdata.rmse_west_europe(naive, orc.barnes_s2_opt(pts, val, 1.0, x0, step, size, num_iter=k), x0, step)
         for k in (1, 2, 4)]
    assert r[0] > r[1] > r[2] and r[2] < 0.2
    lam = orc.barnes_s2_opt(pts, val, 1.0, x0, step, size, num_iter=4, resample=False)
    assert lam.shape == (176, 256)
