"""
CPU pins of the oracle's adjoint (oracle/barnes_oracle.c: orc_filter_grid, orc_adjoint).

The reference has no gradient, so nothing in it pins the adjoint; the mathematics does:
  * the line filters are self-adjoint: <B x, y> = <x, B y> for random grids (a transposed
    operand or a one-sided padding breaks this);
  * central finite differences of the FP64 forward (f = vg/wg + offset, before the float32
    cast) in every value and every coordinate, with points kept away from cell edges
    where the bilinear weights have kinks; the error falls as h^2 (checked at two h).
A dropped term of the adjoint (the dependence of W on the positions, the offset-free
value val - offset, the 1/step factor) fails the coordinate check by orders of magnitude.
"""
import numpy as np
import pytest

from oracle import oracle as orc

GEOM = {
    1: dict(x0=np.array([-2.0]), step=np.array([0.25]), size=[60]),
    2: dict(x0=np.array([-2.0, -1.0]), step=np.array([0.25, 0.2]), size=[60, 40]),
    3: dict(x0=np.array([-2.0, -1.0, 0.5]), step=np.array([0.25, 0.2, 0.3]), size=[40, 30, 20]),
}


@pytest.mark.parametrize('dim', [1, 2, 3])
@pytest.mark.parametrize('tail', [True, False])
def test_filter_is_self_adjoint(dim, tail):
    rng = np.random.default_rng(dim)
    shape = tuple(GEOM[dim]['size'][::-1])
    x, y = rng.standard_normal(shape), rng.standard_normal(shape)
    bx = orc.filter_grid(x, 0.8, GEOM[dim]['step'], 4, tail)
    by = orc.filter_grid(y, 0.8, GEOM[dim]['step'], 4, tail)
    lhs, rhs = float((bx * y).sum()), float((x * by).sum())
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # and B is not the identity or a multiple of it
    assert not np.allclose(bx / bx.std(), x / x.std())


def _problem(dim, n, seed):
    g = GEOM[dim]
    rng = np.random.default_rng(seed)
    hi = g['x0'] + g['step'] * (np.array(g['size']) - 1)
    pts = rng.uniform(g['x0'] + 0.6, hi - 0.6, (n, dim))
    c = (pts - g['x0']) / g['step']
    pts = g['x0'] + (np.floor(c) + np.clip(c % 1, 0.1, 0.9)) * g['step']
    val = rng.uniform(5.0, 25.0, n)
    kw = dict(sigma=1.0, x0=g['x0'], step=g['step'], size=g['size'], num_iter=4, max_dist=3.5)
    return pts, val, kw, rng


def _f64(pts, val, kw):
    _, vg, wg = orc.interpolate_convol(pts, val, return_fields=True, **kw)
    return vg / wg + (val.min() + val.max()) / 2


@pytest.mark.parametrize('dim', [1, 2, 3])
def test_adjoint_matches_central_differences(dim):
    n = 24
    pts, val, kw, rng = _problem(dim, n, 10 + dim)
    f0 = _f64(pts, val, kw)
    mask = ~np.isnan(f0)
    G = np.where(mask, rng.standard_normal(f0.shape), 0.0)

    def loss(p, v):
        f = _f64(p, v, kw)
        return float(np.sum(np.where(mask & ~np.isnan(f), f, 0.0) * G))

    gval, gpts = orc.adjoint(pts, val, g=np.where(mask, G, np.nan), **kw)
    assert gval.shape == (n,) and gpts.shape == (n, dim)
    h = 1e-3
    fdv = np.array([(loss(pts, val + h * e) - loss(pts, val - h * e)) / (2 * h) for e in np.eye(n)])
    assert np.abs(fdv - gval).max() <= 1e-8 * np.abs(gval).max()

    def fd_pts(hh):
        out = np.zeros((n, dim))
        for k in range(n):
            for m in range(dim):
                e = np.zeros((n, dim))
                e[k, m] = hh
                out[k, m] = (loss(pts + e, val) - loss(pts - e, val)) / (2 * hh)
        return out

    err4 = np.abs(fd_pts(1e-4) - gpts).max() / np.abs(gpts).max()
    assert err4 <= 1e-6
    if dim == 2:
        # second order: ten times the step, about a hundred times the error (not a constant bias)
        err3 = np.abs(fd_pts(1e-3) - gpts).max() / np.abs(gpts).max()
        assert err3 > 20 * err4


def test_adjoint_drops_outside_points():
    pts, val, kw, rng = _problem(2, 10, 3)
    pts = np.vstack([pts, [[-5.0, 0.0], [kw['x0'][0] + kw['step'][0] * (kw['size'][0] - 1), 1.0]]])
    val = np.concatenate([val, [7.0, 9.0]])
    f0 = _f64(pts, val, kw)
    gval, gpts = orc.adjoint(pts, val, g=np.ones_like(f0), **kw)
    assert np.all(gval[-2:] == 0.0) and np.all(gpts[-2:] == 0.0)
    assert np.all(gval[:-2] != 0.0)
