"""
Seeded randomized batches (marked `gpu`): every field of a batch of mixed shapes, through the
batch plan, against the CPU oracle on its own.

Each batch mixes grid shapes (widths that are and are not a multiple of 4, heights that are
and are not a multiple of 8: the 2D blocked intermediate and its row-major fallback), kernel
sizes, iteration counts and point counts, with points outside the grid, crowded rows and
masked regions.  One field of each batch has more than 8192 points, so the plan sorts by row
buckets; a batch of small fields sorts each field in LDS.  Bar: the FP64 path equals the oracle bit for bit; the
FP32C path has the oracle's NaN mask (except at nodes whose weight lies within 1e-5 of the
threshold), RMSE <= 1e-5 and max-abs <= 2e-4 of the oracle's FP64 result.
"""
import numpy as np
import pytest

from test_gpu_parity import MAXABS_TOL, RMSE_TOL, _thr, assert_bitwise, fast_errors

pytestmark = pytest.mark.gpu


def _batch(dim, seed, large):
    rng = np.random.default_rng(seed)
    cases = []
    for f in range(5 if dim < 3 else 3):
        if dim == 1:
            size = (int(rng.integers(60, 3000)),)
        elif dim == 2:
            # even fields: widths a multiple of 4 (the blocked intermediate where the y tiles
            # allow it), odd ones: any width (row-major)
            w = int(rng.integers(24, 420))
            size = (w - w % 4 if f % 2 == 0 else w, int(rng.integers(12, 260)))
        else:
            size = tuple(int(rng.integers(10, 56)) for _ in range(3))
        step = float(rng.uniform(0.05, 0.25))
        sigma = float(rng.uniform(0.4, 1.8)) * step * 8
        x0 = rng.uniform(-5, 5, dim)
        hi = x0 + (np.asarray(size) - 1) * step
        n = int(rng.integers(9000, 12000)) if (large and f == 0) else int(rng.integers(50, 3000))
        # most points inside, some outside the grid, a share on a few crowded rows
        p = rng.uniform(x0 - 2 * step, hi + 2 * step, (n, dim))
        if f % 2:
            # confined to the lower 40 % of the first axis: the rest of the grid is masked
            p[:, 0] = rng.uniform(x0[0] - 2 * step, x0[0] + 0.4 * (hi[0] - x0[0]), n)
        if dim >= 2:
            crowd = rng.random(n) < 0.3
            rows = rng.integers(0, size[1] - 1, 3)
            p[crowd, 1] = x0[1] + (rng.choice(rows, crowd.sum()) + 0.5) * step
        val = np.sin(p[:, 0]) + (np.cos(p[:, 1]) if dim >= 2 else 0.0) + rng.normal(0, 0.05, n)
        cases.append(dict(pts=p if dim > 1 else p[:, 0], val=val, sigma=sigma, x0=x0 if dim > 1 else float(x0[0]),
                          step=step, size=size if dim > 1 else size[0], num_iter=int(rng.integers(3, 7)),
                          max_dist=float(rng.uniform(2.5, 4.0))))
    return cases


def _desc(c):
    size = c['size'] if isinstance(c['size'], tuple) else (c['size'],)
    return dict(sigma=c['sigma'], x0=c['x0'], step=c['step'], size=size, n=len(c['val']), num_iter=c['num_iter'],
                max_dist=c['max_dist'])


@pytest.mark.parametrize('dim,seed,large', [(1, 21, True), (2, 22, True), (2, 23, False), (2, 24, True),
                                            (3, 25, True)])
def test_random_batch_matches_oracle(orc, dim, seed, large):
    from fastbarnes.batched import BarnesBatch
    cases = _batch(dim, seed, large)
    pts = [np.asarray(c['pts'], dtype=np.float64).reshape(len(c['val']), -1) for c in cases]
    vals = [c['val'] for c in cases]
    out64 = BarnesBatch([_desc(c) for c in cases], precision='fp64').interpolate(pts, vals)
    b32 = BarnesBatch([_desc(c) for c in cases], precision='fp32c')
    out32 = b32.interpolate(pts, vals)
    if dim == 2:
        blocked = [b32.intermediate_blocked(f) for f in range(len(cases))]
        print('blocked intermediate:', blocked)
        assert any(blocked) and not all(blocked)
    for c, o64, o32 in zip(cases, out64, out32):
        ref, vg, wg = orc.interpolate_convol(c['pts'], c['val'], c['sigma'], c['x0'], c['step'], c['size'],
                                             num_iter=c['num_iter'], max_dist=c['max_dist'], return_fields=True)
        assert_bitwise(o64.reshape(ref.shape), ref)
        rmse, maxabs, mism_far, mism = fast_errors(o32.reshape(ref.shape), ref, wg, _thr(orc, c))
        assert mism_far == 0 and rmse <= RMSE_TOL and maxabs <= MAXABS_TOL, (c['size'], rmse, maxabs, mism)
