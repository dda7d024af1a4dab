"""
Seeded synthetic inputs of the benchmark workloads (used by bench.py, smoke() and tests).

The generator follows the workload description of the benchmark configurations:
a uniform + Gaussian-cluster mix of station locations and a smooth analytic
field plus noise as observation values.
"""
import numpy as np

# paper geometry of the 2D workloads: x in [-26, 49], y in [34, 72]
DOMAIN_2D = (-26.0, 49.0, 34.0, 72.0)


def stations_2d(n, seed, domain=DOMAIN_2D, cluster_frac=0.4, nclusters=6, cluster_std=2.0):
    """N points: 40 % from 6 isotropic Gaussian clusters (std 2.0), the rest uniform."""
    rng = np.random.default_rng(seed)
    x0, x1, y0, y1 = domain
    nc = int(round(cluster_frac * n))
    nu = n - nc
    centres = np.column_stack([rng.uniform(x0, x1, nclusters), rng.uniform(y0, y1, nclusters)])
    which = rng.integers(0, nclusters, nc)
    pc = centres[which] + rng.normal(0.0, cluster_std, (nc, 2))
    pu = np.column_stack([rng.uniform(x0, x1, nu), rng.uniform(y0, y1, nu)])
    pts = np.concatenate([pc, pu])
    perm = rng.permutation(n)
    return np.ascontiguousarray(pts[perm])


def values_cfg(pts, seed):
    """val = 15 + 10 sin(x/8) cos(y/6) + N(0, 0.5)."""
    rng = np.random.default_rng(seed + 7919)
    return 15.0 + 10.0 * np.sin(pts[:, 0] / 8.0) * np.cos(pts[:, 1] / 6.0) + rng.normal(0.0, 0.5, len(pts))


def smooth_field(pts, seed, nwaves=8):
    """Random smooth field at the points: 10 + sum of 8 sinusoids with random wave vectors,
    amplitudes and phases."""
    rng = np.random.default_rng(seed + 104729)
    v = np.zeros(len(pts))
    for _ in range(nwaves):
        k = rng.uniform(0.05, 0.4, pts.shape[1])
        a = rng.uniform(0.5, 3.0)
        phi = rng.uniform(0.0, 2 * np.pi)
        v += a * np.sin(pts @ k + phi)
    return 10.0 + v


def values_sinusoids(pts, seed, nwaves=8):
    """Random smooth field (smooth_field) plus N(0, 0.5) noise."""
    rng = np.random.default_rng(seed + 7)
    return smooth_field(pts, seed, nwaves) + rng.normal(0.0, 0.5, len(pts))


def config1(seed=1):
    """Plain 2D Barnes, N=3490, grid 300x150, step 0.25, x0=(-26,34), sigma=1, n=4, max_dist=3.5."""
    pts = stations_2d(3490, seed)
    return dict(pts=pts, val=values_cfg(pts, seed), sigma=1.0, x0=np.array([-26.0, 34.0]), step=0.25,
                size=(300, 150), num_iter=4, max_dist=3.5)


def config2(seed=2):
    """Same generator at 2400x1200, step 1/32."""
    pts = stations_2d(3490, seed)
    return dict(pts=pts, val=values_cfg(pts, seed), sigma=1.0, x0=np.array([-26.0, 34.0]), step=1.0 / 32,
                size=(2400, 1200), num_iter=4, max_dist=3.5)


def config3_field(i, seed=3):
    """One field of the uneven batch: N~logU[500,50000], grid from {600x300,1200x600,2400x1200},
    step = 75/W, sigma~U[0.5,2], num_iter~{3,4,5,6}, values = 8 random sinusoids + noise."""
    rng = np.random.default_rng(seed * 1000003 + i)
    n = int(round(np.exp(rng.uniform(np.log(500), np.log(50000)))))
    W = [600, 1200, 2400][rng.integers(0, 3)]
    H = W // 2
    pts = stations_2d(n, seed * 7 + i)
    return dict(pts=pts, val=values_sinusoids(pts, seed * 11 + i), sigma=float(rng.uniform(0.5, 2.0)),
                x0=np.array([-26.0, 34.0]), step=75.0 / W, size=(W, H), num_iter=int(rng.integers(3, 7)),
                max_dist=3.5)


def config4(seed=4, n=200000, size=(512, 512, 128)):
    """3D: points ~U over [0,64]x[0,64]x[0,16], val = sin(x/5)+cos(y/7)+z/16+N(0,0.1), step 0.125."""
    rng = np.random.default_rng(seed)
    ext = np.array([size[0], size[1], size[2]]) * 0.125
    pts = rng.uniform(0.0, 1.0, (n, 3)) * ext
    val = np.sin(pts[:, 0] / 5) + np.cos(pts[:, 1] / 7) + pts[:, 2] / 16 + rng.normal(0.0, 0.1, n)
    return dict(pts=pts, val=val, sigma=1.5, x0=np.zeros(3), step=0.125, size=size, num_iter=4, max_dist=5.0)


def config5(seed=5, n=100000, size=(2400, 1200)):
    """Differentiable 2D workload: N=100000 uniform+cluster stations, grid 2400x1200, step 75/W,
    sigma=1, n=4, values as in config1; plus an independently seeded smooth target field on
    the grid (shape size[::-1]) for the loss 0.5*||f - target||^2."""
    W, H = size
    pts = stations_2d(n, seed)
    step = 75.0 / W
    x0 = np.array([-26.0, 34.0])
    xs = x0[0] + step * np.arange(W)
    ys = x0[1] + step * np.arange(H)
    gx, gy = np.meshgrid(xs, ys)
    tp = np.column_stack([gx.ravel(), gy.ravel()])
    target = smooth_field(tp, seed + 31337).reshape(H, W)
    return dict(pts=pts, val=values_cfg(pts, seed), sigma=1.0, x0=x0, step=step, size=(W, H), num_iter=4,
                max_dist=3.5, target=target)
