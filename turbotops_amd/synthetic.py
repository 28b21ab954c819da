"""Synthetic point clouds of the benchmark configurations.

Every generator draws from numpy's default_rng(seed) in float64 and casts to
float32 at the end.
"""
import numpy as np


def torus_c1(n=1000, seed=1):
    """n points on a torus in R^3 (major radius 1.0, minor 0.3, angles
    U[0, 2pi)^2) plus N(0, 0.01^2) noise."""
    rng = np.random.default_rng(seed)
    th, ph = rng.uniform(0, 2 * np.pi, size=(2, n))
    x = np.stack([(1 + 0.3 * np.cos(th)) * np.cos(ph), (1 + 0.3 * np.cos(th)) * np.sin(ph), 0.3 * np.sin(th)], 1)
    return (x + rng.normal(scale=0.01, size=x.shape)).astype(np.float32)


def gauss(n, d, seed):
    return np.random.default_rng(seed).normal(size=(n, d)).astype(np.float32)


def uniform(n, d, seed):
    return np.random.default_rng(seed).uniform(size=(n, d)).astype(np.float32)


def _manifold_sample(rng, n, d):
    """A noisy sample of a random union of 1-4 circles (in random 2-planes)
    or spheres S^{d-1}: radii U[0.5, 2], centres U[-3, 3]^d, points split
    uniformly across the components, plus N(0, 0.02^2)."""
    comps = int(rng.integers(1, 5))
    sizes = np.full(comps, n // comps)
    sizes[: n - sizes.sum()] += 1
    parts = []
    for m in sizes:
        radius = rng.uniform(0.5, 2.0)
        centre = rng.uniform(-3.0, 3.0, size=d)
        if rng.uniform() < 0.5:       # circle in a random 2-plane
            q, _ = np.linalg.qr(rng.normal(size=(d, 2)))
            t = rng.uniform(0, 2 * np.pi, size=m)
            pts = np.cos(t)[:, None] * q[:, 0] + np.sin(t)[:, None] * q[:, 1]
        else:                          # sphere S^{d-1}
            v = rng.normal(size=(m, d))
            pts = v / np.linalg.norm(v, axis=1, keepdims=True)
        parts.append(centre + radius * pts)
    x = np.concatenate(parts, 0)
    return x + rng.normal(scale=0.02, size=x.shape)


def uneven_clouds(count=512, nmin=200, nmax=4000, dims=(2, 3, 8), seed=6):
    """The batched configuration: `count` clouds, n_i ~ U{nmin..nmax},
    d_i uniform over dims, each a noisy sample of circles/spheres."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        d = int(rng.choice(dims))
        out.append(_manifold_sample(rng, n, d).astype(np.float32))
    return out


def noisy_circle(n=2048, sigma=0.05, seed=7, dim=3):
    """A circle of radius 1 in the xy-plane of R^dim plus N(0, sigma^2)."""
    rng = np.random.default_rng(seed)
    t = rng.uniform(0, 2 * np.pi, size=n)
    x = np.zeros((n, dim))
    x[:, 0] = np.cos(t)
    x[:, 1] = np.sin(t)
    return (x + rng.normal(scale=sigma, size=x.shape)).astype(np.float32)


def noisy_circles(count=64, nmin=512, nmax=4096, sigma=0.05, seed=9, dim=3):
    """`count` noisy circles as noisy_circle, n_i ~ U{nmin..nmax}, drawn from
    one default_rng(seed) stream."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        t = rng.uniform(0, 2 * np.pi, size=n)
        x = np.zeros((n, dim))
        x[:, 0] = np.cos(t)
        x[:, 1] = np.sin(t)
        out.append((x + rng.normal(scale=sigma, size=x.shape)).astype(np.float32))
    return out
