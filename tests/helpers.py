"""Shared test inputs (host numpy)."""
import numpy as np


def maxwellian(xi, n, T, u):
    """n (2 pi T)^(-3/2) exp(-|xi - u|^2 / 2T) at the nodes xi [N][3]."""
    d2 = np.sum((xi - np.asarray(u, dtype=np.float64)) ** 2, axis=1)
    return n * (2 * np.pi * T) ** -1.5 * np.exp(-d2 / (2 * T))


def mixture(xi, rng, comps=3, nrange=(0.2, 1.0), Trange=(0.3, 1.5), umax=1.2):
    """Seeded sum of Maxwellians: n ~ U(nrange), T ~ U(Trange), u uniform in the ball |u| <= umax."""
    f = np.zeros(xi.shape[0])
    for _ in range(comps):
        n = rng.uniform(*nrange)
        T = rng.uniform(*Trange)
        while True:
            u = rng.uniform(-umax, umax, 3)
            if np.dot(u, u) <= umax * umax:
                break
        f += maxwellian(xi, n, T, u)
    return f


def moments(xi, f, d3v):
    e = np.sum(xi * xi, axis=1)
    psi = np.stack([np.ones(len(e)), xi[:, 0], xi[:, 1], xi[:, 2], e])
    return psi @ np.asarray(f).T * d3v


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
