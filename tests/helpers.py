"""Shared inputs for the parity tests (seeded, deterministic)."""
import numpy as np

from oracle import oracle as O


def cell_axes(g):
    """Cell-centre coordinates of every mesh cell (ghosts included) per axis."""
    E = O.mesh_extent(g)
    out = []
    for d in range(3):
        ng = (E[d] - g.ncell[d]) // 2
        i = np.arange(1 - ng, g.ncell[d] + ng + 1, dtype=np.float64)
        out.append(g.xmin[d] + ((i - 1.0) + 0.5) * g.dx[d])
    return out


def blob_field(g, seed=0, nblob=3, width=0.05):
    """Sum of Gaussian blobs per component on the mesh (ghosts included)."""
    rng = np.random.default_rng(seed)
    X, Y, Z = np.meshgrid(*cell_axes(g), indexing="ij")
    w = np.zeros((3,) + X.shape, order="F")
    lo = np.array(g.xmin)
    hi = np.array(g.xmax)
    for c in range(3):
        for _ in range(nblob):
            cx = lo + (hi - lo) * rng.uniform(0.3, 0.7, 3)
            w[c] += rng.normal() * np.exp(-((X - cx[0]) ** 2 + (Y - cx[1]) ** 2 + (Z - cx[2]) ** 2) / width)
    return w


def gaussian_ring(g, R=0.5, sigma=0.1, gamma=1.0, zc=0.0):
    """omega_theta = Gamma/(pi sigma^2) exp(-((rho-R)^2 + (z-zc)^2)/sigma^2), axis z."""
    X, Y, Z = np.meshgrid(*cell_axes(g), indexing="ij")
    rho = np.sqrt(X * X + Y * Y)
    wt = gamma / (np.pi * sigma ** 2) * np.exp(-((rho - R) ** 2 + (Z - zc) ** 2) / sigma ** 2)
    th = np.arctan2(Y, X)
    w = np.zeros((3,) + X.shape, order="F")
    w[0] = -np.sin(th) * wt
    w[1] = np.cos(th) * wt
    return w


def rel_linf(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rel_l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b * b), 1e-300)))
