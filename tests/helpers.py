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


def ring_vorticity(X, Y, Z, R, sigma, gamma=1.0, zc=0.0):
    """(3, ...) vorticity of a Gaussian-core ring about the z axis:
    omega_theta = Gamma/(pi sigma^2) exp(-((rho-R)^2 + (z-zc)^2)/sigma^2)."""
    rho = np.sqrt(X * X + Y * Y)
    wt = gamma / (np.pi * sigma ** 2) * np.exp(-((rho - R) ** 2 + (Z - zc) ** 2) / sigma ** 2)
    th = np.arctan2(Y, X)
    return np.stack([-np.sin(th) * wt, np.cos(th) * wt, np.zeros_like(wt)])


def ring_particles(n, seed=2, jitter=0.1):
    """Particles of the two-ring vortex step on [-1, 1]^3 with n^3 cells:
    one particle per cell centre, jittered by U(-jitter h, jitter h) per axis
    (seeded), carrying the vorticity of two coaxial rings (R = 0.4 and 0.5,
    sigma = 0.08, Gamma = 1, centres z = 0 and 0.2) at its position.
    Returns pos, vort as (3, n^3) Fortran-ordered arrays, cells in (k, j, i)
    order."""
    h = 2.0 / n
    c = -1.0 + (np.arange(n) + 0.5) * h
    Z, Y, X = np.meshgrid(c, c, c, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()])
    rng = np.random.default_rng(seed)
    pos = pos + rng.uniform(-jitter * h, jitter * h, size=pos.shape)
    w = (ring_vorticity(pos[0], pos[1], pos[2], 0.4, 0.08, 1.0, 0.0)
         + ring_vorticity(pos[0], pos[1], pos[2], 0.5, 0.08, 1.0, 0.2))
    return np.asfortranarray(pos), np.asfortranarray(w)


def oracle_vortex_step(g, p, pos, vort, dt, nsteps, min_vort, Gh=None, visc=None, fd_order=4):
    """The two-ring vortex step composed from the oracle's routines, in the
    order of the reference driver's step mode: per RK2 stage P2M (clear),
    solve, M2P(vel), advance; then P2M and remesh(vel=.TRUE.).  With visc,
    each stage also forms the vorticity right-hand side on the mesh and
    interpolates it to the particles' dvort (the client's order,
    vortex_codes/vortex_structures/source/main.f:453-471)."""
    if Gh is None:
        Gh = O.poisson_setup(g, p)
    z = np.zeros_like(pos)
    vel, dvort, vel_rk, dvort_rk = z.copy(), z.copy(), z.copy(), z.copy()
    for _ in range(nsteps):
        for s in (1, 2):
            mesh = O.p2m(g, pos, vort)
            vel_m = O.poisson_solve(g, p, Gh, mesh)[0]
            vel = O.m2p(g, vel_m, pos)
            if visc is None:
                dvort = np.zeros_like(pos)
            else:
                dvort = O.m2p(g, O.vort_rhs(g, mesh, vel_m, visc, fd_order), pos)
            pos, vel, vort, dvort, vel_rk, dvort_rk = O.advance(g, 2, s, dt, pos, vel, vort, dvort, vel_rk,
                                                                dvort_rk)
        mesh = O.p2m(g, pos, vort)
        pos, vort, vel, dvort = O.remesh(g, mesh, None, vel_m, min_vort, 0.0, False, dvort=False, vel=True)
        vel_rk, dvort_rk = np.zeros_like(pos), np.zeros_like(pos)
    return pos, vort, vel
