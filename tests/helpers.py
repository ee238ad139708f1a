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


def plane_checksums(u, seed=12345):
    """Size-independent fingerprint of a (3, E0, E1, E2) mesh field: per
    component c and z plane k the weighted sum S[c, k] = sum_ij u[c,i,j,k] r[i,j]
    with seeded normal weights r, plus the bound B = max|u| * sum|r| that a
    relative L-inf difference of 1 between two fields would reach in any S."""
    u = np.asarray(u)
    r = np.random.default_rng(seed).normal(size=u.shape[1:3])
    S = np.einsum("cijk,ij->ck", u, r)
    return {"S": S.tolist(), "B": float(np.abs(u).max() * np.abs(r).sum()), "seed": seed,
            "l2": [float(np.sqrt((u[c] ** 2).sum())) for c in range(u.shape[0])]}


def particle_checksums(pos, vort, vel, seed=54321):
    """Fingerprint of a (3, np) particle set: SHA-256 of the positions'
    bytes (remeshed particles sit exactly on cell centres), and per component
    weighted sums of vort and vel with seeded normal weights, with their
    relative-L-inf bounds (max|a| * sum|r|)."""
    import hashlib
    r = np.random.default_rng(seed).normal(size=pos.shape[1])
    out = {"np": int(pos.shape[1]), "seed": seed,
           "pos_sha256": hashlib.sha256(np.ascontiguousarray(np.asfortranarray(pos).ravel(order="F"))
                                        .tobytes()).hexdigest()}
    for name, a in (("vort", vort), ("vel", vel)):
        out[name] = {"S": (np.asarray(a) @ r).tolist(), "B": float(np.abs(a).max() * np.abs(r).sum())}
    return out


def turbulence_modes(kmax=32, seed=5):
    """Seeded random Fourier coefficients of a solenoidal vector field on the
    periodic box [-1, 1]^3: c[q, nx, ny, nz] for integer wave indices
    nx, ny in [-kmax, kmax], nz in [0, kmax] (wavenumber pi |n|), complex
    normal with variance E(k) / (4 pi k^2), E(k) = k^4 exp(-2 (k/8)^2),
    k = |n|, zero outside |n| <= kmax, then projected onto the plane
    perpendicular to n (divergence-free)."""
    rng = np.random.default_rng(seed)
    m = 2 * kmax + 1
    shape = (3, m, m, kmax + 1)
    c = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    n1 = np.arange(-kmax, kmax + 1, dtype=np.float64)
    NX, NY, NZ = np.meshgrid(n1, n1, np.arange(kmax + 1, dtype=np.float64), indexing="ij")
    k = np.sqrt(NX ** 2 + NY ** 2 + NZ ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = np.where((k > 0) & (k <= kmax), np.sqrt(k ** 4 * np.exp(-2.0 * (k / 8.0) ** 2)
                                                      / (4 * np.pi * k ** 2)), 0.0)
        kk = np.where(k > 0, k, 1.0)
    c = c * amp
    nvec = np.stack([NX / kk, NY / kk, NZ / kk])
    c = c - nvec * np.sum(nvec * c, axis=0)
    return c


def turbulence_vorticity(n, kmax=32, seed=5, zwidth=0.3, torch_device=None):
    """Vorticity of the mixed-BC configuration on n^3 cell centres of
    [-1, 1]^3: the real field sum_n c_n exp(i pi n . x) of turbulence_modes
    (all modes with |n| <= kmax, Hermitian completion implied), multiplied
    by exp(-z^2 / zwidth^2) so it decays toward the free z boundaries.
    Returns (3, n, n, n) float64 in Fortran order (numpy), or a flat
    Fortran-ordered torch tensor on torch_device."""
    assert 2 * kmax < n, "grid too coarse for the modes"
    c = turbulence_modes(kmax, seed)
    full = np.zeros((3, n, n, n // 2 + 1), dtype=np.complex128)
    idx = np.arange(-kmax, kmax + 1) % n
    # sample at cell centres x_j = -1 + (j + 1/2) 2/n: exp(i pi n x_j) =
    # exp(2 pi i n j / n_grid) * exp(i pi n (1/n_grid - 1))
    n1 = np.arange(-kmax, kmax + 1, dtype=np.float64)
    ph1 = np.exp(1j * np.pi * n1 * (1.0 / n - 1.0))
    phz = np.exp(1j * np.pi * np.arange(kmax + 1) * (1.0 / n - 1.0))
    c = c * ph1[None, :, None, None] * ph1[None, None, :, None] * phz[None, None, None, :]
    full[:, idx[:, None], idx[None, :], :kmax + 1] = c
    x = -1.0 + (np.arange(n) + 0.5) * (2.0 / n)
    env = np.exp(-x * x / (zwidth * zwidth))
    if torch_device is None:
        w = np.fft.irfftn(full, s=(n, n, n), axes=(1, 2, 3), norm="forward")
        w *= env[None, None, None, :]
        return np.asfortranarray(w)
    import torch
    out = []
    envt = torch.from_numpy(env).to(torch_device)
    for q in range(3):
        f = torch.from_numpy(full[q]).to(torch_device)
        w = torch.fft.irfftn(f, s=(n, n, n), norm="forward") * envt[None, None, :]
        out.append(w.permute(2, 1, 0).reshape(-1))  # Fortran order of (i, j, k)
        del f
    # interleave the components fastest: (3, n, n, n) Fortran order
    return torch.stack(out, dim=1).reshape(-1)
