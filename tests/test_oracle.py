"""Pins the oracle (oracle/pmoracle.c) before it is trusted as the GPU checker:
  * against the golden vectors written by the compiled reference
    (tests/golden/ref_*.npz, tests/golden/make_golden.py),
  * against closed forms and invariants that a plausible mistake would break
    (textbook DFT, single Fourier mode, kernel limits, partition of unity,
    interpolation moments, RK coefficients, NINT rounding at exact halves),
  * and, when oracle/_ref is built in this container, against a live run of
    the reference on a fresh random case.
"""
import glob
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import blob_field, cell_axes, rel_linf

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ----------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 9, 12, 14, 15, 16, 20, 22, 26, 28, 30, 32, 48, 64, 96,
                               100, 128, 160])
def test_glassman_is_the_textbook_dft(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    f = O.glassman(x, -1)
    assert np.abs(f - np.fft.fft(x)).max() <= 1e-13 * np.abs(np.fft.fft(x)).max() * max(1, np.log2(n))
    b = O.glassman(x, +1)
    assert np.abs(b - np.fft.ifft(x) * n).max() <= 1e-13 * np.abs(b).max() * max(1, np.log2(n))


def test_glassman_brute_force_and_impulse():
    n = 12
    x = np.exp(1j * np.arange(n) ** 1.3)
    dft = np.array([sum(x[m] * np.exp(-2j * np.pi * k * m / n) for m in range(n)) for k in range(n)])
    assert np.abs(O.glassman(x) - dft).max() < 1e-12
    e = np.zeros(n, complex)
    e[3] = 1.0
    assert np.abs(O.glassman(e) - np.exp(-2j * np.pi * 3 * np.arange(n) / n)).max() < 1e-14


# ------------------------------------------------------- Green's function
@pytest.mark.parametrize("order,g0", [(2, 1.0), (4, 1.5), (6, 1.875), (8, 2.1875), (10, 2.4609375)])
def test_green_r0_value_is_the_limit(order, g0):
    """G(0) is the r -> 0 limit of the regularised kernel (closed form in
    poisson/pmlib_poisson_setup.f, e.g. :562-570): check against the series
    of the kernel at a tiny radius evaluated in extended precision."""
    h = 1.0 / 16
    g = O.make_geom((8, 8, 8), 0, [0, 0, 0], h)
    G = O.green_real(g, O.make_params(order, 1, 2.0)).reshape((16, 16, 16), order="F")
    eps = 2.0 * h
    c = np.sqrt(2.0) / (4.0 * np.pi ** 1.5 * eps)
    assert abs(G[8, 8, 8] - g0 * c) <= 1e-15 * g0 * c
    # continuity: value at r = h against the kernel formula at r = h
    from math import erf, exp, pi, sqrt
    rho = h / eps
    E = exp(-0.5 * rho * rho) / sqrt(2 * pi)
    poly = {2: 0.0, 4: rho, 6: 1.75 * rho - 0.25 * rho ** 3,
            8: 2.375 * rho - (2 / 3) * rho ** 3 + rho ** 5 / 24,
            10: 2.921875 * rho - 233 / 192 * rho ** 3 + 29 / 192 * rho ** 5 - rho ** 7 / 192}[order]
    ref = (poly * E + erf(rho / sqrt(2))) / (4 * pi * h)
    assert abs(G[9, 8, 8] - ref) <= 1e-14 * ref
    # far field: G -> 1/(4 pi r) (3.5 eps away the smoothing error is < 1%)
    r = 7 * h
    assert abs(G[15, 8, 8] * 4 * pi * r - 1.0) < 1e-2


def test_green_spectrum_is_real_and_even():
    g = O.make_geom((6, 5, 4), 1, [0, 0, 0], 0.1)
    Gh = O.poisson_setup(g, O.make_params(6, 1, 2.0)).reshape((16, 14, 12), order="F")
    assert np.abs(Gh.imag).max() <= 1e-14 * np.abs(Gh).max()
    assert np.abs(Gh - Gh[::-1][np.r_[-1, 0:15]][:, ::-1][:, np.r_[-1, 0:13]]).max() <= 1e-14 * np.abs(Gh).max()


# --------------------------------------------------------------- solve
def test_periodic_single_mode_is_exact():
    """All-periodic solve (poisson_solve.f:799-808): omega_z = sin(2 pi x / Lx)
    has psi_z = omega_z / k^2 and u_y = -dpsi_z/dx exactly."""
    g = O.make_geom((12, 8, 8), 0, [-0.5, -0.6, -0.7], 1.0 / 8, (1, 1, 1))
    p = O.make_params(6, 1, 2.0)
    X, Y, Z = np.meshgrid(*cell_axes(g), indexing="ij")
    k = 2 * np.pi / 1.5
    w = np.zeros((3,) + X.shape, order="F")
    w[2] = np.sin(k * X)
    vel, _, psi = O.poisson_solve(g, p, O.poisson_setup(g, p), w, calc_vecpot=True)
    assert np.abs(psi[2] - np.sin(k * X) / k ** 2).max() < 1e-14
    assert np.abs(vel[1] + np.cos(k * X) / k).max() < 1e-14
    assert np.abs(vel[0]).max() < 1e-14 and np.abs(vel[2]).max() < 1e-14


def test_free_space_solve_converges_to_closed_form_ring():
    """The compact ring of validation/poisson (closed-form omega and u, see
    tests/golden/kat_ring.py): the rms velocity error of the order-6, eps = 2h
    solve falls by more than 2^3 when h halves (a dropped h^3, a wrong
    wavenumber scale or a sign error in the curl breaks this)."""
    import sys
    sys.path.insert(0, GOLD)
    import kat_ring as K
    errs = []
    for n in (16, 32):
        h = 1.0 / n
        ncell = (n, 2 * n, 2 * n)
        xmin = (-0.5, -1.0, -1.0)
        g = O.make_geom(ncell, 4, xmin, h)
        p = O.make_params(6, 1, 2.0)
        w, u = K.ring_fields(xmin, h, ncell, 4)
        vel, _, _ = O.poisson_solve(g, p, O.poisson_setup(g, p), w)
        errs.append(K.rms_velocity_error(vel, u, 4))
    assert errs[1] < 0.05 and errs[0] / errs[1] > 8.0, errs


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_solve_*.npz"))),
                         ids=lambda p: os.path.basename(p)[10:-4])
def test_oracle_solve_matches_reference_golden(path):
    d = np.load(path, allow_pickle=False)
    ncell, ng, bc, order, (rv, rp, cv), kernel = solve_spec(json.loads(str(d["spec"])))
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    p = O.make_params(order, kernel, 2.0)
    vel, vo, ps = O.poisson_solve(g, p, O.poisson_setup(g, p), d["vort_in"], rv, rp, cv)
    assert rel_linf(vel, d["vel"]) <= 2e-15
    if rv or rp:
        assert rel_linf(vo, d["vort"]) <= 2e-15
    if cv and kernel == 1:
        assert rel_linf(ps, d["psi"]) <= 2e-15


def solve_spec(spec):
    """Golden solve spec with the kernel (1 when absent)."""
    return tuple(spec) if len(spec) == 6 else tuple(spec) + (1,)


def _green_closed_form(order, r, eps):
    """Kernel-1 G(r) for r > 0 from the order table (poly(rho) E + F) / (4 pi r);
    complex r allowed (scipy's erf is analytic)."""
    from scipy.special import erf
    rho = r / eps
    E = np.exp(-0.5 * rho * rho) / np.sqrt(2 * np.pi)
    F = erf(rho / np.sqrt(2))
    poly = {0: None, 2: 0 * rho, 4: rho, 6: 1.75 * rho - 0.25 * rho ** 3,
            8: 2.375 * rho - (2 / 3) * rho ** 3 + rho ** 5 / 24,
            10: 2.921875 * rho - (233 / 192) * rho ** 3 + (29 / 192) * rho ** 5 - rho ** 7 / 192}[order]
    if poly is None:
        return 1 / (4 * np.pi * r)
    return (poly * E + F) / (4 * np.pi * r)


@pytest.mark.parametrize("order", [0, 2, 4, 6, 8, 10])
def test_green_kernel2_is_the_gradient_of_kernel1(order):
    """Kernel 2 (grad G) against the complex-step derivative of the kernel-1
    closed form (itself checked against the oracle's kernel-1 grid first):
    K = (x / r) dG/dr at every grid point, 0 at r = 0, and -x / (4 pi r^3)
    in the far field.  A wrong polynomial coefficient fails here."""
    g = O.make_geom((6, 5, 4), 0, [-0.3, -0.2, -0.25], 0.05, (0, 0, 0))
    L = [2 * n for n in g.ncell]
    x = [(np.arange(L[d]) - L[d] // 2) * g.dx[d] for d in range(3)]
    X, Y, Z = np.meshgrid(*x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    eps = 2.0 * max(g.dx)
    G1 = O.green_real(g, O.make_params(order, 1, 2.0)).reshape(L, order="F")
    K = O.green_real(g, O.make_params(order, 2, 2.0)).reshape([3] + L, order="F")
    nz = r > 0
    g_np = _green_closed_form(order, r[nz], eps)
    assert np.abs(g_np - G1[nz]).max() <= 1e-14 * np.abs(G1[nz]).max()
    h = 1e-30
    dG = (_green_closed_form(order, r[nz] + 1j * h, eps)).imag / h
    for c, P in enumerate((X, Y, Z)):
        want = P[nz] / r[nz] * dG
        assert np.abs(K[c][nz] - want).max() <= 1e-12 * np.abs(want).max(), (order, c)
        assert K[c][~nz].tolist() == [0.0]
    far = r > 8 * eps
    if far.any():
        for c, P in enumerate((X, Y, Z)):
            want = -P[far] / (4 * np.pi * r[far] ** 3)
            assert np.abs(K[c][far] - want).max() <= 1e-6 * np.abs(want).max()


# -------------------------------------------------------- interpolation
def test_cell_index_rounds_half_away_from_zero():
    """NINT((x - xmin) * (1/h)) (interpolation/pmlib_interp_particle_mesh.f:153)."""
    f = O.lib().or_cell_index
    assert f(1.5, 0.0, 1.0) == 2 and f(2.5, 0.0, 1.0) == 3 and f(-0.5, 0.0, 1.0) == -1
    assert f(-1.5, 0.0, 1.0) == -2
    x = 0.0 + 2.5 * 0.1
    assert f(np.nextafter(2.5, 0), 0.0, 1.0) == 2 and f(np.nextafter(2.5, 3), 0.0, 1.0) == 3
    # the subtraction and the multiply are each correctly rounded (no FMA):
    xmin, h = -1.0, 2.0 / 128
    for i in range(-3, 131):
        x = xmin + ((i - 1) + 0.5) * h
        t = (x - xmin) * (1.0 / h)
        assert f(x, xmin, 1.0 / h) == int(np.sign(t) * np.floor(abs(t) + 0.5))


@pytest.mark.parametrize("io", [3, 4])
def test_interpolation_partition_of_unity_and_moments(io):
    """One particle deposited away from walls: weights sum to 1, first and
    second moments vanish (M'4 and M6* reproduce quadratics)."""
    g = O.make_geom((16, 16, 16), 0, [0, 0, 0], 1.0 / 16, (0, 0, 0), io)
    rng = np.random.default_rng(io)
    for _ in range(20):
        x = rng.uniform(0.3, 0.7, 3)
        m = O.p2m(g, x.reshape(3, 1), np.array([[1.0], [0.0], [0.0]]))[0]
        X, Y, Z = np.meshgrid(*cell_axes(g), indexing="ij")
        assert abs(m.sum() - 1.0) < 1e-13
        for A, c in ((X, x[0]), (Y, x[1]), (Z, x[2])):
            assert abs((m * (A - c)).sum()) < 1e-13
            assert abs((m * (A - c) ** 2).sum()) < 1e-13


def _geom_from(d):
    ncell, ng, bc, io = json.loads(str(d["geom"]))
    return O.make_geom(ncell, ng, [-0.5, -0.5, -0.5], 1.0 / 8, bc, io)


PART = sorted(glob.glob(os.path.join(GOLD, "ref_particles_*.npz")))


@pytest.mark.parametrize("path", PART, ids=lambda p: os.path.basename(p)[14:-4])
def test_oracle_particles_match_reference_golden_bitwise(path):
    d = np.load(path, allow_pickle=False)
    g = _geom_from(d)
    assert np.array_equal(O.p2m(g, d["pos"], d["val"]), d["p2m"])
    assert np.array_equal(O.m2p(g, d["mesh"], d["pos"]), d["m2p"])
    for ig in (0, 1):
        P, W, V, D = O.remesh(g, d["rm_vort"], d["rm_dvort"], d["rm_vel"], 1.2, 0.15, bool(ig))
        assert np.array_equal(np.concatenate([P, W, V], axis=1), d["remesh%d" % ig])
    for st in (1, 2):
        out = O.advance(g, 2, st, 0.2, d["a_pos"], d["a_vel"], d["a_vort"], d["a_dvort"], d["a_vel_rk"],
                        d["a_dvort_rk"])
        mine = np.concatenate([out[0], out[2], out[4], out[5]], axis=1)
        assert np.array_equal(mine, d["advance%d" % st])


def test_rk_coefficients_closed_form():
    """Constant velocity: every scheme moves x by dt*u over its stages
    (particles/pmlib_particles_advance.f:49-88)."""
    g = O.make_geom((8, 8, 8), 0, [-10, -10, -10], 2.5, (0, 0, 0))
    rng = np.random.default_rng(0)
    x0 = rng.uniform(-1, 1, (3, 5))
    u = rng.uniform(-1, 1, (3, 5))
    for order in (1, 2, 3):
        arrs = [x0, u, np.zeros((3, 5)), np.zeros((3, 5)), np.zeros((3, 5)), np.zeros((3, 5))]
        for st in range(1, order + 1):
            arrs = O.advance(g, order, st, 0.1, *arrs)
            arrs[1] = u.copy()
        assert np.abs(arrs[0] - (x0 + 0.1 * u)).max() < 1e-15


def test_remesh_order_and_threshold():
    g = O.make_geom((3, 2, 2), 0, [0, 0, 0], 1.0)
    vm = np.zeros((3, 3, 2, 2), order="F")
    vm[0, 2, 0, 0] = 2.0
    vm[1, 0, 1, 1] = 3.0
    vm[2, 1, 0, 1] = 0.5
    P, W, V, D = O.remesh(g, vm, None, None, 1.0, 0.0)
    assert P.shape[1] == 2
    assert np.array_equal(P[:, 0], [2.5, 0.5, 0.5]) and np.array_equal(P[:, 1], [0.5, 1.5, 1.5])


def test_remesh_mask_and_copy_flags_reduce_to_the_plain_remesh():
    """mask=.TRUE. keeps exactly the unmasked particles of the plain remesh,
    in the same order (remesh/pmlib_remesh.f:107-113); dvort/vel = .FALSE.
    zero those particle fields and change nothing else (:208-219)."""
    g = O.make_geom((9, 7, 8), 2, [-0.3, -0.2, -0.1], 0.1, (0, 1, 0))
    rng = np.random.default_rng(21)
    ext = (3,) + O.mesh_extent(g)
    vm, dvm, velm = (np.asfortranarray(rng.normal(size=ext) * s) for s in (1.0, 0.1, 1.0))
    mask = rng.random(ext[1:]) < 0.6
    for ig in (False, True):
        P, W, V, D = O.remesh(g, vm, dvm, velm, 1.2, 0.15, ig)
        cells = [np.rint((P[d] - g.xmin[d]) / g.dx[d] - 0.5).astype(int) + (ext[1 + d] - g.ncell[d]) // 2
                 for d in range(3)]
        keep = mask[cells[0], cells[1], cells[2]]
        Pm, Wm, Vm, Dm = O.remesh(g, vm, dvm, velm, 1.2, 0.15, ig, mask=mask)
        for a, b in zip((Pm, Wm, Vm, Dm), (P, W, V, D)):
            assert np.array_equal(a, b[:, keep])
        P2, W2, V2, D2 = O.remesh(g, vm, dvm, velm, 1.2, 0.15, ig, dvort=False, vel=False)
        assert np.array_equal(P2, P) and np.array_equal(W2, W)
        assert not V2.any() and not D2.any()
        # the vorticity really is copied from the cell the particle sits in
        assert np.array_equal(W, vm[:, cells[0], cells[1], cells[2]])
        assert np.array_equal(V, velm[:, cells[0], cells[1], cells[2]])


# ------------------------------------------------------ live reference
@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")
def test_oracle_matches_live_reference_random_case():
    rng = np.random.default_rng(99)
    ncell = tuple(int(v) for v in rng.integers(6, 13, 3))
    g = O.make_geom(ncell, 1, [-0.4, -0.5, -0.6], 1.0 / 10, (0, 0, 0))
    p = O.make_params(8, 1, 2.0)
    w = blob_field(g, seed=5)
    raw, _ = O.ref_run(dict(O.geom_job(g, p), mode="solve", reg_vort=True), [w])
    ref = np.frombuffer(raw, dtype=np.float64)
    vel, vo, _ = O.poisson_solve(g, p, O.poisson_setup(g, p), w, True)
    assert rel_linf(vel, ref[:w.size].reshape(w.shape, order="F")) <= 2e-15
    assert rel_linf(vo, ref[w.size:].reshape(w.shape, order="F")) <= 2e-15


# ----------------------------------------------------------- penalisation
def _poly_fields(g, deg):
    """A velocity field whose components are polynomials of total degree
    `deg`, its analytic curl, and the cell-centre coordinates."""
    X, Y, Z = np.meshgrid(*cell_axes(g), indexing="ij")
    rng = np.random.default_rng(deg)
    a = rng.normal(size=(3, 4))
    u = np.zeros((3,) + X.shape, order="F")
    du = np.zeros((3, 3) + X.shape)   # du[c, d] = d u_c / d x_d
    V = (X, Y, Z)
    for c in range(3):
        # u_c = a0 x^deg + a1 y^deg + a2 z^deg + a3 x y z^(deg-2)
        u[c] = a[c, 0] * X ** deg + a[c, 1] * Y ** deg + a[c, 2] * Z ** deg + a[c, 3] * X * Y * Z ** (deg - 2)
        du[c, 0] = deg * a[c, 0] * X ** (deg - 1) + a[c, 3] * Y * Z ** (deg - 2)
        du[c, 1] = deg * a[c, 1] * Y ** (deg - 1) + a[c, 3] * X * Z ** (deg - 2)
        du[c, 2] = deg * a[c, 2] * Z ** (deg - 1) + (deg - 2) * a[c, 3] * X * Y * Z ** max(deg - 3, 0)
    curl = np.stack([du[2, 1] - du[1, 2], du[0, 2] - du[2, 0], du[1, 0] - du[0, 1]])
    return u, curl, V


@pytest.mark.parametrize("fd_order,deg", [(4, 4), (4, 3), (2, 2)])
def test_penalisation_curl_exact_on_polynomials(fd_order, deg):
    """Central differences of order p are exact on polynomials of degree p
    (4th order: c1 = 1/6, c2 = 4/3 over 2h, penalisation_solve.f:127,195-273),
    so with chi = 1 and a solid at rest the increment is -relax curl(u)
    exactly up to rounding; a sign, index, coefficient or transposition
    mistake breaks it.  The force and enstrophy sums follow from it."""
    g = O.make_geom((10, 9, 8), 2, [-0.3, -0.2, -0.45], 1.0, (0, 0, 0))
    g.dx[0], g.dx[1], g.dx[2] = 0.1, 0.12, 0.09
    u, curl, (X, Y, Z) = _poly_fields(g, deg)
    E = O.mesh_extent(g)
    w0 = np.asfortranarray(np.random.default_rng(1).normal(size=(3,) + E))
    relax = 0.7
    w1, diag = O.penalise(g, u, np.ones(E), w0, relax=relax, fd_order=fd_order)
    inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
    dw = -relax * curl[inner]
    scale = np.abs(dw).max()
    assert np.abs((w1 - w0)[inner] - dw).max() <= 1e-11 * scale
    ghost = np.ones(E, dtype=bool)
    ghost[2:-2, 2:-2, 2:-2] = False
    assert np.array_equal(w1[:, ghost], w0[:, ghost])
    px, py, pz = X[inner[1:]], Y[inner[1:]], Z[inner[1:]]
    F = [np.sum(py * dw[2] - pz * dw[1]), np.sum(pz * dw[0] - px * dw[2]), np.sum(px * dw[1] - py * dw[0])]
    assert np.allclose(diag[:3], F, rtol=1e-10, atol=1e-10 * scale * dw[0].size)
    assert abs(diag[3] - np.sum(dw ** 2)) <= 1e-10 * np.sum(dw ** 2)
    assert abs(diag[4] - np.sum(w1[inner] ** 2)) <= 1e-12 * np.sum(w1[inner] ** 2)


def test_penalisation_brinkman_factor_and_solid_velocity():
    """f = chi (us - u) is what gets differentiated (penalisation_solve.f:96-111):
    the same increment comes from chi = 1, us = 0 and u = -f; and with chi = 1
    a uniform solid velocity drops out of the curl."""
    g = O.make_geom((9, 8, 10), 2, [-0.3, -0.2, -0.45], 0.1, (0, 0, 0))
    u, _, (X, Y, Z) = _poly_fields(g, 3)
    chi = 0.5 + 0.3 * X
    us = np.zeros_like(u)
    us[0] = 1.25   # free stream seen from the solid
    E = O.mesh_extent(g)
    w0 = np.asfortranarray(np.random.default_rng(3).normal(size=(3,) + E))
    w_a, d_a = O.penalise(g, u, chi, w0, solid_vel=us, relax=0.9)
    w_b, d_b = O.penalise(g, -(chi * (us - u)), np.ones(E), w0, relax=0.9)
    assert np.array_equal(w_a, w_b) and np.array_equal(d_a, d_b)
    inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
    w_c, _ = O.penalise(g, u, np.ones(E), w0, solid_vel=us, relax=0.9)
    w_d, _ = O.penalise(g, u, np.ones(E), w0, relax=0.9)
    assert np.abs(w_c[inner] - w_d[inner]).max() <= 1e-12 * np.abs(w_d[inner]).max()


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_step_*.npz"))),
                         ids=lambda p: os.path.basename(p)[9:-4])
def test_oracle_vortex_step_matches_reference_golden(path):
    """Whole two-ring vortex steps (P2M, order-6 solve, M2P, RK2, remesh)
    composed from the oracle's routines against the reference driver."""
    from tests.helpers import oracle_vortex_step
    d = np.load(path, allow_pickle=False)
    n, bc, nsteps, min_vort = json.loads(str(d["spec"]))
    h = 2.0 / n
    g = O.make_geom((n, n, n), 0, [-1.0] * 3, h, bc, 3)
    p = O.make_params(6, 1, 2.0)
    pos, vort, vel = oracle_vortex_step(g, p, d["pos_in"], d["vort_in"], 0.5 * h, nsteps, min_vort)
    assert pos.shape == d["pos"].shape
    assert np.array_equal(pos, d["pos"])
    assert rel_linf(vort, d["vort"]) <= 1e-13
    assert rel_linf(vel, d["vel"]) <= 1e-13


# ------------------------------------------------------ vorticity right-hand side
def _poly_vort(g, deg, seed):
    """w_l = b0 x^deg + b1 y^deg + b2 z^deg + b3 x y z^(deg-2) and its
    Laplacian, per component."""
    X, Y, Z = np.meshgrid(*cell_axes(g), indexing="ij")
    b = np.random.default_rng(seed).normal(size=(3, 4))
    w = np.zeros((3,) + X.shape, order="F")
    lap = np.zeros((3,) + X.shape)
    for c in range(3):
        w[c] = b[c, 0] * X ** deg + b[c, 1] * Y ** deg + b[c, 2] * Z ** deg + b[c, 3] * X * Y * Z ** (deg - 2)
        lap[c] = deg * (deg - 1) * (b[c, 0] * X ** (deg - 2) + b[c, 1] * Y ** (deg - 2) + b[c, 2] * Z ** (deg - 2))
        if deg >= 4:
            lap[c] += (deg - 2) * (deg - 3) * b[c, 3] * X * Y * Z ** (deg - 4)
    return w, lap


def _poly_jacobian(g, deg):
    """_poly_fields' velocity and its Jacobian du[c, d] = d u_c / d x_d."""
    X, Y, Z = np.meshgrid(*cell_axes(g), indexing="ij")
    a = np.random.default_rng(deg).normal(size=(3, 4))
    u = np.zeros((3,) + X.shape, order="F")
    du = np.zeros((3, 3) + X.shape)
    for c in range(3):
        u[c] = a[c, 0] * X ** deg + a[c, 1] * Y ** deg + a[c, 2] * Z ** deg + a[c, 3] * X * Y * Z ** (deg - 2)
        du[c, 0] = deg * a[c, 0] * X ** (deg - 1) + a[c, 3] * Y * Z ** (deg - 2)
        du[c, 1] = deg * a[c, 1] * Y ** (deg - 1) + a[c, 3] * X * Z ** (deg - 2)
        du[c, 2] = deg * a[c, 2] * Z ** (deg - 1) + (deg - 2) * a[c, 3] * X * Y * Z ** max(deg - 3, 0)
    return u, du


@pytest.mark.parametrize("fd_order,deg", [(4, 4), (4, 3), (2, 2)])
def test_vort_rhs_exact_on_polynomials(fd_order, deg):
    """part_deriv_vorticity.f:57-196: the 4th-order stencils are exact on
    quartics, the 2nd-order ones on quadratics, so dvort_l = visc lap(w_l) +
    sum_m w_m du_l/dx_m holds to rounding on the interior; the velocity
    Jacobian is not symmetric, so a transposed du_m/dx_l, a sign, a
    coefficient or a dropped term fails.  Ghost cells stay as given."""
    g = O.make_geom((10, 9, 8), 2, [-0.3, -0.2, -0.45], 1.0, (0, 0, 0))
    g.dx[0], g.dx[1], g.dx[2] = 0.1, 0.12, 0.09
    w, lap = _poly_vort(g, deg, 5)
    u, du = _poly_jacobian(g, deg)
    visc = 0.013
    want = visc * lap + np.einsum("mxyz,lmxyz->lxyz", w, du)
    E = O.mesh_extent(g)
    sentinel = np.full((3,) + tuple(E), 7.5, order="F")
    got = O.vort_rhs(g, w, u, visc, fd_order, dvort=sentinel)
    inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
    assert np.abs(got[inner] - want[inner]).max() <= 1e-10 * np.abs(want[inner]).max()
    ghost = np.ones(E, dtype=bool)
    ghost[2:-2, 2:-2, 2:-2] = False
    assert (got[:, ghost] == 7.5).all()


def test_vort_rhs_second_order_is_second_order():
    """On quartics the 2nd-order stencils are not exact (the error is the
    O(h^2) term), while the 4th-order ones are: the two paths differ."""
    g = O.make_geom((10, 9, 8), 2, [-0.3, -0.2, -0.45], 0.1, (0, 0, 0))
    w, lap = _poly_vort(g, 4, 6)
    u, du = _poly_jacobian(g, 4)
    want = 0.02 * lap + np.einsum("mxyz,lmxyz->lxyz", w, du)
    inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
    e2 = np.abs(O.vort_rhs(g, w, u, 0.02, 2)[inner] - want[inner]).max()
    e4 = np.abs(O.vort_rhs(g, w, u, 0.02, 4)[inner] - want[inner]).max()
    assert e4 <= 1e-10 * np.abs(want[inner]).max() < e2


@pytest.mark.parametrize("fd_order", [2, 4])
def test_vort_rhs_periodic_modified_wavenumbers(fd_order):
    """All-periodic patch with single Fourier modes: central differences of
    sin and cos give the modified wavenumbers in closed form,
    2nd order: -(2 - 2 cos kh) / h^2 and sin(kh) / h, 4th order:
    -(30 - 32 cos kh + 2 cos 2kh) / (12 h^2) and (8 sin kh - sin 2kh) / (6 h);
    the ghost layers of dvort are the interior's periodic images."""
    n = (12, 10, 8)
    g = O.make_geom(n, 0, [0.0, 0.0, 0.0], 1.0, (1, 1, 1), 3)
    g.dx[0], g.dx[1], g.dx[2] = 0.1, 0.125, 0.15
    for d in range(3):
        g.xmax[d] = g.xmin[d] + n[d] * g.dx[d]
    X, Y, Z = np.meshgrid(*cell_axes(g), indexing="ij")
    V = (X, Y, Z)
    k = [2 * np.pi / (n[d] * g.dx[d]) for d in range(3)]
    h = [g.dx[d] for d in range(3)]
    if fd_order == 2:
        k2 = [-(2 - 2 * np.cos(k[d] * h[d])) / h[d] ** 2 for d in range(3)]
        k1 = [np.sin(k[d] * h[d]) / h[d] for d in range(3)]
    else:
        k2 = [-(30 - 32 * np.cos(k[d] * h[d]) + 2 * np.cos(2 * k[d] * h[d])) / (12 * h[d] ** 2) for d in range(3)]
        k1 = [(8 * np.sin(k[d] * h[d]) - np.sin(2 * k[d] * h[d])) / (6 * h[d]) for d in range(3)]
    # w_l = sin(k_l x_l) (one mode per component, along its own axis)
    # u_l = cos(k_m x_m) along axis m = l + 1 (mod 3)
    w = np.zeros((3,) + X.shape, order="F")
    u = np.zeros((3,) + X.shape, order="F")
    want = np.zeros((3,) + X.shape)
    visc = 0.05
    for l in range(3):
        w[l] = np.sin(k[l] * V[l])
        u[l] = np.cos(k[(l + 1) % 3] * V[(l + 1) % 3])
    for l in range(3):
        m = (l + 1) % 3
        # lap(w_l) = k2_l sin(k_l x_l); du_l/dx_m = -k1_m sin(k_m x_m)
        want[l] = visc * k2[l] * np.sin(k[l] * V[l]) + w[m] * (-k1[m] * np.sin(k[m] * V[m]))
    got = O.vort_rhs(g, w, u, visc, fd_order)
    ng = 2
    inner = (slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    assert np.abs(got[inner] - want[inner]).max() <= 1e-12 * np.abs(want[inner]).max()
    # ghosts: mapped images of the interior (x, then y, then z)
    for d in range(3):
        N = n[d]
        lo = [slice(None)] * 4
        src = [slice(None)] * 4
        lo[d + 1] = slice(0, ng)
        src[d + 1] = slice(N, N + ng)
        assert np.array_equal(got[tuple(lo)], got[tuple(src)])
        hi = [slice(None)] * 4
        src2 = [slice(None)] * 4
        hi[d + 1] = slice(ng + N, ng + N + ng)
        src2[d + 1] = slice(ng, 2 * ng)
        assert np.array_equal(got[tuple(hi)], got[tuple(src2)])


# ---------------------------------------------------------------- diagnostics
def _ring_particles(n, R, x0, gamma, h, U):
    """n particles on a circle of radius R in the plane x = x0 about the x
    axis, vorticity gamma ds / h^3 along e_theta (ds = 2 pi R / n), velocity U."""
    th = 2 * np.pi * (np.arange(n) + 0.25) / n
    pos = np.stack([np.full(n, x0), R * np.cos(th), R * np.sin(th)])
    s = gamma * (2 * np.pi * R / n) / h ** 3
    vort = np.stack([np.zeros(n), -s * np.sin(th), s * np.cos(th)])
    vel = np.tile(np.asarray(U, dtype=np.float64)[:, None], (1, n))
    return np.asfortranarray(pos), np.asfortranarray(vort), np.asfortranarray(vel), s


def test_diag_particles_discrete_ring_closed_forms():
    """A discrete circular ring (axis x, radius R, circulation gamma):
    impulse (pi R^2 gamma, 0, 0), zero angular impulse and helicity,
    kinetic energy 2 U . impulse for a uniform U, enstrophy n s^2 h^3,
    centroid (x0, 0, 0), ring radius R, core radius 0, circulation gamma,
    max |vort| = s.  A transposed cross product, a sign or a swapped y/z
    fails one of these."""
    n, R, x0, gamma, h = 720, 0.37, 0.11, 1.3, 0.02
    U = (0.7, 0.2, -0.3)
    pos, vort, vel, s = _ring_particles(n, R, x0, gamma, h, U)
    d = dict(zip(O.DIAG_PARTICLE_NAMES, O.diag_particles((h, h, h), pos, vort, vel)))
    imp = np.pi * R * R * gamma
    assert abs(d["impulse_x"] - imp) <= 1e-13 * imp
    scale = n * s * h ** 3  # sum of |terms| of the vanishing sums
    for k in ("impulse_y", "impulse_z", "ang_impulse_x", "ang_impulse_y", "ang_impulse_z", "helicity"):
        assert abs(d[k]) <= 1e-13 * scale, (k, d[k])
    assert abs(d["vort_core"]) <= 1e-13 * R
    assert abs(d["kin_energy"] - 2 * U[0] * imp) <= 1e-13 * imp
    assert abs(d["enstrophy"] - n * s * s * h ** 3) <= 1e-13 * n * s * s * h ** 3
    assert abs(d["centroid_x"] - x0) <= 1e-14 and abs(d["centroid_y"]) <= 1e-14 and abs(d["centroid_z"]) <= 1e-14
    assert abs(d["vort_radius"] - R) <= 1e-14
    assert abs(d["circulation"] - gamma) <= 1e-13 * gamma
    assert abs(d["vort_max"] - s) <= 1e-14 * s


def test_diag_particles_shifted_ring_and_scaling():
    """Shifting the ring in y, z moves the centroid and leaves radius and
    circulation; reversing the vorticity flips impulse and circulation and
    keeps enstrophy.  (Impulse depends on the origin; it is taken about 0.)"""
    n, R, h = 400, 0.25, 0.03
    pos, vort, vel, s = _ring_particles(n, R, 0.0, 1.0, h, (0.0, 0.0, 0.0))
    pos2 = pos.copy()
    pos2[1] += 0.2
    pos2[2] -= 0.15
    d = dict(zip(O.DIAG_PARTICLE_NAMES, O.diag_particles((h, h, h), pos2, vort, vel)))
    assert abs(d["centroid_y"] - 0.2) <= 1e-14 and abs(d["centroid_z"] + 0.15) <= 1e-14
    assert abs(d["vort_radius"] - R) <= 1e-13 and abs(d["circulation"] - 1.0) <= 1e-13
    r = dict(zip(O.DIAG_PARTICLE_NAMES, O.diag_particles((h, h, h), pos, -vort, vel)))
    assert abs(r["circulation"] + 1.0) <= 1e-13
    assert abs(r["impulse_x"] + np.pi * R * R) <= 1e-13
    assert abs(r["enstrophy"] - n * s * s * h ** 3) <= 1e-13 * n * s * s * h ** 3


def test_diag_mesh_uniform_fields_closed_forms():
    """Uniform u and w on the interior (ghosts filled with junk that must be
    ignored): kinetic energy |u|^2 V / 2, enstrophy |w|^2 V, helicity
    (w.u)^2 / |w|^2 V, Lamb |w x u|^2 / |w|^2 V (so helicity + Lamb = 2 kin),
    max |w|."""
    g = O.make_geom((7, 5, 6), 2, [-0.3, 0.1, -0.2], 0.05, (0, 0, 0))
    E = O.mesh_extent(g)
    u = np.array([0.3, -1.1, 0.7])
    w = np.array([2.0, 0.5, -1.5])
    vel = np.full((3,) + E, 1e3, order="F")
    vort = np.full((3,) + E, -1e3, order="F")
    vel[:, 2:-2, 2:-2, 2:-2] = u[:, None, None, None]
    vort[:, 2:-2, 2:-2, 2:-2] = w[:, None, None, None]
    V = 7 * 5 * 6 * 0.05 ** 3
    d = dict(zip(O.DIAG_MESH_NAMES, O.diag_mesh(g, vel, vort)))
    ww, uu = w @ w, u @ u
    assert abs(d["kin_energy"] - 0.5 * uu * V) <= 1e-13 * uu * V
    assert abs(d["enstrophy"] - ww * V) <= 1e-13 * ww * V
    assert abs(d["helicity"] - (w @ u) ** 2 / ww * V) <= 1e-13 * uu * V
    assert abs(d["lamb"] - np.cross(w, u) @ np.cross(w, u) / ww * V) <= 1e-13 * uu * V
    assert abs(d["helicity"] + d["lamb"] - 2 * d["kin_energy"]) <= 1e-13 * uu * V
    assert abs(d["vort_max"] - np.sqrt(ww)) <= 1e-15 * np.sqrt(ww)


# ------------------------------------------------------ penalisation_solve
def _penal_case(n=16, ng=2):
    g = O.make_geom((n, n, n), ng, [-1.0, -1.0, -1.0], 2.0 / n, (0, 0, 0))
    E = O.mesh_extent(g)
    axes = [g.xmin[d] + (np.arange(E[d]) - ng + 0.5) * g.dx[d] for d in range(3)]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    chi = np.asfortranarray((r < 0.45).astype(np.float64))
    vel = np.zeros((3,) + E, order="F")
    vel[0] = 1.0
    return g, chi, vel, r


def test_penalisation_first_force_is_twice_the_penalised_momentum():
    """One iteration, uniform U = e_x through a body at rest: dvort = curl(f)
    with f = -chi U, and sum x x curl(f) h^3 = 2 sum f h^3 for compactly
    supported f (summation by parts of the central differences), so
    force = -(sum x x dvort) h^3 / dt = 2 U V_chi / dt along the flow, zero
    across it.  A wrong sign or a missing h^3 or 1/dt fails."""
    g, chi, vel, _ = _penal_case()
    p = O.make_params(4, 1, 2.0)
    Gh = O.poisson_setup(g, p)
    dt = 0.05
    vort = np.zeros_like(vel, order="F")
    _, _, force, npen, _ = O.penalisation_solve(g, p, Gh, vel, vort, chi, max_iter=1, dtime=dt, fd_order=4)
    V = chi[2:-2, 2:-2, 2:-2].sum() * np.prod(g.dx)
    assert npen == 1
    assert abs(force[0] - 2 * V / dt) <= 1e-10 * (2 * V / dt), (force, 2 * V / dt)
    assert abs(force[1]) <= 1e-10 * force[0] and abs(force[2]) <= 1e-10 * force[0]


def test_penalisation_iterations_reduce_the_slip_and_stop_at_the_tolerance():
    """Each iteration lowers the slip |chi (us - u - u_pen)| inside the body,
    the loop stops at the first iteration whose relative force increment is
    <= tolerance (or at penal_max), and the free stream is added as the
    reference writes it (vel_inf(3) onto the second component)."""
    g, chi, vel, r = _penal_case()
    p = O.make_params(4, 1, 2.0)
    Gh = O.poisson_setup(g, p)
    vort = np.zeros_like(vel, order="F")
    slips, incs = [], []
    for it in (1, 2, 4, 8):
        v, w, f, n, res = O.penalisation_solve(g, p, Gh, vel, vort, chi, tolerance=0.0, max_iter=it, dtime=0.1)
        assert n == it
        # with no mesh vorticity the output is the velocity induced by the
        # penalisation vorticity, which cancels more of the unit stream
        # inside the body with every iteration
        slips.append(np.sqrt(((vel + v) ** 2).sum(0))[r < 0.45].mean())
        incs.append(res)
    assert all(a > b for a, b in zip(slips, slips[1:])), slips
    assert slips[-1] < 0.3 * slips[0]
    assert all(a > b for a, b in zip(incs[1:], incs[2:])), incs
    tol = incs[2] * 1.0001
    v, w, f, n, res = O.penalisation_solve(g, p, Gh, vel, vort, chi, tolerance=tol, max_iter=50, dtime=0.1)
    assert n <= 4 and res <= tol
    vq, _, _, _, _ = O.penalisation_solve(g, p, Gh, vel, vort, chi, max_iter=1, vel_inf=(0.0, 0.0, 2.5))
    v0, _, _, _, _ = O.penalisation_solve(g, p, Gh, vel, vort, chi, max_iter=1)
    assert np.array_equal(vq[0], v0[0]) and np.array_equal(vq[2], v0[2])
    assert np.allclose(vq[1] - v0[1], 2.5, rtol=0, atol=1e-13)
    vf, _, _, _, _ = O.penalisation_solve(g, p, Gh, vel, vort, chi, max_iter=1, vel_inf=(0.0, 0.0, 2.5),
                                          fix_free_stream=True)
    assert np.allclose(vf[2] - v0[2], 2.5, rtol=0, atol=1e-13) and np.array_equal(vf[1], v0[1])


# ---------------------------------------------------------------------------
# Multiresolution: refinement-patch kernels, parent topology, mesh-to-mesh
# (poisson/pmlib_poisson_setup.f:446-528 and the other orders' refinement
# blocks, topology/pmlib_topology_parent.f:5-100,
# interpolation/pmlib_interp_mesh_mesh.f:5-2078)

def _refine_setup(spec):
    ncell, ng, bc, order, kernel, level, parent = spec
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc)
    return g, O.make_params(order, kernel, 2.0, refine=2.0 ** (level - parent))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_refine_*.npz"))),
                         ids=lambda p: os.path.basename(p)[11:-4])
def test_oracle_refinement_solve_matches_reference_golden(path):
    d = np.load(path, allow_pickle=False)
    g, p = _refine_setup(json.loads(str(d["spec"])))
    vel, _, _ = O.poisson_solve(g, p, O.poisson_setup(g, p), d["vort_in"])
    assert rel_linf(vel, d["vel"]) <= 2e-15


@pytest.mark.parametrize("order,kernel", [(2, 1), (4, 1), (6, 2), (8, 1), (10, 2)])
def test_refinement_kernel_is_the_difference_of_two_regularisations(order, kernel):
    """G_refine(eps, a) = G(eps) - G(a eps) on a free-space kernel grid: the
    base kernel built with the regularisation radius scaled by a, including
    the r = 0 values (1 - 1/a) G(0) of kernel 1."""
    g = O.make_geom((8, 6, 10), 1, [-0.3, -0.2, -0.4], 0.05)
    a = 4.0
    ref = O.green_real(g, O.make_params(order, kernel, 2.0, refine=a))
    g1 = O.green_real(g, O.make_params(order, kernel, 2.0))
    ga = O.green_real(g, O.make_params(order, kernel, 2.0 * a))
    diff = g1 - ga
    assert np.abs(ref - diff).max() <= 1e-13 * np.abs(g1).max()
    # a = 1: the two regularisations cancel
    assert np.abs(O.green_real(g, O.make_params(order, kernel, 2.0, refine=1.0))).max() == 0.0


def test_refinement_kernel_rejects_order_0():
    g = O.make_geom((8, 8, 8), 0, [-0.5] * 3, 1.0 / 8)
    with pytest.raises(ValueError):
        O.green_real(g, O.make_params(0, 1, 2.0, refine=2.0))


def test_topology_parent_geometry():
    g = O.make_geom((16, 12, 8), 2, [-0.5, -0.6, -0.7], 1.0 / 8, (0, 1, 0), 4)
    gp = O.topology_parent(g, 3, 1)
    assert tuple(gp.ncell) == (4, 3, 2) and gp.dx[0] == 0.5 and gp.nghost == 3
    assert tuple(gp.bound_cond) == (0, 1, 0) and tuple(gp.xmin) == tuple(g.xmin)
    assert O.topology_parent(O.make_geom((16, 12, 8), 0, [0.0] * 3, 0.1), 2, 1).nghost == 2
    with pytest.raises(ValueError):
        O.topology_parent(O.make_geom((16, 10, 8), 0, [0.0] * 3, 0.1), 3, 1)   # 10 % 4 != 0
    with pytest.raises(ValueError):
        O.topology_parent(g, 2, 0)


def _m2m_geoms(spec):
    ncell, ng, bc, io, level, clear = spec
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc, io)
    return g, O.topology_parent(g, level, 1), clear


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_m2m_*.npz"))),
                         ids=lambda p: os.path.basename(p)[8:-4])
def test_oracle_m2m_matches_reference_golden_bitwise(path):
    d = np.load(path, allow_pickle=False)
    g, gp, clear = _m2m_geoms(json.loads(str(d["spec"])))
    assert tuple(gp.ncell) == tuple(d["parent_ncell"])
    np.testing.assert_array_equal(O.interp_mesh_mesh(g, gp, d["fine"], clear=True), d["down"])
    np.testing.assert_array_equal(O.interp_mesh_mesh(gp, g, d["coarse"], d["base"], clear=clear), d["up"])


def _interior_blob(g, seed, margin):
    """A random field that vanishes within `margin` cells of every mesh edge."""
    rng = np.random.default_rng(seed)
    E = O.mesh_extent(g)
    f = np.zeros((3,) + E, order="F")
    sl = tuple(slice(margin, e - margin) for e in E)
    f[(slice(None),) + sl] = rng.normal(size=(3,) + tuple(e - 2 * margin for e in E))
    return np.asfortranarray(f)


@pytest.mark.parametrize("io,level", [(3, 2), (4, 2), (3, 3), (4, 3)])
def test_m2m_deposit_is_the_scaled_adjoint_of_the_interpolation(io, level):
    """<D f, c>_parent = reldx^3 <f, I c>_patch in free space when no stencil
    is clamped (f vanishes near the patch edges, every target of I is
    computed): the deposit weights reldx W(d) and the interpolation weights
    W(d) are the same kernel of the same offset, so a transposed operand, a
    wrong stencil slot or index fails it.  The deposit also keeps the
    total: sum D f = reldx^3 sum f."""
    g = O.make_geom((16, 16, 16), 3, [-0.5, -0.6, -0.7], 1.0 / 16, (0, 0, 0), io)
    gp = O.topology_parent(g, level, 1)
    reldx = g.dx[0] / gp.dx[0]
    f = _interior_blob(g, 31 + io, 6)
    c = np.asfortranarray(np.random.default_rng(41 + level).normal(size=(3,) + O.mesh_extent(gp)))
    Df = O.interp_mesh_mesh(g, gp, f, clear=True)
    Ic = O.interp_mesh_mesh(gp, g, c, clear=True)
    lhs = (Df * c).sum()
    rhs = reldx ** 3 * (f * Ic).sum()
    assert abs(lhs - rhs) <= 1e-12 * (np.abs(Df * c).sum())
    assert abs(Df.sum() - reldx ** 3 * f.sum()) <= 1e-12 * np.abs(f).sum()


@pytest.mark.parametrize("io,deg", [(3, 2), (4, 3)])
def test_m2m_interpolation_reproduces_polynomials(io, deg):
    """Parent -> patch reproduces polynomials up to the kernel's degree
    (M'4: quadratics, M6*: cubics) wherever the stencil lies inside the
    parent mesh."""
    g = O.make_geom((16, 16, 16), 2, [-0.5, -0.6, -0.7], 1.0 / 16, (0, 0, 0), io)
    gp = O.topology_parent(g, 2, 1)

    def field(gg):
        E = O.mesh_extent(gg)
        x = [gg.xmin[d] + (np.arange(E[d]) - O.ng_side(gg, d) + 0.5) * gg.dx[d] for d in range(3)]
        X, Y, Z = np.meshgrid(*x, indexing="ij")
        p = 0.3 + X - 2 * Y + 0.5 * Z + X * Y - Z * Z
        if deg == 3:
            p = p + 0.7 * X ** 3 - Y * Z * Z
        return np.asfortranarray(np.stack([p, 2 * p, -p]))

    got = O.interp_mesh_mesh(gp, g, field(gp), clear=True)
    want = field(g)
    ngp, ng = O.ng_side(gp, 0), O.ng_side(g, 0)
    # patch cells whose parent stencil stays inside the parent mesh: those
    # within the patch interior, two parent cells from the parent's edge
    sl = slice(ng + 4, O.mesh_extent(g)[0] - ng - 4)
    assert np.abs(got[:, sl, sl, sl] - want[:, sl, sl, sl]).max() <= 1e-12 * np.abs(want).max()
    assert ngp >= 2



# ---------------------------------------------------------------------------
# Order 100: the band-limited kernel (poisson/pmlib_poisson_setup.f:1419-1505)

def test_sine_integral_restates_cisi():
    """or_si follows the cisi method to its 1e-10 truncation: within 1e-10 of
    scipy's Si, odd, 0 at 0, pi/2 at infinity, and continuous across the
    switch from series to continued fraction at |x| = 2."""
    import ctypes as C
    from scipy.special import sici
    L = O.lib()
    L.or_si.restype = C.c_double
    L.or_si.argtypes = [C.c_double]
    xs = np.concatenate([np.linspace(-40, 40, 801), [1e-20, 1e-8, 1.9999999, 2.0000001, 123.4, 1e4]])
    got = np.array([L.or_si(x) for x in xs])
    assert np.abs(got - sici(xs)[0]).max() <= 1e-10
    assert L.or_si(0.0) == 0.0 and L.or_si(-1.3) == -L.or_si(1.3)
    assert abs(L.or_si(1e8) - np.pi / 2) <= 1e-7
    assert abs(L.or_si(2.0 - 1e-12) - L.or_si(2.0 + 1e-12)) <= 1e-10


@pytest.mark.parametrize("kernel", [1, 2])
def test_order100_kernel_closed_form(kernel):
    """G = Si(pi r / h) / (2 pi^2 r), G(0) = 1 / (2 pi h); grad G = x (sin rho -
    Si(rho)) / (2 pi^2 r^3), 0 at r = 0 (free space, scipy's Si)."""
    from scipy.special import sici
    g = O.make_geom((6, 4, 8), 0, [-0.3, -0.2, -0.4], 0.05)
    G = O.green_real(g, O.make_params(100, kernel, 2.0))
    L = [2 * n for n in g.ncell]
    h = max(g.dx)
    x = [(-(0.5 * L[d] + 0.5) + np.arange(L[d]) + 0.5) * g.dx[d] for d in range(3)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    rho = np.pi * r / h
    with np.errstate(divide="ignore", invalid="ignore"):
        if kernel == 1:
            want = np.where(r == 0, 1.0 / (2 * np.pi * h), sici(rho)[0] / (2 * np.pi ** 2 * r)).ravel()
            got = G
        else:
            f = np.where(r == 0, 0.0, (np.sin(rho) - sici(rho)[0]) / (2 * np.pi ** 2 * r ** 3))
            want = np.stack([X * f, Y * f, Z * f], axis=-1).ravel()
            got = G
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
    with pytest.raises(ValueError):    # no refinement kernel for order 100
        O.green_real(g, O.make_params(100, kernel, 2.0, refine=2.0))



@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "ref_rhs_*.npz"))),
                         ids=lambda p: os.path.basename(p)[8:-4])
def test_oracle_rhs_matches_client_routine_bitwise(path):
    """or_vort_rhs against the vortex client's own part_deriv_vorticity
    (vortex_codes/vortex_structures/source/particle_derivatives/
    part_deriv_vorticity.f), compiled with the reference library."""
    d = np.load(path, allow_pickle=False)
    ncell, ng, bc, io, fd = json.loads(str(d["spec"]))
    g = O.make_geom(ncell, ng, [-0.5, -0.6, -0.7], 1.0 / 8, bc, io)
    got = O.vort_rhs(g, d["vort"], d["vel"], float(d["visc"]), fd, dvort=d["dvort_in"])
    np.testing.assert_array_equal(got, d["dvort"])


def test_ring_kat_is_what_the_reference_validation_program_prints():
    """tests/golden/validation_poisson_convergence.dat is the table the
    reference's own validation/poisson program prints when built with the
    reference library here (oracle/Makefile.ref validation_poisson): the
    rms velocity errors 0.02293 / 7.59e-4 of the harness's ring fixture, not
    the 1.619e-6 / 2.049e-10 of the shipped validation/poisson/convergence.dat."""
    rows = [list(map(float, l.split())) for l in open(os.path.join(GOLD, "validation_poisson_convergence.dat"))
            if l.strip()]
    kat = json.load(open(os.path.join(GOLD, "kat_ring_ref.json")))["cases"]
    assert len(rows) == len(kat) == 2
    for (h, err), case in zip(rows, kat):
        assert h == case["h"]
        # the table prints 12 significant digits
        assert abs(err - case["rms_error"]) <= 1e-11 * case["rms_error"]
        assert abs(err / case["convergence_dat"]) > 1e3
