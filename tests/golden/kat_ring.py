"""Known-answer test of the free-space Poisson solve: the compact C-infinity
vortex ring of validation/poisson/sources/poisson.f:193-263 (vorticity) and
:315-340 (velocity), whose rms relative velocity errors are stored in
validation/poisson/convergence.dat:1-2.

The fields are derived here from the ring's stream function (the commented
"stream function check" of poisson.f:342-347):

    psi = psi_theta(x, rho) e_theta,  psi_theta = exp(-a / D),
    a = c r0^2,  D = 2 r0 rho - rho^2 - x^2 = r0^2 - phi^2,
    phi^2 = (rho - r0)^2 + x^2,  rho = sqrt(y^2 + z^2),  theta = atan2(z, y),

so u = curl psi and omega = curl u in cylindrical coordinates around the x
axis, with support phi < r0.  Parameters: c = 10, r0 = 0.5 (poisson.f:182-186).

The grid follows the patch that pmlib_patch_adjust produces on the second and
third call of that program (patch/pmlib_patch_adjust.f:47-66 mutates xmin/xmax
each call): h = 1/32 -> ncell (40,72,72), xmin (-0.625,-1.125,-1.125);
h = 1/64 -> ncell (88,152,152), xmin (-0.6875,-1.1875,-1.1875) (the latter is
confirmed by validation/poisson/out.pvti:3).  nghost = 4, order 6, eps = 2h,
kernel 1, reg_vort = .TRUE. (poisson.f:101-116, :268-269).
"""
import numpy as np

C_RING = 10.0
R0_RING = 0.5

# (h, ncell, xmin) of the two resolutions and the stored rms errors
KAT_CASES = [
    (1.0 / 32, (40, 72, 72), (-0.625, -1.125, -1.125), 0.161906725038e-05),
    (1.0 / 64, (88, 152, 152), (-0.6875, -1.1875, -1.1875), 0.204917513782e-09),
]
KAT_NGHOST = 4
KAT_ORDER = 6
KAT_EPS_RADIUS = 2.0


def cell_centres(xmin, dx, ncell, ng):
    """Cell-centre coordinates xmin + (i - 1 + 1/2) dx, i = 1-ng .. N+ng,
    evaluated in the same operation order as the reference."""
    i = np.arange(1 - ng, ncell + ng + 1, dtype=np.float64)
    return xmin + ((i - 1.0) + 0.5) * dx


def _fields(px, py, pz):
    a = C_RING * R0_RING ** 2
    rho = np.sqrt(pz * pz + py * py)
    phi = np.sqrt((rho - R0_RING) ** 2 + px ** 2)
    theta = np.arctan2(pz, py)
    inside = phi < R0_RING
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        D = 2.0 * R0_RING * rho - rho ** 2 - px ** 2
        f = np.where(inside, np.exp(-a / np.where(inside, D, 1.0)), 0.0)
        Dn = np.where(inside, D, 1.0)
        rn = np.where(inside, rho, 1.0)
        E = a / Dn ** 2
        g = 2.0 * R0_RING - 2.0 * rn
        # u = curl psi
        u_rho = 2.0 * px * f * E
        u_x = f / rn + f * E * g
        # omega_theta = d(u_rho)/dx - d(u_x)/drho
        dur_dx = 2.0 * f * E - 4.0 * px ** 2 * f * E ** 2 + 8.0 * a * px ** 2 * f / Dn ** 3
        dux_drho = (f * E * g / rn - f / rn ** 2
                    + g ** 2 * f * (E ** 2 - 2.0 * a / Dn ** 3) - 2.0 * f * E)
        w_t = dur_dx - dux_drho
    z = np.zeros_like(px)
    u = np.stack([np.where(inside, u_x, 0.0),
                  np.where(inside, np.cos(theta) * u_rho, 0.0),
                  np.where(inside, np.sin(theta) * u_rho, 0.0)])
    w = np.stack([z, np.where(inside, -np.sin(theta) * w_t, 0.0),
                  np.where(inside, np.cos(theta) * w_t, 0.0)])
    return u, w


def ring_fields(xmin, dx, ncell, ng):
    """(vorticity, velocity) as (3, Mx, My, Mz) arrays over the mesh including
    ghosts, component first (the reference layout, Fortran order when
    flattened with order='F' after a transpose is not needed: index [c,i,j,k])."""
    xs = [cell_centres(xmin[d], dx, ncell[d], ng) for d in range(3)]
    X, Y, Z = np.meshgrid(*xs, indexing="ij")
    u, w = _fields(X, Y, Z)
    return w, u


def rms_velocity_error(vel, u_exact, ng):
    """rms relative velocity error over the interior cells
    (validation/poisson/sources/poisson.f:285-431): sqrt(sum|u-ue|^2 / sum|ue|^2)."""
    sl = (slice(None),) + tuple(slice(ng, s - ng) for s in vel.shape[1:])
    d = vel[sl] - u_exact[sl]
    return float(np.sqrt(np.sum(d * d) / np.sum(u_exact[sl] ** 2)))
