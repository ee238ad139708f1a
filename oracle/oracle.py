"""ctypes front end of the C oracle (oracle/pmoracle.c) and of the compiled
reference harness (oracle/_ref/ref_driver).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Arrays follow the reference layout: meshes are numpy arrays of shape
(3, Ex, Ey, Ez) indexed [c, i, j, k] (component first, ghosts included) and
stored in Fortran order, particles are (3, np) Fortran-ordered.
"""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libpmoracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
MPIEXEC = "/opt/conda/bin/mpiexec"


class Geom(C.Structure):
    _fields_ = [("ncell", C.c_int32 * 3), ("nghost", C.c_int32),
                ("bound_cond", C.c_int32 * 3), ("interp_order", C.c_int32),
                ("xmin", C.c_double * 3), ("xmax", C.c_double * 3), ("dx", C.c_double * 3)]


class Params(C.Structure):
    _fields_ = [("regularisation_radius", C.c_double), ("poisson_order", C.c_int32),
                ("poisson_kernel", C.c_int32), ("refine", C.c_double)]


def make_geom(ncell, nghost, xmin, dx, bound_cond=(0, 0, 0), interp_order=3):
    g = Geom()
    for d in range(3):
        g.ncell[d] = int(ncell[d])
        g.bound_cond[d] = int(bound_cond[d])
        g.xmin[d] = float(xmin[d])
        g.dx[d] = float(dx if np.isscalar(dx) else dx[d])
        g.xmax[d] = g.xmin[d] + g.ncell[d] * g.dx[d]
    g.nghost = int(nghost)
    g.interp_order = int(interp_order)
    return g


def make_params(order=4, kernel=1, eps_radius=2.0, refine=0.0):
    """refine: a = 2^(level - parent level) of a refinement patch (0: base level)."""
    p = Params()
    p.regularisation_radius = float(eps_radius)
    p.poisson_order = int(order)
    p.poisson_kernel = int(kernel)
    p.refine = float(refine)
    return p


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run make -f oracle/Makefile")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        G = C.POINTER(Geom)
        Pp = C.POINTER(Params)
        L.or_glassman.argtypes = [P, P, C.c_int, C.c_int]
        L.or_green_real.argtypes = [G, Pp, P]
        L.or_green_real.restype = C.c_int
        L.or_kernel_size.argtypes = [G, Pp]
        L.or_kernel_size.restype = C.c_int64
        L.or_poisson_setup.argtypes = [G, Pp, P]
        L.or_poisson_setup.restype = C.c_int
        L.or_poisson_solve.argtypes = [G, Pp, P, P, P, P, P, C.c_int, C.c_int, C.c_int]
        L.or_poisson_solve.restype = C.c_int
        L.or_cell_index.argtypes = [C.c_double, C.c_double, C.c_double]
        L.or_cell_index.restype = C.c_int32
        L.or_cell_indices.argtypes = [G, P, C.c_int64, P]
        L.or_p2m.argtypes = [G, P, P, C.c_int64, P, C.c_int]
        L.or_p2m.restype = C.c_int
        L.or_m2p.argtypes = [G, P, P, C.c_int64, P]
        L.or_m2p.restype = C.c_int
        L.or_remesh.argtypes = [G, P, P, P, P, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                P, P, P, P, C.c_int64]
        L.or_remesh.restype = C.c_int64
        L.or_advance.argtypes = [G, C.c_int, C.c_int, C.c_double, C.c_int, P, P, P, P, P, P, C.c_int64]
        L.or_advance.restype = C.c_int64
        L.or_penalise.argtypes = [G, P, P, P, C.c_double, C.c_int, P, P]
        L.or_penalise.restype = C.c_int
        L.or_vort_rhs.argtypes = [G, P, P, C.c_double, C.c_int, P]
        L.or_vort_rhs.restype = C.c_int
        L.or_topology_parent.argtypes = [G, C.c_int, C.c_int, G]
        L.or_topology_parent.restype = C.c_int
        L.or_interp_mesh_mesh.argtypes = [G, G, P, P, C.c_int]
        L.or_interp_mesh_mesh.restype = C.c_int
        L.or_diag_particles.argtypes = [P, P, P, P, C.c_int64, P]
        L.or_diag_mesh.argtypes = [G, P, P, P]
        L.or_diag_mesh.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def ng_side(g, d):
    """Ghost layers on each side of axis d (free: nghost; periodic: 2 for
    M'4, 3 for M6*)."""
    if g.bound_cond[d] == 0:
        return g.nghost
    return 3 if g.interp_order == 4 else 2


def mesh_extent(g):
    return tuple(g.ncell[d] + 2 * ng_side(g, d) for d in range(3))


def new_mesh(g):
    return np.zeros((3,) + mesh_extent(g), dtype=np.float64, order="F")


def glassman(x, sign=-1):
    u = np.array(x, dtype=np.complex128)
    w = np.empty_like(u)
    lib().or_glassman(_p(u), _p(w), len(u), sign)
    return u


def green_real(g, p):
    n = lib().or_kernel_size(g, p)
    G = np.zeros(n, dtype=np.float64)
    if lib().or_green_real(g, p, _p(G)):
        raise ValueError("bad kernel order")
    return G


def poisson_setup(g, p):
    n = lib().or_kernel_size(g, p)
    Gh = np.zeros(n, dtype=np.complex128)
    if lib().or_poisson_setup(g, p, _p(Gh)):
        raise ValueError("bad kernel order")
    return Gh


def poisson_solve(g, p, Ghat, vort, reg_vort=False, reproj=False, calc_vecpot=False):
    vort = np.asfortranarray(vort, dtype=np.float64)
    vel = new_mesh(g)
    vo = new_mesh(g) if (reg_vort or reproj) else None
    ps = new_mesh(g) if calc_vecpot else None
    e = lib().or_poisson_solve(g, p, _p(Ghat), _p(vort), _p(vel), _p(vo), _p(ps),
                               int(reg_vort), int(reproj), int(calc_vecpot))
    if e:
        raise ValueError("or_poisson_solve failed %d" % e)
    return vel, vo, ps


def cell_indices(g, pos):
    pos = np.asfortranarray(pos, dtype=np.float64)
    out = np.zeros(pos.shape, dtype=np.int32, order="F")
    lib().or_cell_indices(g, _p(pos), pos.shape[1], _p(out))
    return out


def p2m(g, pos, val, mesh=None, clear=True):
    pos = np.asfortranarray(pos, dtype=np.float64)
    val = np.asfortranarray(val, dtype=np.float64)
    if mesh is None:
        mesh = new_mesh(g)
    lib().or_p2m(g, _p(pos), _p(val), pos.shape[1], _p(mesh), int(clear))
    return mesh


def m2p(g, mesh, pos):
    pos = np.asfortranarray(pos, dtype=np.float64)
    mesh = np.asfortranarray(mesh, dtype=np.float64)
    out = np.zeros(pos.shape, dtype=np.float64, order="F")
    lib().or_m2p(g, _p(mesh), _p(pos), pos.shape[1], _p(out))
    return out


def remesh(g, vort_m, dvort_m=None, vel_m=None, min_vort=0.0, min_dvort=0.0, incl_ghost=False,
           mask=None, dvort=True, vel=True):
    """pos, vort, vel, dvort of the remeshed particles; mask is a boolean
    (Ex, Ey, Ez) array or None, dvort/vel the reference's copy flags."""
    vort_m = np.asfortranarray(vort_m, dtype=np.float64)
    dv = None if dvort_m is None else np.asfortranarray(dvort_m, dtype=np.float64)
    vl = None if vel_m is None else np.asfortranarray(vel_m, dtype=np.float64)
    mk = None if mask is None else np.asfortranarray(mask, dtype=np.uint8)
    args = (g, _p(vort_m), _p(dv), _p(vl), _p(mk), min_vort, min_dvort, int(dvort), int(vel),
            int(incl_ghost))
    n = lib().or_remesh(*args, None, None, None, None, 0)
    out = [np.zeros((3, n), order="F") for _ in range(4)]
    lib().or_remesh(*args, *[_p(a) for a in out], n)
    return out  # pos, vort, vel, dvort


def advance(g, rk_order, stage, dt, pos, vel, vort, dvort, vel_rk, dvort_rk, euler_vort=False):
    arrs = [np.array(a, dtype=np.float64, order="F", copy=True)
            for a in (pos, vel, vort, dvort, vel_rk, dvort_rk)]
    n = lib().or_advance(g, rk_order, stage, dt, int(euler_vort), *[_p(a) for a in arrs],
                         arrs[0].shape[1])
    return [a[:, :n] for a in arrs]


def penalise(g, vel, chi, vort, solid_vel=None, relax=1.0, fd_order=4):
    """One penalisation iteration: returns (vort + relax curl(chi (us - u)),
    diag = [sum_dforce(3), res_enst, sum_enst_pen])."""
    vel = np.asfortranarray(vel, dtype=np.float64)
    chi = np.asfortranarray(chi, dtype=np.float64)
    vort = np.array(vort, dtype=np.float64, order="F", copy=True)
    sv = None if solid_vel is None else np.asfortranarray(solid_vel, dtype=np.float64)
    diag = np.zeros(5)
    if lib().or_penalise(g, _p(vel), _p(chi), _p(sv), relax, int(fd_order), _p(vort), _p(diag)):
        raise ValueError("penalise needs fd_order 2 or 4 and fd_order/2 ghost layers")
    return vort, diag


def vort_rhs(g, vort, vel, visc, fd_order=4, dvort=None):
    """Vorticity right-hand side visc lap(w) + (w . grad) u on the interior
    cells, periodic ghosts mapped; free-side ghosts of dvort (zeros when not
    given) are left as they are."""
    vort = np.asfortranarray(vort, dtype=np.float64)
    vel = np.asfortranarray(vel, dtype=np.float64)
    out = np.zeros_like(vort, order="F") if dvort is None else np.array(dvort, dtype=np.float64, order="F", copy=True)
    if lib().or_vort_rhs(g, _p(vort), _p(vel), float(visc), int(fd_order), _p(out)):
        raise ValueError("vort_rhs needs fd_order 2 or 4 and fd_order/2 ghost layers")
    return out


def _ghost_fill_periodic(g, m):
    """Periodic forward ghost map of a (3, E...) mesh, x then y then z
    (mesh/pmlib_mesh_map_ghost.f, incl_edges = .FALSE., op = replace)."""
    E = mesh_extent(g)
    for d in range(3):
        if g.bound_cond[d] != 1:
            continue
        ng = (E[d] - g.ncell[d]) // 2
        N = g.ncell[d]
        for t in range(ng):
            lo = [slice(None)] * 4
            src = [slice(None)] * 4
            lo[d + 1], src[d + 1] = t, N + t
            m[tuple(lo)] = m[tuple(src)]
            lo[d + 1], src[d + 1] = ng + N + t, ng + t
            m[tuple(lo)] = m[tuple(src)]
    return m


def penalisation_solve(g, p, Ghat, vel, vort, chi, solid_vel=None, relaxation=1.0, tolerance=1e-3, max_iter=10,
                       fd_order=4, dtime=1.0, vel_inf=(0.0, 0.0, 0.0), reproject=False, fix_free_stream=False):
    """penalisation_solve(patch, topo_all, mesh, ierr)
    (vortex_codes/vortex_penalisation/source/penalisation/penalisation_solve.f:
    set-up :61-111, iteration :133-343, map back :348-365, final solve
    :372-398, free stream :405-414) on one rank with the penalisation mesh on
    the patch itself, composed of or_penalise and or_poisson_solve.  solid_vel
    None: the solid at rest.  Returns (vel, vort, force, iterations,
    residual_force)."""
    vel = np.asfortranarray(vel, dtype=np.float64)
    us = np.zeros_like(vel) if solid_vel is None else np.asfortranarray(solid_vel, dtype=np.float64)
    vel_pen = np.asfortranarray(us - vel)
    vort_pen = np.zeros_like(vel, order="F")
    u_pen = np.zeros_like(vel, order="F")
    c_1_dt = 1.0 / dtime
    force = np.zeros(3)
    residual = 0.0
    npen = 0
    while (residual > tolerance and npen < max_iter) or npen == 0:
        npen += 1
        # chi (vel_pen - u_pen): in the first iteration u_pen is zero
        vort_pen, d = penalise(g, u_pen, chi, vort_pen, solid_vel=vel_pen, relax=relaxation, fd_order=fd_order)
        u_pen = poisson_solve(g, p, Ghat, vort_pen)[0]
        dforce = np.array([-d[c] * g.dx[0] * g.dx[1] * g.dx[2] * c_1_dt for c in range(3)])
        force = force + dforce
        residual = float(np.sqrt(dforce[0] ** 2 + dforce[1] ** 2 + dforce[2] ** 2) /
                         np.sqrt(force[0] ** 2 + force[1] ** 2 + force[2] ** 2))
    vort = _ghost_fill_periodic(g, np.asfortranarray(np.asarray(vort, dtype=np.float64) + vort_pen))
    out_v, out_w, _ = poisson_solve(g, p, Ghat, vort, False, bool(reproject), False)
    if reproject:
        vort = out_w
    out_v = np.asfortranarray(out_v)
    out_v[0] = out_v[0] + vel_inf[0]
    out_v[1] = out_v[1] + vel_inf[1]
    if fix_free_stream:
        out_v[2] = out_v[2] + vel_inf[2]
    else:
        out_v[1] = out_v[1] + vel_inf[2]
    return out_v, vort, force, npen, residual


def topology_parent(g, level=2, parent_level=1):
    """pmlib_topology_parent of a one-rank cuboid (a new geometry)."""
    out = Geom()
    if lib().or_topology_parent(g, int(level), int(parent_level), C.byref(out)):
        raise ValueError("Number mesh cells does not match (or parent level 0)")
    return out


def interp_mesh_mesh(g_from, g_to, mesh_from, mesh_to=None, clear=True):
    """pmlib_interp_mesh_mesh of a (3, E...) mesh onto another geometry."""
    mesh_from = np.asfortranarray(mesh_from, dtype=np.float64)
    out = new_mesh(g_to) if mesh_to is None else np.array(mesh_to, dtype=np.float64, order="F", copy=True)
    lib().or_interp_mesh_mesh(g_from, g_to, _p(mesh_from), _p(out), int(bool(clear)))
    return out


DIAG_PARTICLE_NAMES = ("impulse_x", "impulse_y", "impulse_z", "ang_impulse_x", "ang_impulse_y", "ang_impulse_z",
                       "kin_energy", "enstrophy", "helicity", "circulation", "centroid_x", "centroid_y",
                       "centroid_z", "vort_radius", "vort_core", "vort_max")
DIAG_MESH_NAMES = ("kin_energy", "enstrophy", "helicity", "lamb", "vort_max")


def diag_particles(dx, pos, vort, vel):
    """Ring diagnostics of a particle set (diagnostics_ring.f), 16 values in
    the order of DIAG_PARTICLE_NAMES."""
    pos = np.asfortranarray(pos, dtype=np.float64)
    vort = np.asfortranarray(vort, dtype=np.float64)
    vel = np.asfortranarray(vel, dtype=np.float64)
    out = np.zeros(16)
    d = np.asarray(dx, dtype=np.float64)
    lib().or_diag_particles(_p(d), _p(pos), _p(vort), _p(vel), pos.shape[1], _p(out))
    return out


def diag_mesh(g, vel, vort):
    """Mesh energy diagnostics (diagnostics_energy_mesh.f), 5 values in the
    order of DIAG_MESH_NAMES."""
    vel = np.asfortranarray(vel, dtype=np.float64)
    vort = np.asfortranarray(vort, dtype=np.float64)
    out = np.zeros(5)
    lib().or_diag_mesh(g, _p(vel), _p(vort), _p(out))
    return out


# --------------------------------------------------------------------------
# compiled reference (oracle/_ref), run in this container only


def ref_available():
    return os.path.exists(REF_DRIVER)


def _fmt(v):
    if isinstance(v, bool):
        return ".true." if v else ".false."
    if isinstance(v, str):
        return "'%s'" % v
    if isinstance(v, (list, tuple)):
        return ",".join(_fmt(x) for x in v)
    if isinstance(v, float):
        return repr(v)
    return str(v)


def ref_run(job, inputs, nproc=1, timeout=600, driver=None):
    """Run the reference harness with namelist `job` (dict); `inputs` is a list
    of arrays written back to back (Fortran order) as the input file.
    `driver` replaces the harness binary (the same harness built on another
    implementation of the subroutines).  Returns (raw float64 output, stdout)."""
    drv = REF_DRIVER if driver is None else driver
    with tempfile.TemporaryDirectory() as td:
        inf = os.path.join(td, "in.bin")
        outf = os.path.join(td, "out.bin")
        with open(inf, "wb") as f:
            for a in inputs:
                np.asarray(a).ravel(order="F").tofile(f)
        job = dict(job, infile=inf, outfile=outf)
        nml = "&job " + ", ".join("%s=%s" % (k, _fmt(v)) for k, v in job.items()) + " /\n"
        cmd = [drv] if nproc == 1 else [MPIEXEC, "-n", str(nproc), drv]
        r = subprocess.run(cmd, input=nml, text=True, capture_output=True, timeout=timeout)
        if r.returncode != 0:
            raise RuntimeError("%s failed: %s %s" % (os.path.basename(drv), r.stdout[-2000:], r.stderr[-2000:]))
        raw = open(outf, "rb").read() if os.path.exists(outf) else b""
        return raw, r.stdout


def geom_job(g, p=None):
    j = dict(ncell=list(g.ncell), nghost=g.nghost, bound_cond=list(g.bound_cond),
             xmin=list(g.xmin), xmax=list(g.xmax), dx=list(g.dx), interp_order=g.interp_order)
    if p is not None:
        j.update(order=p.poisson_order, kernel=p.poisson_kernel, eps_radius=p.regularisation_radius)
    return j
