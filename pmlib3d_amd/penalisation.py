"""Host mirror of one Brinkman penalisation iteration.

Reference: penalisation_solve(patch, topo_all, mesh, ierr)
(vortex_codes/vortex_penalisation/source/penalisation/penalisation_solve.f:6;
the iteration body :96-111 and the curl of :146-273).  The reference's module
globals (mesh_penal%solid, mesh_penal%solid_vel, penal_relaxation,
pmlib_fd_order) become arguments here.
"""
import ctypes as C

import numpy as np

from . import _lib
from .core import mesh_extent


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def penalise(patch, vel, chi, vort, solid_vel=None, us_const=None, relax=1.0, fd_order=4,
             interp_order=None):
    """vort + relax curl(chi (us - vel)) on the interior cells and the sums
    [sum_dforce(3), res_enst, sum_enst_pen]; vel, vort (3, E...), chi (E...)."""
    _lib.init()
    g = patch.geom(interp_order)
    ext = mesh_extent(patch, interp_order)
    vel = np.asfortranarray(vel, dtype=np.float64)
    chi = np.asfortranarray(chi, dtype=np.float64)
    out = np.array(vort, dtype=np.float64, order="F", copy=True)
    if vel.shape != (3,) + ext or out.shape != (3,) + ext or chi.shape != ext:
        raise _lib.PmgpuError("penalise: mesh shapes do not match the patch")
    sv = None if solid_vel is None else np.asfortranarray(solid_vel, dtype=np.float64)
    uc = None if us_const is None else (C.c_double * 3)(*[float(v) for v in us_const])
    diag = np.zeros(5)
    _lib.check(_lib.load().pmgpu_penalise(C.byref(g), _ptr(vel), _ptr(chi), _ptr(sv), uc, float(relax),
                                          int(fd_order), _ptr(out), _ptr(diag)), "pmgpu_penalise")
    return out, diag


def penalise_device(patch, precision, d_vel, d_chi, d_vort, d_solid_vel=None, us_const=None, relax=1.0,
                    fd_order=4, diag=True, stream=None, interp_order=None):
    """Device form (addresses of arrays in the given precision); returns the
    FP64 sums when diag, else None without synchronising."""
    _lib.init()
    g = patch.geom(interp_order)
    uc = None if us_const is None else (C.c_double * 3)(*[float(v) for v in us_const])
    out = np.zeros(5) if diag else None
    _lib.check(_lib.load().pmgpu_penalise_device(C.byref(g), int(precision), d_vel, d_chi, d_solid_vel, uc,
                                                 float(relax), int(fd_order), d_vort, _ptr(out), stream),
               "pmgpu_penalise_device")
    return out


def sphere_mask(patch, radius, centre=(0.0, 0.0, 0.0), width=None, interp_order=None):
    """Smoothed solid indicator chi_s = (1 - tanh((r - R) / w)) / 2 on every
    mesh cell (ghosts included), w = 2h by default."""
    ext = mesh_extent(patch, interp_order)
    axes = []
    for d in range(3):
        ng = (ext[d] - patch.ncell[d]) // 2
        i = np.arange(1 - ng, patch.ncell[d] + ng + 1, dtype=np.float64)
        axes.append(patch.xmin[d] + ((i - 1.0) + 0.5) * patch.dx[d])
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt((X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2)
    w = 2.0 * max(patch.dx) if width is None else width
    return np.asfortranarray(0.5 * (1.0 - np.tanh((r - radius) / w)))


def brinkman_factor(chi_s, lam, dt):
    """lambda dt chi / (1 + lambda dt chi): the implicit finite-lambda form of
    the penalisation factor that replaces relax * chi."""
    a = lam * dt * chi_s
    return a / (1.0 + a)


def penalise_device_noalloc(patch, precision, d_vel, d_chi, d_vort, d_diag=None, d_solid_vel=None, us_const=None,
                            relax=1.0, fd_order=4, stream=None, interp_order=None):
    """Allocation-free device form: no device allocation and no host
    synchronisation; the FP64 sums go to the device array d_diag (5
    doubles), if given.  Partial sums use the library's fixed per-GPU
    scratch (at most 262144 blocks of 64 x 4 cells); calls that overlap on
    one GPU race."""
    _lib.init()
    g = patch.geom(interp_order)
    uc = None if us_const is None else (C.c_double * 3)(*[float(v) for v in us_const])
    _lib.check(_lib.load().pmgpu_penalise_device_noalloc(C.byref(g), int(precision), d_vel, d_chi, d_solid_vel,
                                                         uc, float(relax), int(fd_order), d_vort, d_diag, stream),
               "pmgpu_penalise_device_noalloc")


def penalisation_solve_device(solver, d_chi, d_vel, d_vort, relaxation=1.0, tolerance=1e-3, max_iter=10,
                              fd_order=4, dtime=1.0, vel_inf=(0.0, 0.0, 0.0), reproject=False,
                              d_solid_vel=None, us_const=None, fix_free_stream=False):
    """penalisation_solve(patch, topo_all, mesh, ierr) on device meshes of the
    solver's patch and precision (vortex_codes/vortex_penalisation/source/
    penalisation/penalisation_solve.f:6): iterates penalisation and solve
    until the relative force increment is at most `tolerance` (at least once,
    at most `max_iter` times), adds the penalisation vorticity to d_vort,
    solves for d_vel and adds vel_inf (as the reference does, vel_inf[2] goes
    to the second component unless fix_free_stream).  The module globals
    penal_relaxation, penal_tolerance, penal_max, pmlib_fd_order, dtime,
    vel_inf and reproject are arguments.  Returns (force[3], iterations,
    last relative force increment)."""
    pp = _lib.PenalParams(float(relaxation), float(tolerance), int(max_iter), int(fd_order), float(dtime),
                          (C.c_double * 3)(*[float(v) for v in vel_inf]), int(bool(reproject)),
                          int(bool(fix_free_stream)))
    uc = None if us_const is None else (C.c_double * 3)(*[float(v) for v in us_const])
    force = np.zeros(3)
    npen = C.c_int32(0)
    res = C.c_double(0.0)
    _lib.check(_lib.load().pmgpu_penalisation_solve_device(solver.h, C.byref(pp), d_chi, d_solid_vel, uc, d_vel,
                                                           d_vort, _ptr(force), C.byref(npen), C.byref(res)),
               "pmgpu_penalisation_solve_device")
    return force, npen.value, res.value
