"""Host mirror of the vorticity right-hand side (stretching and diffusion).

Reference: part_deriv_vorticity(topo, mesh, ierr)
(vortex_codes/vortex_penalisation/source/particle_derivatives/
part_deriv_vorticity.f:6, body :57-208): dvort = visc lap(w) + (w . grad) u
on the interior cells with central differences of order pmlib_fd_order
(pmlib_parameters.f:57), then the periodic ghost map.  The client's
viscosity visc becomes an argument.
"""
import ctypes as C

import numpy as np

from . import _lib
from .core import mesh_extent, params, pmlib_write


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _order(fd_order):
    o = params.pmlib_fd_order if fd_order is None else fd_order
    if o not in (2, 4):
        raise _lib.PmgpuError("pmlib_fd_order must be 2 or 4")
    return int(o)


def vorticity_rhs(patch, vort, vel, visc, fd_order=None, dvort=None, interp_order=None):
    """dvort (3, E...) from vort and vel (3, E...) meshes, FP64 host arrays;
    the free-side ghost cells of dvort keep the values of `dvort` (zeros when
    not given)."""
    _lib.init()
    g = patch.geom(interp_order)
    ext = (3,) + mesh_extent(patch, interp_order)
    vort = np.asfortranarray(vort, dtype=np.float64)
    vel = np.asfortranarray(vel, dtype=np.float64)
    out = np.zeros(ext, order="F") if dvort is None else np.array(dvort, dtype=np.float64, order="F", copy=True)
    if vort.shape != ext or vel.shape != ext or out.shape != ext:
        raise _lib.PmgpuError("vorticity_rhs: mesh shapes do not match the patch")
    _lib.check(_lib.load().pmgpu_vorticity_rhs(C.byref(g), _ptr(vort), _ptr(vel), float(visc), _order(fd_order),
                                               _ptr(out)), "pmgpu_vorticity_rhs")
    return out


def vorticity_rhs_device(patch, precision, d_vort, d_vel, d_dvort, visc, fd_order=None, stream=None,
                         interp_order=None):
    """Device form: addresses of (3, E...) meshes in the given precision;
    stream-ordered."""
    _lib.init()
    g = patch.geom(interp_order)
    _lib.check(_lib.load().pmgpu_vorticity_rhs_device(C.byref(g), int(precision), d_vort, d_vel, float(visc),
                                                      _order(fd_order), d_dvort, stream),
               "pmgpu_vorticity_rhs_device")


def part_deriv_vorticity(patch, mesh, visc, interp_order=None):
    """mesh.dvort from mesh.vort and mesh.vel (allocated when missing);
    returns ierr (0, or -1 after a message)."""
    try:
        ext = (3,) + mesh_extent(patch, interp_order)
        dv = np.zeros(ext, order="F") if mesh.dvort is None else mesh.dvort
        mesh.dvort = vorticity_rhs(patch, mesh.vort, mesh.vel, visc, dvort=dv, interp_order=interp_order)
        return 0
    except _lib.PmgpuError as e:
        pmlib_write(params.mpi_rank, "part_deriv_vorticity", str(e))
        return -1
