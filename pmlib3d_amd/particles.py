"""Host-side mirror of the reference's particle subroutines.

References (paths in the pmlib3d tree):
  pmlib_interp_particle_mesh  source/interpolation/pmlib_interp_particle_mesh.f:5-6
  pmlib_interp_mesh_particle  source/interpolation/pmlib_interp_mesh_particle.f:5-6
  pmlib_remesh                source/remesh/pmlib_remesh.f:5-6
  pmlib_particles_advance     source/particles/pmlib_particles_advance.f:5-6

Each subroutine returns ierr (0, or -1 after writing "[rank](caller) :
message" as pmlib_write does).  Host arrays are float64 numpy arrays in the
reference's layout: particles (3, np), meshes (3, Ex, Ey, Ez) in Fortran
order.  The work runs in the HIP kernels of libpmgpu.so; ParticleOps exposes
the device entry points for data that already lives in HBM.
"""
import ctypes as C

import numpy as np

from . import _lib
from .core import Particles, mesh_extent, params, pmlib_write


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    a = np.asfortranarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise _lib.PmgpuError("array shape %s, expected %s" % (a.shape, shape))
    return a


def _geom(patch, interp_order=None):
    _lib.init()
    return patch.geom(interp_order)


def _ext(patch, interp_order=None):
    return (3,) + mesh_extent(patch, interp_order)


def cell_index(patch, pos, interp_order=None):
    """NINT((pos - xmin) / dx) per particle and axis, int32 (3, np)
    (interpolation/pmlib_interp_particle_mesh.f:153-155)."""
    g = _geom(patch, interp_order)
    pos = _f64(pos)
    n = pos.shape[1]
    out = np.zeros((3, n), dtype=np.int32, order="F")
    _lib.check(_lib.load().pmgpu_cell_index(C.byref(g), _ptr(pos), n, _ptr(out)), "pmgpu_cell_index")
    return out


def pmlib_interp_particle_mesh(patch, xpart, part, mesh, clear=False, clear_ghost=False,
                               interp_order=None):
    """Deposit part (3, np) at xpart (3, np) onto mesh (3, E...) in place."""
    try:
        g = _geom(patch, interp_order)
        xpart = _f64(xpart)
        part = _f64(part, xpart.shape)
        if not (isinstance(mesh, np.ndarray) and mesh.dtype == np.float64 and mesh.flags.f_contiguous
                and mesh.shape == _ext(patch, interp_order)):
            raise _lib.PmgpuError("Dimension of interpolation variables does not correspond.")
        _lib.check(_lib.load().pmgpu_interp_particle_mesh(C.byref(g), _ptr(xpart), _ptr(part),
                                                          xpart.shape[1], _ptr(mesh), int(bool(clear)),
                                                          int(bool(clear_ghost))),
                   "pmgpu_interp_particle_mesh")
        return 0
    except _lib.PmgpuError as e:
        pmlib_write(params.mpi_rank, "pmlib_interp_particle_mesh", str(e))
        return -1


def interp_mesh_field(patch, field, pos, interp_order=None):
    """One mesh field (3, E...) interpolated to the particles: (3, np)."""
    g = _geom(patch, interp_order)
    field = _f64(field, _ext(patch, interp_order))
    pos = _f64(pos)
    out = np.zeros(pos.shape, order="F")
    _lib.check(_lib.load().pmgpu_interp_mesh_particle(C.byref(g), _ptr(field), _ptr(pos), pos.shape[1],
                                                      _ptr(out)), "pmgpu_interp_mesh_particle")
    return out


def pmlib_interp_mesh_particle(patch, mesh, part, vort=False, dvort=False, vel=False, clear=False,
                               interp_order=None):
    """Interpolate the selected mesh fields onto part (a Particles); each
    selected particle field is overwritten (clear only zeroes it first)."""
    try:
        for flag, name in ((vort, "vort"), (dvort, "dvort"), (vel, "vel")):
            if flag:
                src = getattr(mesh, name)
                if src is None:
                    raise _lib.PmgpuError("mesh%%%s is not allocated" % name)
                setattr(part, name, interp_mesh_field(patch, src, part.pos, interp_order))
        return 0
    except _lib.PmgpuError as e:
        pmlib_write(params.mpi_rank, "pmlib_interp_mesh_particle", str(e))
        return -1


def pmlib_remesh(patch, mesh, part, min_vort, min_dvort, dvort=False, vel=False, mask=False,
                 incl_ghost=False, interp_order=None):
    """Replace part's arrays with one particle per cell above threshold."""
    try:
        g = _geom(patch, interp_order)
        ext = _ext(patch, interp_order)
        vm = _f64(mesh.vort, ext)
        dvm = None if mesh.dvort is None else _f64(mesh.dvort, ext)
        velm = None if mesh.vel is None else _f64(mesh.vel, ext)
        if vel and velm is None:
            raise _lib.PmgpuError("mesh%vel is not allocated")
        mk = None
        if mask:
            if getattr(mesh, "patch_mask", None) is None:
                raise _lib.PmgpuError("mesh%patch_mask is not allocated")
            mk = np.asfortranarray(mesh.patch_mask, dtype=np.uint8)
            if mk.shape != ext[1:]:
                raise _lib.PmgpuError("mesh%patch_mask has the wrong shape")
        lib = _lib.load()
        n = C.c_int64()
        args = (C.byref(g), _ptr(vm), _ptr(dvm), _ptr(velm), _ptr(mk), float(min_vort), float(min_dvort),
                int(bool(dvort)), int(bool(vel)), int(bool(incl_ghost)))
        _lib.check(lib.pmgpu_remesh(*args, None, None, None, None, 0, C.byref(n)), "pmgpu_remesh")
        arrs = [np.zeros((3, n.value), order="F") for _ in range(4)]
        _lib.check(lib.pmgpu_remesh(*args, *[_ptr(a) for a in arrs], n.value, C.byref(n)), "pmgpu_remesh")
        part.pos, part.vort, part.vel, part.dvort = arrs
        part.vel_rk = np.zeros((3, n.value), order="F")
        part.dvort_rk = np.zeros((3, n.value), order="F")
        return 0
    except _lib.PmgpuError as e:
        pmlib_write(params.mpi_rank, "pmlib_remesh_particles", str(e))
        return -1


def pmlib_particles_advance(patch, part, itimestage, dtime, euler_vort=False, rk_order=None,
                            interp_order=None):
    """One low-storage RK stage of order pmlib_time_integration_order, then
    the periodic wrap / free-side removal; part's arrays are replaced.  The
    map reorders the arrays the reference's map carries (pos and vort, and
    above first order vel_rk and dvort_rk); vel and dvort keep the update's
    order (particles/pmlib_particles_advance.f:121-166)."""
    try:
        g = _geom(patch, interp_order)
        order = params.pmlib_time_integration_order if rk_order is None else rk_order
        arrs = [np.array(getattr(part, k), dtype=np.float64, order="F", copy=True)
                for k in ("pos", "vel", "vort", "dvort", "vel_rk", "dvort_rk")]
        npart = arrs[0].shape[1]
        for a in arrs:
            if a.shape != (3, npart):
                raise _lib.PmgpuError("Discrepancy in number of particles")
        n = C.c_int64()
        _lib.check(_lib.load().pmgpu_particles_advance(C.byref(g), int(order), int(itimestage), float(dtime),
                                                       int(bool(euler_vort)), *[_ptr(a) for a in arrs],
                                                       npart, C.byref(n)), "pmgpu_particles_advance")
        for k, a in zip(("pos", "vel", "vort", "dvort", "vel_rk", "dvort_rk"), arrs):
            setattr(part, k, np.asfortranarray(a[:, :n.value]))
        return 0
    except _lib.PmgpuError as e:
        pmlib_write(params.mpi_rank, "pmlib_particles_advance", str(e))
        return -1


class ParticleOps:
    """Device entry points for one patch; arguments are device addresses
    (ints, e.g. torch tensor.data_ptr()) and an optional hipStream_t (int).
    Particle arrays are float64; meshes are in `precision` (FP64 or FP32)."""

    def __init__(self, patch, interp_order=None, precision=_lib.FP64):
        self.g = _geom(patch, interp_order)
        self.lib = _lib.load()
        self.ext = mesh_extent(patch, interp_order)
        self.precision = precision

    def p2m(self, d_xpart, d_part, npart, d_mesh, clear=True, clear_ghost=False, stream=None):
        _lib.check(self.lib.pmgpu_interp_particle_mesh_device(C.byref(self.g), self.precision, d_xpart, d_part,
                                                              npart, d_mesh, int(clear), int(clear_ghost), stream),
                   "pmgpu_interp_particle_mesh_device")

    def m2p(self, d_mesh, d_pos, npart, d_out, stream=None):
        _lib.check(self.lib.pmgpu_interp_mesh_particle_device(C.byref(self.g), self.precision, d_mesh, d_pos,
                                                              npart, d_out, stream),
                   "pmgpu_interp_mesh_particle_device")

    def remesh(self, d_vort_m, d_dvort_m, d_vel_m, min_vort, min_dvort, out=None, cap=0, d_mask=None,
               dvort=False, vel=False, incl_ghost=False, stream=None):
        """out = (pos, vort, vel, dvort, vel_rk, dvort_rk) device addresses of
        (3, cap) arrays, or None to count only; returns the particle count."""
        n = C.c_int64()
        o = list(out) if out is not None else [None] * 6
        _lib.check(self.lib.pmgpu_remesh_device(C.byref(self.g), self.precision, d_vort_m, d_dvort_m, d_vel_m,
                                                d_mask,
                                                float(min_vort), float(min_dvort), int(dvort), int(vel),
                                                int(incl_ghost), *o, int(cap), C.byref(n), stream),
                   "pmgpu_remesh_device")
        return n.value

    def advance(self, rk_order, stage, dt, arrays, npart, euler_vort=False, stream=None, spare=None):
        """arrays = (pos, vel, vort, dvort, vel_rk, dvort_rk) device addresses;
        returns the new particle count.  With spare (a second such set, at
        least npart particles each), returns (count, in_spare): in_spare True
        when the reordered carried arrays were written to spare, which then
        holds the particle set."""
        n = C.c_int64()
        ptrs = (C.c_void_p * 6)(*arrays)
        if spare is None:
            _lib.check(self.lib.pmgpu_particles_advance_device(C.byref(self.g), int(rk_order), int(stage),
                                                               float(dt), int(euler_vort), ptrs, int(npart),
                                                               C.byref(n), stream), "pmgpu_particles_advance_device")
            return n.value
        sp = (C.c_void_p * 6)(*spare)
        flag = C.c_int32()
        _lib.check(self.lib.pmgpu_particles_advance_device_spare(C.byref(self.g), int(rk_order), int(stage),
                                                                 float(dt), int(euler_vort), ptrs, sp, int(npart),
                                                                 C.byref(n), C.byref(flag), stream),
                   "pmgpu_particles_advance_device_spare")
        return n.value, bool(flag.value)


__all__ = ["cell_index", "pmlib_interp_particle_mesh", "pmlib_interp_mesh_particle", "interp_mesh_field",
           "pmlib_remesh", "pmlib_particles_advance", "ParticleOps", "Particles"]
