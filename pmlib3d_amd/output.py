"""On-disk formats: VTK mesh output and the particle checkpoint.

References: pmlib_output_mesh(outdir, tag, patch, topo, mesh, ierr)
(source/output/pmlib_output_mesh.f:5-188), output_particles(tag, topo,
part, ierr) (vortex_codes/vortex_structures/source/output/
output_particles.f:50-69) and the particle reading of input_particles
(vortex_codes/vortex_structures/source/input/input_particles.f:81-135,
200-230), one rank.  The files are written and read by libpmgpu.so (C ABI),
byte for byte in the reference's formats.
"""
import ctypes as C

import numpy as np

from . import _lib


def _b(s):
    return s.encode() if isinstance(s, str) else s


def pmlib_output_mesh(outdir, tag, patch, vort, vel, interp_order=None):
    """Writes outdir + tag + '.pvti' and outdir + tag + '_P00000.vti' from
    host (3, E...) FP64 meshes; returns ierr (0 or -1).  Host I/O only (no
    device is touched)."""
    g = patch.geom(interp_order)
    w = np.asfortranarray(vort, dtype=np.float64)
    u = np.asfortranarray(vel, dtype=np.float64)
    return _lib.load().pmgpu_output_mesh(_b(outdir), _b(tag), C.byref(g), _lib.FP64,
                                         w.ctypes.data_as(C.c_void_p), u.ctypes.data_as(C.c_void_p), 0)


def output_mesh_device(outdir, tag, patch, precision, d_vort, d_vel, interp_order=None):
    """The same from device meshes (addresses) in the given precision."""
    _lib.init()
    g = patch.geom(interp_order)
    _lib.check(_lib.load().pmgpu_output_mesh(_b(outdir), _b(tag), C.byref(g), int(precision), d_vort, d_vel, 1),
               "pmgpu_output_mesh")


def output_particles(path, itime, time, dtime, pos, vort):
    """The particle checkpoint (records itime + 1, time + dtime, npart, then
    pos(3), vort(3) per particle) from host (3, np) FP64 arrays."""
    pos = np.asfortranarray(pos, dtype=np.float64)
    vort = np.asfortranarray(vort, dtype=np.float64)
    _lib.check(_lib.load().pmgpu_output_particles(_b(path), int(itime), float(time), float(dtime),
                                                  pos.ctypes.data_as(C.c_void_p),
                                                  vort.ctypes.data_as(C.c_void_p), pos.shape[1], 0),
               "pmgpu_output_particles")


def output_particles_device(path, itime, time, dtime, d_pos, d_vort, np_):
    _lib.init()
    _lib.check(_lib.load().pmgpu_output_particles(_b(path), int(itime), float(time), float(dtime), d_pos, d_vort,
                                                  int(np_), 1), "pmgpu_output_particles")


def input_particles(path, trunc=0.0):
    """Reads a particle checkpoint: (itime, time, pos (3, np), vort (3, np),
    vort_max) keeping the particles with |vort| >= trunc in file order
    (host I/O only)."""
    L = _lib.load()
    it, t, n, vm = C.c_int32(0), C.c_double(0.0), C.c_int64(0), C.c_double(0.0)
    _lib.check(L.pmgpu_input_particles(_b(path), float(trunc), C.byref(it), C.byref(t), C.byref(n), C.byref(vm),
                                       None, None, 0, 0), "pmgpu_input_particles")
    pos = np.zeros((3, n.value), order="F")
    vort = np.zeros((3, n.value), order="F")
    _lib.check(L.pmgpu_input_particles(_b(path), float(trunc), C.byref(it), C.byref(t), C.byref(n), C.byref(vm),
                                       pos.ctypes.data_as(C.c_void_p), vort.ctypes.data_as(C.c_void_p), n.value, 0),
               "pmgpu_input_particles")
    return it.value, t.value, pos, vort, vm.value
