"""Global diagnostics in FP64 on the GPU.

References (client code of the reference's vortex solvers):
diagnostics_ring(topo, part, ierr)
(vortex_codes/vortex_structures/source/diagnostics/diagnostics_ring.f:5,
sums :72-158, normalisations :166-226) and diagnostics_energy_mesh(topo,
mesh, ierr) (.../diagnostics/diagnostics_energy_mesh.f:5, :72-134), on one
rank.  Both return their totals as a dict (and, in the device forms, may
leave them in a device array without synchronising).
"""
import ctypes as C

import numpy as np

from . import _lib

PARTICLE_NAMES = ("impulse_x", "impulse_y", "impulse_z", "ang_impulse_x", "ang_impulse_y", "ang_impulse_z",
                  "kin_energy", "enstrophy", "helicity", "circulation", "centroid_x", "centroid_y", "centroid_z",
                  "vort_radius", "vort_core", "vort_max")
MESH_NAMES = ("kin_energy", "enstrophy", "helicity", "lamb", "vort_max")


def diagnostics_particles_device(dx, d_pos, d_vort, d_vel, np_, d_out=None, host=True, stream=None):
    """Ring diagnostics of np_ particles whose (3, np_) FP64 positions,
    vorticities and velocities are at the device addresses d_pos, d_vort,
    d_vel.  Returns the 16 totals (PARTICLE_NAMES order) when host, else
    None (the totals then only go to the device array d_out)."""
    _lib.init()
    h = (C.c_double * 3)(*[float(v) for v in dx])
    out = np.zeros(16) if host else None
    _lib.check(_lib.load().pmgpu_diagnostics_particles(h, d_pos, d_vort, d_vel, int(np_), d_out,
                                                       None if out is None else out.ctypes.data_as(C.c_void_p),
                                                       stream), "pmgpu_diagnostics_particles")
    return out


def diagnostics_ring(dx, pos, vort, vel):
    """diagnostics_ring on host (3, np) FP64 arrays: a dict of the 16
    totals (the arrays are copied to the GPU for the call)."""
    import torch
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).to(dev)
         for a in (pos, vort, vel)]
    out = diagnostics_particles_device(dx, *[x.data_ptr() for x in t], np.asarray(pos).shape[1])
    return dict(zip(PARTICLE_NAMES, out))


def diagnostics_mesh_device(patch, precision, d_vel, d_vort, d_out=None, host=True, stream=None,
                            interp_order=None):
    """Mesh energy diagnostics of (3, E...) device meshes in the given
    precision: the 5 totals (MESH_NAMES order) when host, else None."""
    _lib.init()
    g = patch.geom(interp_order)
    out = np.zeros(5) if host else None
    _lib.check(_lib.load().pmgpu_diagnostics_mesh(C.byref(g), int(precision), d_vel, d_vort, d_out,
                                                  None if out is None else out.ctypes.data_as(C.c_void_p),
                                                  stream), "pmgpu_diagnostics_mesh")
    return out


def diagnostics_energy_mesh(patch, vel, vort, interp_order=None):
    """diagnostics_energy_mesh on host (3, E...) FP64 meshes: a dict of the
    5 totals."""
    import torch
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))).to(dev)
         for a in (vel, vort)]
    out = diagnostics_mesh_device(patch, _lib.FP64, t[0].data_ptr(), t[1].data_ptr(), interp_order=interp_order)
    return dict(zip(MESH_NAMES, out))
