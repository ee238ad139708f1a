"""Multiresolution coupling: refinement patches and their parents.

References: pmlib_topology_parent(topo, topo_parent, ierr)
(source/topology/pmlib_topology_parent.f:5-100), pmlib_interp_mesh_mesh(
topo_from, topo_to, mesh_from, mesh_to, ierr, clear)
(source/interpolation/pmlib_interp_mesh_mesh.f:5-2078) and the
refinement-patch kernels of pmlib_poisson_setup
(source/poisson/pmlib_poisson_setup.f:446-528 and the other orders' blocks),
which a PoissonSolver on a patch with level != 1 builds
(pmgpu_poisson_set_level).  One rank; the work runs in libpmgpu.so (C ABI).
"""
import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from .core import Patch, mesh_extent, params


def pmlib_topology_parent(patch, interp_order=None):
    """The parent cuboid of a refinement patch as a Patch: the same box and
    boundary conditions, cells 2^(level - parent(1)) times larger, 2 (M'4) or
    3 (M6*) ghost layers; level and parent copied.  Raises on parent level 0
    or a cell count that is not a multiple of the ratio."""
    g = patch.geom(interp_order)
    out = _lib.Geom()
    _lib.check(_lib.load().pmgpu_topology_parent(C.byref(g), int(patch.level), int(patch.parent[0]),
                                                 C.byref(out)), "pmgpu_topology_parent")
    return dataclasses.replace(patch, ncell=tuple(out.ncell), dx=tuple(out.dx), nghost=int(out.nghost))


def pmlib_interp_mesh_mesh(patch_from, patch_to, mesh_from, mesh_to, clear=False, interp_order=None):
    """Interpolates a host (3, E...) FP64 mesh onto another patch's mesh in
    place (mesh_to, F-ordered float64) and returns it."""
    _lib.init()
    io = params.pmlib_interpolation_order if interp_order is None else interp_order
    gf, gt = patch_from.geom(io), patch_to.geom(io)
    mesh_from = np.asfortranarray(mesh_from, dtype=np.float64)
    assert mesh_from.shape == (3,) + mesh_extent(patch_from, io), mesh_from.shape
    assert mesh_to.shape == (3,) + mesh_extent(patch_to, io), mesh_to.shape
    assert mesh_to.dtype == np.float64 and mesh_to.flags.f_contiguous
    _lib.check(_lib.load().pmgpu_interp_mesh_mesh(C.byref(gf), C.byref(gt), mesh_from.ctypes.data_as(C.c_void_p),
                                                  mesh_to.ctypes.data_as(C.c_void_p), int(bool(clear))),
               "pmgpu_interp_mesh_mesh")
    return mesh_to


def interp_mesh_mesh_device(patch_from, patch_to, precision, d_from, d_to, clear=False, stream=None,
                            interp_order=None):
    """The same on device meshes (addresses) in the given precision, on a
    HIP stream (None: the null stream), without host synchronisation."""
    _lib.init()
    io = params.pmlib_interpolation_order if interp_order is None else interp_order
    gf, gt = patch_from.geom(io), patch_to.geom(io)
    _lib.check(_lib.load().pmgpu_interp_mesh_mesh_device(C.byref(gf), C.byref(gt), int(precision), d_from, d_to,
                                                         int(bool(clear)), stream), "pmgpu_interp_mesh_mesh_device")
