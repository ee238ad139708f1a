"""Host-side mirror of the reference types and Poisson subroutines.

References (paths in the pmlib3d tree):
  class_patch       source/pmlib_patch.f:15-27
  class_mesh        source/pmlib_mesh.f:15-36
  class_particles   source/pmlib_particles.f:14-28
  module globals    source/pmlib_parameters.f:46-66
  pmlib_poisson_setup / pmlib_poisson_solve
                    source/poisson/pmlib_poisson_setup.f:5-19,
                    source/poisson/pmlib_poisson_solve.f:15-34
Meshes are numpy float64 arrays of shape (3, Ex, Ey, Ez) indexed [c, i, j, k]
in Fortran order, i.e. exactly the reference's mesh%vort(1:3, 1-ng:N+ng, ...).
"""
import ctypes as C
import sys
from dataclasses import dataclass, field

import numpy as np

from . import _lib


@dataclass
class Parameters:
    """The library globals of pmlib_parameters.f:46-66 that the path reads."""
    pmlib_regularisation_radius: float = 0.0
    pmlib_poisson_kernel: int = -99
    pmlib_poisson_order: int = -99
    pmlib_interpolation_order: int = -99
    pmlib_time_integration_order: int = -99
    pmlib_fd_order: int = -99
    mpi_rank: int = 0


# module globals, as in pmlib_mod_parameters
params = Parameters()


@dataclass
class Patch:
    """class_patch (pmlib_patch.f:15-27)."""
    xmin: tuple = (-1.0, -1.0, -1.0)
    xmax: tuple = (1.0, 1.0, 1.0)
    ncell: tuple = (32, 32, 32)
    dx: tuple = (1.0 / 16, 1.0 / 16, 1.0 / 16)
    bound_cond: tuple = (0, 0, 0)
    nghost: int = 0
    fft_seq: tuple = (1, 2, 3)
    ptype: int = 1
    level: int = 1
    parent: tuple = (0, 0)

    def geom(self, interp_order=None):
        g = _lib.Geom()
        io = params.pmlib_interpolation_order if interp_order is None else interp_order
        for d in range(3):
            g.ncell[d] = int(self.ncell[d])
            g.bound_cond[d] = int(self.bound_cond[d])
            g.xmin[d] = float(self.xmin[d])
            g.xmax[d] = float(self.xmax[d])
            g.dx[d] = float(self.dx[d])
        g.nghost = int(self.nghost)
        g.interp_order = int(io if io in (3, 4) else 3)
        return g


def mesh_extent(patch, interp_order=None):
    """(Ex, Ey, Ez) of a mesh on one rank, ghosts included
    (mesh/pmlib_mesh_allocate.f:46-76, topology/pmlib_topology_cuboid.f:71-80)."""
    io = params.pmlib_interpolation_order if interp_order is None else interp_order
    out = []
    for d in range(3):
        ng = patch.nghost if patch.bound_cond[d] == 0 else (3 if io == 4 else 2)
        out.append(int(patch.ncell[d]) + 2 * ng)
    return tuple(out)


@dataclass
class Mesh:
    """class_mesh (pmlib_mesh.f:15-36); pkernel_fft holds the GPU solver,
    which owns G-hat on the device."""
    vel: np.ndarray = None
    vort: np.ndarray = None
    vec_pot: np.ndarray = None
    dvort: np.ndarray = None
    patch_mask: np.ndarray = None
    pkernel_fft: object = None

    @staticmethod
    def allocate(patch, vort=True, dvort=False, vel=True):
        ext = (3,) + mesh_extent(patch)
        m = Mesh()
        if vort:
            m.vort = np.zeros(ext, order="F")
        if dvort:
            m.dvort = np.zeros(ext, order="F")
        if vel:
            m.vel = np.zeros(ext, order="F")
        return m


@dataclass
class Particles:
    """class_particles (pmlib_particles.f:14-28): (3, np) arrays."""
    pos: np.ndarray = None
    vel: np.ndarray = None
    vort: np.ndarray = None
    dvort: np.ndarray = None
    vel_rk: np.ndarray = None
    dvort_rk: np.ndarray = None


def pmlib_write(rank, caller, mesg):
    """pmlib_write.f:35-43 message format."""
    sys.stdout.write("[%d](%s) : %s\n" % (rank, caller, mesg))
    sys.stdout.flush()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class PoissonSolver:
    """Device-side state of one patch's solve (plans, buffers, G-hat)."""

    def __init__(self, patch, precision=_lib.FP64, order=None, kernel=None, eps_radius=None,
                 interp_order=None):
        _lib.init()
        self.lib = _lib.load()
        self.patch = patch
        self.precision = precision
        g = patch.geom(interp_order)
        p = _lib.Params()
        p.regularisation_radius = float(params.pmlib_regularisation_radius if eps_radius is None
                                        else eps_radius)
        p.poisson_order = int(params.pmlib_poisson_order if order is None else order)
        p.poisson_kernel = int(params.pmlib_poisson_kernel if kernel is None else kernel)
        h = C.c_void_p()
        _lib.check(self.lib.pmgpu_poisson_create(C.byref(g), C.byref(p), precision, None,
                                                 C.byref(h)), "pmgpu_poisson_create")
        self.h = h
        self.ext = mesh_extent(patch, interp_order)
        if int(patch.level) != 1:
            _lib.check(self.lib.pmgpu_poisson_set_level(h, int(patch.level), int(patch.parent[0])),
                       "pmgpu_poisson_set_level")

    def setup(self):
        _lib.check(self.lib.pmgpu_poisson_setup(self.h), "pmgpu_poisson_setup")

    def solve(self, vort, reg_vort=False, reproj=False, calc_vecpot=False):
        vort = np.asfortranarray(vort, dtype=np.float64)
        assert vort.shape == (3,) + self.ext, (vort.shape, self.ext)
        vel = np.zeros_like(vort, order="F")
        vo = np.zeros_like(vort, order="F") if (reg_vort or reproj) else None
        ps = np.zeros_like(vort, order="F") if calc_vecpot else None
        _lib.check(self.lib.pmgpu_poisson_solve(self.h, _ptr(vort), _ptr(vel), _ptr(vo), _ptr(ps),
                                                int(reg_vort), int(reproj), int(calc_vecpot)),
                   "pmgpu_poisson_solve")
        return vel, vo, ps

    def solve_device(self, d_vort, d_vel, d_vort_out=None, d_vecpot=None, reg_vort=False,
                     reproj=False, calc_vecpot=False):
        """Device pointers (ints, e.g. torch tensor.data_ptr()) in the handle's precision."""
        _lib.check(self.lib.pmgpu_poisson_solve_device(self.h, d_vort, d_vel, d_vort_out, d_vecpot,
                                                       int(reg_vort), int(reproj), int(calc_vecpot)),
                   "pmgpu_poisson_solve_device")

    def kernel_octant(self):
        dims = (C.c_int32 * 3)()
        itemsize = 8 if self.precision == _lib.FP64 else 4
        # query dims first with a zero-capacity call (fails but fills dims)
        self.lib.pmgpu_poisson_kernel_octant(self.h, C.c_void_p(1), 0, dims)
        n = dims[0] * dims[1] * dims[2]
        buf = np.zeros(n, dtype=np.float64 if itemsize == 8 else np.float32)
        _lib.check(self.lib.pmgpu_poisson_kernel_octant(self.h, _ptr(buf), buf.nbytes, dims),
                   "pmgpu_poisson_kernel_octant")
        return buf.reshape((dims[1], dims[0], dims[2]))  # [ky', kx, kz']

    def tune_scalar(self):
        _lib.check(self.lib.pmgpu_poisson_tune_scalar(self.h), "pmgpu_poisson_tune_scalar")

    def solve_scalar(self, w):
        """psi = G * w for one scalar field w (Ex, Ey, Ez), FP64 host arrays."""
        w = np.asfortranarray(w, dtype=np.float64)
        assert w.shape == self.ext, (w.shape, self.ext)
        psi = np.zeros_like(w, order="F")
        _lib.check(self.lib.pmgpu_poisson_solve_scalar(self.h, _ptr(w), _ptr(psi)), "pmgpu_poisson_solve_scalar")
        return psi

    def solve_scalar_device(self, d_w, d_psi):
        _lib.check(self.lib.pmgpu_poisson_solve_scalar_device(self.h, d_w, d_psi),
                   "pmgpu_poisson_solve_scalar_device")

    def set_timing(self, on=True):
        _lib.check(self.lib.pmgpu_poisson_set_timing(self.h, int(on)), "pmgpu_poisson_set_timing")

    def last_timing(self):
        a, b = C.c_double(), C.c_double()
        _lib.check(self.lib.pmgpu_poisson_last_timing(self.h, C.byref(a), C.byref(b)), "timing")
        return a.value, b.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                self.lib.pmgpu_poisson_destroy(self.h)
                self.h = None
        except Exception:
            pass


class PoissonBatch:
    """Several set-up solvers (patches of any sizes) whose single-component
    solves run as one captured HIP graph, one branch per patch."""

    def __init__(self, solvers, stream=None):
        self.lib = _lib.load()
        self.solvers = list(solvers)
        hs = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        b = C.c_void_p()
        _lib.check(self.lib.pmgpu_poisson_batch_create(hs, len(self.solvers), stream, C.byref(b)),
                   "pmgpu_poisson_batch_create")
        self.b = b

    def solve_scalar(self, d_ins, d_outs):
        """Device addresses of the input and output scalar meshes, one per
        patch, in the solvers' precision; stream-ordered."""
        n = len(self.solvers)
        ins = (C.c_void_p * n)(*d_ins)
        outs = (C.c_void_p * n)(*d_outs)
        _lib.check(self.lib.pmgpu_poisson_batch_solve_scalar(self.b, ins, outs), "pmgpu_poisson_batch_solve_scalar")

    def __del__(self):
        try:
            if getattr(self, "b", None):
                self.lib.pmgpu_poisson_batch_destroy(self.b)
                self.b = None
        except Exception:
            pass


def pmlib_poisson_setup(patch, mesh, precision=_lib.FP64):
    """pmlib_poisson_setup(patch, topo_all, mesh, ierr): returns ierr."""
    try:
        s = PoissonSolver(patch, precision)
        s.setup()
        mesh.pkernel_fft = s
        return 0
    except _lib.PmgpuError as e:
        pmlib_write(params.mpi_rank, "pmlib_poisson_setup", str(e))
        return -1


def pmlib_poisson_solve(patch, mesh, reg_vort=False, reproj=False, calc_vecpot=False):
    """pmlib_poisson_solve(patch, topo_all, mesh, ierr, reg_vort, reproj,
    calc_vecpot): fills mesh.vel (and mesh.vort when reg_vort/reproj,
    mesh.vec_pot when calc_vecpot with kernel 1); returns ierr."""
    try:
        s = mesh.pkernel_fft
        if s is None:
            raise _lib.PmgpuError("pmlib_poisson_setup has not been called")
        vel, vo, ps = s.solve(mesh.vort, reg_vort, reproj, calc_vecpot)
        mesh.vel = vel
        if vo is not None:
            mesh.vort = vo
        if ps is not None:
            mesh.vec_pot = ps
        return 0
    except _lib.PmgpuError as e:
        pmlib_write(params.mpi_rank, "pmlib_poisson_solve", str(e))
        return -1
