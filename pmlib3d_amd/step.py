"""Device-resident vortex time step built from the library's routines.

This is the caller side of the boundary: the client time loop of the
reference's vortex codes (vortex_codes/*/source/main.f) calls, per RK stage,
pmlib_interp_particle_mesh -> pmlib_poisson_solve -> (part_deriv_vorticity) ->
pmlib_interp_mesh_particle -> pmlib_particles_advance, then deposits again and
calls pmlib_remesh.
VortexStep runs that sequence with every array resident in HBM (torch
tensors as device memory, the HIP kernels through the C ABI); only the
particle count of the remesh crosses to the host.
"""
import numpy as np

from . import _lib
from .core import PoissonSolver, mesh_extent
from .derivatives import vorticity_rhs_device
from .particles import ParticleOps

_FIELDS = ("pos", "vel", "vort", "dvort", "vel_rk", "dvort_rk")


class VortexStep:
    def __init__(self, patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3, capacity=None,
                 device="cuda:0", precision=_lib.FP64, visc=None, fd_order=4):
        """precision selects the mesh side (deposition, solve, gather,
        remesh thresholds): FP64, or FP32 (compensated deposition sums:
        per-row partials, Kahan-added; FP64 gather sums).  Particles are FP64 either way.  visc: each stage also forms
        the vorticity right-hand side visc lap(w) + (w . grad) u on the mesh
        (finite differences of order fd_order; needs fd_order / 2 ghost
        layers) and interpolates it to the particles' dvort; None: dvort = 0
        (advection only)."""
        import torch
        self.torch = torch
        self.dev = torch.device(device)
        self.patch = patch
        self.rk_order = rk_order
        self.solver = PoissonSolver(patch, precision, order=order, kernel=1, eps_radius=eps_radius,
                                    interp_order=interp_order)
        self.solver.setup()
        self.ops = ParticleOps(patch, interp_order, precision)
        ext = mesh_extent(patch, interp_order)
        nm = 3 * int(np.prod(ext))
        self.ext = ext
        mdt = torch.float64 if precision == _lib.FP64 else torch.float32
        self.mesh_vort = torch.zeros(nm, dtype=mdt, device=self.dev)
        self.mesh_vel = torch.zeros(nm, dtype=mdt, device=self.dev)
        self.precision = precision
        self.interp_order = interp_order
        self.visc, self.fd_order = visc, fd_order
        # free-side ghost cells of the right-hand side mesh stay zero
        self.mesh_dvort = torch.zeros(nm, dtype=mdt, device=self.dev) if visc is not None else None
        self.cap = int(capacity if capacity is not None else np.prod(ext))
        self.buf = [{k: torch.zeros(3 * self.cap, dtype=torch.float64, device=self.dev) for k in _FIELDS}
                    for _ in range(2)]
        self.cur = 0
        self.np = 0

    @property
    def arrays(self):
        return self.buf[self.cur]

    def load(self, pos, vort):
        pos = np.asfortranarray(pos, dtype=np.float64)
        n = pos.shape[1]
        if n > self.cap:
            raise _lib.PmgpuError("more particles than the step's capacity")
        a = self.arrays
        for k in _FIELDS:
            a[k].zero_()
        a["pos"][:3 * n].copy_(self.torch.from_numpy(pos.ravel(order="F")))
        a["vort"][:3 * n].copy_(self.torch.from_numpy(np.asfortranarray(vort).ravel(order="F")))
        self.np = n

    def load_device(self, d_pos, d_vort, n):
        """The same as load from device tensors: flat (3, n) Fortran-ordered
        float64 positions and vorticities (device-to-device copies)."""
        if n > self.cap:
            raise _lib.PmgpuError("more particles than the step's capacity")
        a = self.arrays
        for k in _FIELDS:
            a[k].zero_()
        a["pos"][:3 * n].copy_(d_pos[:3 * n])
        a["vort"][:3 * n].copy_(d_vort[:3 * n])
        self.np = n

    def seed_from_mesh(self, d_mesh_vort, min_vort):
        """Particles from a vorticity mesh in the step's precision (3, E...):
        one at each interior cell centre with |omega| > min_vort, carrying
        the cell's value (pmlib_remesh, remesh/pmlib_remesh.f:84-226).
        Returns the particle count."""
        a = self.arrays
        n = self.ops.remesh(d_mesh_vort, None, None, min_vort, 0.0,
                            out=[a[k].data_ptr() for k in ("pos", "vort", "vel", "dvort", "vel_rk", "dvort_rk")],
                            cap=self.cap, dvort=False, vel=False)
        if n > self.cap:
            raise _lib.PmgpuError("remesh made %d particles, capacity %d" % (n, self.cap))
        self.np = n
        return n

    def fetch(self):
        a = self.arrays
        n = self.np
        return tuple(a[k][:3 * n].cpu().numpy().reshape((3, n), order="F") for k in ("pos", "vort", "vel"))

    def stage(self, s, dt):
        a = self.arrays
        n = self.np
        self.ops.p2m(a["pos"].data_ptr(), a["vort"].data_ptr(), n, self.mesh_vort.data_ptr(), clear=True)
        self.solver.solve_device(self.mesh_vort.data_ptr(), self.mesh_vel.data_ptr())
        self.ops.m2p(self.mesh_vel.data_ptr(), a["pos"].data_ptr(), n, a["vel"].data_ptr())
        # Without a right-hand side dvort stays zero in both buffer sets:
        # they start zeroed, load zeroes every field, remesh writes zero
        # dvort, the RK update only reads it and the map does not carry it.
        if self.visc is not None:
            vorticity_rhs_device(self.patch, self.precision, self.mesh_vort.data_ptr(), self.mesh_vel.data_ptr(),
                                 self.mesh_dvort.data_ptr(), self.visc, fd_order=self.fd_order,
                                 interp_order=self.interp_order)
            self.ops.m2p(self.mesh_dvort.data_ptr(), a["pos"].data_ptr(), n, a["dvort"].data_ptr())
        # the map writes a reordered set into the other buffer set, which then
        # becomes the current one (no copy back)
        spare = self.buf[1 - self.cur]
        ptrs = [a[k].data_ptr() for k in _FIELDS]
        sp = [spare[k].data_ptr() for k in _FIELDS]
        if self.visc is None:
            # a zero right-hand side: the update neither reads dvort nor
            # reads, writes or carries dvort_rk (both stay zero)
            ptrs[3] = ptrs[5] = sp[3] = sp[5] = None
        self.np, swapped = self.ops.advance(self.rk_order, s, dt, ptrs, n, spare=sp)
        if swapped:
            self.cur = 1 - self.cur

    def remesh(self, min_vort, min_dvort=0.0):
        a = self.arrays
        self.ops.p2m(a["pos"].data_ptr(), a["vort"].data_ptr(), self.np, self.mesh_vort.data_ptr(), clear=True)
        nxt = self.buf[1 - self.cur]
        out = [nxt[k].data_ptr() for k in ("pos", "vort", "vel", "dvort", "vel_rk", "dvort_rk")]
        if self.visc is None:
            out[3] = out[5] = None   # zero already and never written without a right-hand side
        n = self.ops.remesh(self.mesh_vort.data_ptr(), None, self.mesh_vel.data_ptr(), min_vort, min_dvort,
                            out=out, cap=self.cap, dvort=False, vel=True)
        if n > self.cap:
            raise _lib.PmgpuError("remesh made %d particles, capacity %d" % (n, self.cap))
        self.cur = 1 - self.cur
        self.np = n

    def step(self, dt, min_vort, min_dvort=0.0):
        for s in range(1, self.rk_order + 1):
            self.stage(s, dt)
        self.remesh(min_vort, min_dvort)


class GraphVortexStep:
    """The same vortex step with the particle count kept in device memory and
    the whole step (every stage's deposition, solve, gather, RK update and
    migration, then deposition and remesh) recorded once as a HIP graph and
    replayed: no host synchronisation inside a step (pmgpu_vstep_*; the
    client loop of vortex_codes/vortex_structures/source/main.f:453-505
    without the vorticity right-hand side)."""

    def __init__(self, patch, order=6, eps_radius=2.0, rk_order=2, interp_order=3, capacity=None,
                 precision=_lib.FP64, device="cuda:0"):
        import ctypes as C
        import torch
        self.torch, self.C = torch, C
        self.dev = torch.device(device)
        _lib.init()
        self.lib = _lib.load()
        g = patch.geom(interp_order)
        p = _lib.Params(float(eps_radius), int(order), 1)
        self.cap = int(capacity if capacity is not None else np.prod(mesh_extent(patch, interp_order)))
        h = C.c_void_p()
        _lib.check(self.lib.pmgpu_vstep_create(C.byref(g), C.byref(p), int(precision), int(rk_order), self.cap,
                                               C.byref(h)), "pmgpu_vstep_create")
        self.h = h
        arr = (C.c_void_p * 6)()
        st = C.c_void_p()
        _lib.check(self.lib.pmgpu_vstep_arrays(self.h, arr, C.byref(st)), "pmgpu_vstep_arrays")
        self.ptrs = [arr[i] for i in range(6)]
        self.stream = st.value

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.pmgpu_vstep_destroy(self.h)
            self.h = None

    def load(self, pos, vort):
        pos = np.asfortranarray(pos, dtype=np.float64)
        n = pos.shape[1]
        t = [self.torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)).to(self.dev)
             for a in (pos, vort)]
        self.torch.cuda.synchronize()
        _lib.check(self.lib.pmgpu_vstep_load(self.h, t[0].data_ptr(), t[1].data_ptr(), n), "pmgpu_vstep_load")

    def load_device(self, d_pos, d_vort, n):
        """The same as load from device tensors: flat (3, n) Fortran-ordered
        float64 positions and vorticities."""
        self.torch.cuda.synchronize()
        _lib.check(self.lib.pmgpu_vstep_load(self.h, d_pos.data_ptr(), d_vort.data_ptr(), int(n)),
                   "pmgpu_vstep_load")

    def step(self, dt, min_vort, graph=True):
        _lib.check(self.lib.pmgpu_vstep_step(self.h, float(dt), float(min_vort), 1 if graph else 0),
                   "pmgpu_vstep_step")

    def sync(self):
        n = self.C.c_int64(0)
        st = self.C.c_int32(0)
        _lib.check(self.lib.pmgpu_vstep_sync(self.h, self.C.byref(n), self.C.byref(st)), "pmgpu_vstep_sync")
        return n.value, st.value

    def fetch(self):
        """(pos, vort, vel) of the live particles as host (3, np) arrays."""
        n, st = self.sync()
        if st:
            raise _lib.PmgpuError("vortex step status %d (1: remesh above capacity, 2: a particle moved more "
                                  "than a period)" % st)
        out = [np.empty((3, n), order="F") for _ in range(3)]
        got = self.C.c_int64(0)
        _lib.check(self.lib.pmgpu_vstep_fetch(self.h, *[o.ctypes.data_as(self.C.c_void_p) for o in out], n,
                                              self.C.byref(got)), "pmgpu_vstep_fetch")
        return tuple(out)
