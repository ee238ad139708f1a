"""ctypes binding of libpmgpu.so (the C ABI declared in include/pmgpu.h).

The library is built in place by ``make -f pmlib3d_amd/Makefile`` (or
``__graft_entry__.build()``).  There is no fallback: if the shared library is
missing or a call fails, an error is raised.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpmgpu.so")

FP64 = 0
FP32 = 1


class PmgpuError(RuntimeError):
    pass


class Geom(C.Structure):
    _fields_ = [("ncell", C.c_int32 * 3), ("nghost", C.c_int32), ("bound_cond", C.c_int32 * 3),
                ("interp_order", C.c_int32), ("xmin", C.c_double * 3), ("xmax", C.c_double * 3),
                ("dx", C.c_double * 3)]


class Params(C.Structure):
    _fields_ = [("regularisation_radius", C.c_double), ("poisson_order", C.c_int32),
                ("poisson_kernel", C.c_int32)]


class PenalParams(C.Structure):
    _fields_ = [("relaxation", C.c_double), ("tolerance", C.c_double), ("max_iter", C.c_int32),
                ("fd_order", C.c_int32), ("dtime", C.c_double), ("vel_inf", C.c_double * 3),
                ("reproject", C.c_int32), ("fix_free_stream", C.c_int32)]


# name -> (restype, argtypes); every exported symbol of include/pmgpu.h
_P = C.c_void_p
_I = C.c_int32
SIGNATURES = {
    "pmgpu_init": (C.c_int, [_I]),
    "pmgpu_finalize": (C.c_int, []),
    "pmgpu_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "pmgpu_mesh_extent": (C.c_int, [C.POINTER(Geom), C.POINTER(C.c_int32)]),
    "pmgpu_poisson_create": (C.c_int, [C.POINTER(Geom), C.POINTER(Params), _I, _P, C.POINTER(_P)]),
    "pmgpu_poisson_destroy": (C.c_int, [_P]),
    "pmgpu_poisson_setup": (C.c_int, [_P]),
    "pmgpu_poisson_set_level": (C.c_int, [_P, _I, _I]),
    "pmgpu_topology_parent": (C.c_int, [C.POINTER(Geom), _I, _I, C.POINTER(Geom)]),
    "pmgpu_interp_mesh_mesh": (C.c_int, [C.POINTER(Geom), C.POINTER(Geom), _P, _P, _I]),
    "pmgpu_interp_mesh_mesh_device": (C.c_int, [C.POINTER(Geom), C.POINTER(Geom), _I, _P, _P, _I, _P]),
    "pmgpu_poisson_solve": (C.c_int, [_P, _P, _P, _P, _P, _I, _I, _I]),
    "pmgpu_poisson_solve_device": (C.c_int, [_P, _P, _P, _P, _P, _I, _I, _I]),
    "pmgpu_poisson_kernel_octant": (C.c_int, [_P, _P, C.c_int64, C.POINTER(C.c_int32)]),
    "pmgpu_poisson_last_timing": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "pmgpu_poisson_set_timing": (C.c_int, [_P, _I]),
    "pmgpu_poisson_tune_scalar": (C.c_int, [_P]),
    "pmgpu_poisson_solve_scalar": (C.c_int, [_P, _P, _P]),
    "pmgpu_poisson_solve_scalar_device": (C.c_int, [_P, _P, _P]),
    "pmgpu_poisson_batch_create": (C.c_int, [C.POINTER(_P), _I, _P, C.POINTER(_P)]),
    "pmgpu_poisson_batch_solve_scalar": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "pmgpu_poisson_batch_destroy": (C.c_int, [_P]),
    "pmgpu_interp_particle_mesh": (C.c_int, [C.POINTER(Geom), _P, _P, C.c_int64, _P, _I, _I]),
    "pmgpu_interp_particle_mesh_device": (C.c_int, [C.POINTER(Geom), _I, _P, _P, C.c_int64, _P, _I, _I, _P]),
    "pmgpu_interp_mesh_particle": (C.c_int, [C.POINTER(Geom), _P, _P, C.c_int64, _P]),
    "pmgpu_interp_mesh_particle_device": (C.c_int, [C.POINTER(Geom), _I, _P, _P, C.c_int64, _P, _P]),
    "pmgpu_remesh": (C.c_int, [C.POINTER(Geom), _P, _P, _P, _P, C.c_double, C.c_double, _I, _I, _I,
                               _P, _P, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "pmgpu_remesh_device": (C.c_int, [C.POINTER(Geom), _I, _P, _P, _P, _P, C.c_double, C.c_double, _I, _I,
                                      _I, _P, _P, _P, _P, _P, _P, C.c_int64, C.POINTER(C.c_int64), _P]),
    "pmgpu_particles_advance": (C.c_int, [C.POINTER(Geom), _I, _I, C.c_double, _I, _P, _P, _P, _P, _P,
                                          _P, C.c_int64, C.POINTER(C.c_int64)]),
    "pmgpu_particles_advance_device": (C.c_int, [C.POINTER(Geom), _I, _I, C.c_double, _I,
                                                 C.POINTER(_P), C.c_int64, C.POINTER(C.c_int64), _P]),
    "pmgpu_particles_advance_device_spare": (C.c_int, [C.POINTER(Geom), _I, _I, C.c_double, _I, C.POINTER(_P),
                                                       C.POINTER(_P), C.c_int64, C.POINTER(C.c_int64),
                                                       C.POINTER(_I), _P]),
    "pmgpu_cell_index": (C.c_int, [C.POINTER(Geom), _P, C.c_int64, _P]),
    "pmgpu_penalise": (C.c_int, [C.POINTER(Geom), _P, _P, _P, _P, C.c_double, _I, _P, _P]),
    "pmgpu_penalise_device": (C.c_int, [C.POINTER(Geom), _I, _P, _P, _P, _P, C.c_double, _I, _P, _P, _P]),
    "pmgpu_diagnostics_particles": (C.c_int, [_P, _P, _P, _P, C.c_int64, _P, _P, _P]),
    "pmgpu_diagnostics_mesh": (C.c_int, [C.POINTER(Geom), _I, _P, _P, _P, _P, _P]),
    "pmgpu_penalisation_solve": (C.c_int, [_P, C.POINTER(PenalParams), _P, _P, _P, _P, _P, _P,
                                           C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "pmgpu_penalisation_solve_device": (C.c_int, [_P, C.POINTER(PenalParams), _P, _P, _P, _P, _P, _P,
                                                  C.POINTER(C.c_int32), C.POINTER(C.c_double)]),
    "pmgpu_output_mesh": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(Geom), _I, _P, _P, _I]),
    "pmgpu_output_particles": (C.c_int, [C.c_char_p, _I, C.c_double, C.c_double, _P, _P, C.c_int64, _I]),
    "pmgpu_input_particles": (C.c_int, [C.c_char_p, C.c_double, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_double), _P, _P, C.c_int64, _I]),
    "pmgpu_vstep_create": (C.c_int, [C.POINTER(Geom), C.POINTER(Params), _I, _I, C.c_int64, C.POINTER(_P)]),
    "pmgpu_vstep_destroy": (C.c_int, [_P]),
    "pmgpu_vstep_load": (C.c_int, [_P, _P, _P, C.c_int64]),
    "pmgpu_vstep_step": (C.c_int, [_P, C.c_double, C.c_double, _I]),
    "pmgpu_vstep_sync": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "pmgpu_vstep_arrays": (C.c_int, [_P, C.POINTER(_P), C.POINTER(_P)]),
    "pmgpu_vstep_fetch": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "pmgpu_penalise_device_noalloc": (C.c_int, [C.POINTER(Geom), _I, _P, _P, _P, _P, C.c_double, _I, _P, _P,
                                                _P]),
    "pmgpu_vorticity_rhs": (C.c_int, [C.POINTER(Geom), _P, _P, C.c_double, _I, _P]),
    "pmgpu_vorticity_rhs_device": (C.c_int, [C.POINTER(Geom), _I, _P, _P, C.c_double, _I, _P, _P]),
}

_lib = None


def load():
    """Load libpmgpu.so and bind every entry point (raises if missing)."""
    global _lib
    if _lib is None:
        # torch ships its own HIP runtime (libamdhip64.so.7 + libhsa-runtime64)
        # and a process must hold only one: loading torch first lets
        # libpmgpu.so bind to that copy (same SONAME), so device pointers and
        # streams are shared.  Without torch the system runtime is used.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise PmgpuError("libpmgpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def last_error():
    buf = C.create_string_buffer(1024)
    load().pmgpu_last_error(buf, 1024)
    return buf.value.decode()


def check(rc, what=""):
    if rc != 0:
        raise PmgpuError("%s: %s" % (what, last_error()))
    return rc


_initialised = False


def init(device=0):
    global _initialised
    if not _initialised:
        check(load().pmgpu_init(device), "pmgpu_init")
        _initialised = True
