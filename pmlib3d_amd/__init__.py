"""pmlib3d_amd: MI355X-native particle-mesh hot path with the pmlib3d interface.

Python mirror of the reference's Fortran interface for the hot path
(the types class_patch / class_mesh / class_particles and the subroutines
pmlib_poisson_setup, pmlib_poisson_solve, pmlib_interp_particle_mesh,
pmlib_interp_mesh_particle, pmlib_remesh, pmlib_particles_advance, the
Brinkman penalisation step, the vorticity right-hand side and the
multiresolution coupling of a refinement patch with its parent).  Every routine runs the HIP kernels of
libpmgpu.so through its C ABI; subroutines return ierr (0 or -1) and write
"[rank](caller) : message" on failure, as pmlib_write does.
"""
from .core import (Patch, Mesh, Particles, Parameters, params, pmlib_write, mesh_extent,
                   pmlib_poisson_setup, pmlib_poisson_solve, PoissonSolver, PoissonBatch)
from .particles import (cell_index, pmlib_interp_particle_mesh, pmlib_interp_mesh_particle,
                        interp_mesh_field, pmlib_remesh, pmlib_particles_advance, ParticleOps)
from .penalisation import (penalise, penalise_device, penalise_device_noalloc, penalisation_solve_device,
                           sphere_mask, brinkman_factor)
from .derivatives import vorticity_rhs, vorticity_rhs_device, part_deriv_vorticity
from .step import VortexStep, GraphVortexStep
from .output import (pmlib_output_mesh, output_mesh_device, output_particles, output_particles_device,
                     input_particles)
from .diagnostics import (diagnostics_ring, diagnostics_energy_mesh, diagnostics_particles_device,
                          diagnostics_mesh_device)
from .multires import pmlib_topology_parent, pmlib_interp_mesh_mesh, interp_mesh_mesh_device
from ._lib import PmgpuError, FP32, FP64

__all__ = ["Patch", "Mesh", "Particles", "Parameters", "params", "pmlib_write", "mesh_extent",
           "pmlib_poisson_setup", "pmlib_poisson_solve", "PoissonSolver", "PoissonBatch", "PmgpuError",
           "FP32", "FP64", "cell_index", "pmlib_interp_particle_mesh", "pmlib_interp_mesh_particle",
           "interp_mesh_field", "pmlib_remesh", "pmlib_particles_advance", "ParticleOps", "penalise",
           "penalise_device", "penalise_device_noalloc", "penalisation_solve_device", "sphere_mask", "brinkman_factor", "vorticity_rhs", "vorticity_rhs_device",
           "part_deriv_vorticity", "VortexStep", "GraphVortexStep", "diagnostics_ring", "diagnostics_energy_mesh",
           "diagnostics_particles_device", "diagnostics_mesh_device", "pmlib_output_mesh", "output_mesh_device",
           "output_particles", "output_particles_device", "input_particles", "pmlib_topology_parent",
           "pmlib_interp_mesh_mesh", "interp_mesh_mesh_device"]
