!-------------------------------------------------------------------------------
! pmgpu_penalisation.f90 -- drop-in for the client module mod_penalisation
! (vortex_codes/vortex_penalisation/source/mod_penalisation.f) on the GPU.
!
! The same module name, the same penalisation classes (patch_penal,
! topo_penal, mesh_penal) and the same subroutine
!   penalisation_solve(patch,topo_all,mesh,ierr)
!                          penalisation/penalisation_solve.f:6
! reading the same globals: penal_relaxation, penal_tolerance, penal_max,
! dtime, vel_inf and reproject of the client's mod_parameters
! (mod_parameters.f:43-74) and pmlib_fd_order of pmlib_mod_parameters.  The
! client compiles its own mod_parameters first, then this file instead of
! mod_penalisation.f.
!
! On one rank the penalisation mesh must cover the patch: topo_penal's cuboid
! equals topo_all's.  mesh_penal%solid(1,:,:,:) is the penalisation factor
! chi and mesh_penal%solid_vel (if allocated) the solid velocity; mesh%vel and
! mesh%vort are updated as the reference updates them, and mesh_penal%vort is
! left cleared.  The penalisation mesh's own velocity fields (vel, vel_pen)
! are device work space here and are not written back.
!-------------------------------------------------------------------------------
MODULE mod_penalisation
  USE, INTRINSIC :: ISO_C_BINDING
  USE mod_parameters, ONLY: penal_relaxation, penal_tolerance, penal_max, dtime, vel_inf, reproject
  USE pmlib_mod_parameters
  USE pmlib_mod_write
  USE pmlib_mod_patch, ONLY: class_patch
  USE pmlib_mod_topology, ONLY: class_topology_all
  USE pmlib_mod_mesh, ONLY: class_mesh
  USE pmgpu_iface
  USE pmgpu_pmlib, ONLY: pmgpu_patch_solver, mesh_ptr
  IMPLICIT NONE

  TYPE(class_patch)                         :: patch_penal
  TYPE(class_topology_all)                  :: topo_penal
  TYPE(class_mesh), SAVE                    :: mesh_penal

  ! results of the last call (no counterpart in the reference, which keeps
  ! them local): the accumulated force, the iterations and the last
  ! relative force increment
  REAL(MK), DIMENSION(3), TARGET, SAVE      :: penal_force = 0.0_MK
  INTEGER(C_INT32_T), TARGET, SAVE          :: penal_iterations = 0
  REAL(MK), TARGET, SAVE                    :: penal_residual = 0.0_MK

CONTAINS

  SUBROUTINE penalisation_solve(patch, topo_all, mesh, ierr)
    TYPE(class_patch)                         :: patch
    TYPE(class_topology_all)                  :: topo_all
    TYPE(class_mesh)                          :: mesh
    INTEGER, INTENT(OUT)                      :: ierr
    CHARACTER(LEN=18) :: caller = 'penalisation'
    TYPE(pmgpu_penal_params) :: pp
    TYPE(C_PTR) :: h
    INTEGER :: d
    ierr = 0
    IF (.NOT. (ASSOCIATED(mesh%vel) .AND. ASSOCIATED(mesh%vort) .AND. ASSOCIATED(mesh_penal%solid))) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'mesh%vel, mesh%vort and mesh_penal%solid must be allocated')
      RETURN
    END IF
    DO d = 1, 3
      IF (topo_penal%cuboid(mpi_rank)%ncell(d) .NE. topo_all%cuboid(mpi_rank)%ncell(d) .OR. &
        & topo_penal%cuboid(mpi_rank)%xmin(d) .NE. topo_all%cuboid(mpi_rank)%xmin(d) .OR. &
        & topo_penal%cuboid(mpi_rank)%dx(d) .NE. topo_all%cuboid(mpi_rank)%dx(d)) THEN
        ierr = -1
        CALL pmlib_write(mpi_rank, caller, 'the penalisation mesh must cover the patch on this rank')
        RETURN
      END IF
    END DO
    h = pmgpu_patch_solver(patch, topo_all, ierr, caller)
    IF (ierr .NE. 0) RETURN
    pp%relaxation = REAL(penal_relaxation, C_DOUBLE)
    pp%tolerance = REAL(penal_tolerance, C_DOUBLE)
    pp%max_iter = INT(penal_max, C_INT32_T)
    pp%fd_order = INT(pmlib_fd_order, C_INT32_T)
    pp%dtime = REAL(dtime, C_DOUBLE)
    pp%vel_inf = REAL(vel_inf, C_DOUBLE)
    pp%reproject = 0
    IF (reproject) pp%reproject = 1
    pp%fix_free_stream = 0
    IF (pmgpu_penalisation_solve(h, pp, mesh_ptr(mesh_penal%solid), mesh_ptr(mesh_penal%solid_vel), C_NULL_PTR, &
      & mesh_ptr(mesh%vel), mesh_ptr(mesh%vort), C_LOC(penal_force), C_LOC(penal_iterations), &
      & C_LOC(penal_residual)) .NE. 0) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, TRIM(pmgpu_error_message()))
      RETURN
    END IF
    IF (ASSOCIATED(mesh_penal%vort)) mesh_penal%vort = 0.0_MK
  END SUBROUTINE penalisation_solve

END MODULE mod_penalisation
