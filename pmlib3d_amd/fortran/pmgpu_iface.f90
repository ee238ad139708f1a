!-------------------------------------------------------------------------------
! pmgpu_iface.f90 -- ISO_C_BINDING interfaces of libpmgpu.so (include/pmgpu.h)
!
! One interface per C entry point, same argument order; pmgpu_geom and
! pmgpu_params are interoperable derived types with the C struct layouts.
! Nothing here depends on pmlib3d: pmgpu_pmlib.f90 builds the reference's
! subroutine interfaces on top of it.
!-------------------------------------------------------------------------------
MODULE pmgpu_iface
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT32_T), PARAMETER :: PMGPU_FP64 = 0, PMGPU_FP32 = 1

  TYPE, BIND(C) :: pmgpu_geom
    INTEGER(C_INT32_T) :: ncell(3)
    INTEGER(C_INT32_T) :: nghost
    INTEGER(C_INT32_T) :: bound_cond(3)
    INTEGER(C_INT32_T) :: interp_order
    REAL(C_DOUBLE)     :: xmin(3)
    REAL(C_DOUBLE)     :: xmax(3)
    REAL(C_DOUBLE)     :: dx(3)
  END TYPE pmgpu_geom

  TYPE, BIND(C) :: pmgpu_params
    REAL(C_DOUBLE)     :: regularisation_radius
    INTEGER(C_INT32_T) :: poisson_order
    INTEGER(C_INT32_T) :: poisson_kernel
  END TYPE pmgpu_params

  TYPE, BIND(C) :: pmgpu_penal_params
    REAL(C_DOUBLE)     :: relaxation
    REAL(C_DOUBLE)     :: tolerance
    INTEGER(C_INT32_T) :: max_iter
    INTEGER(C_INT32_T) :: fd_order
    REAL(C_DOUBLE)     :: dtime
    REAL(C_DOUBLE)     :: vel_inf(3)
    INTEGER(C_INT32_T) :: reproject
    INTEGER(C_INT32_T) :: fix_free_stream
  END TYPE pmgpu_penal_params

  INTERFACE
    INTEGER(C_INT) FUNCTION pmgpu_init(device) BIND(C, NAME='pmgpu_init')
      IMPORT :: C_INT, C_INT32_T
      INTEGER(C_INT32_T), VALUE :: device
    END FUNCTION pmgpu_init

    INTEGER(C_INT) FUNCTION pmgpu_finalize() BIND(C, NAME='pmgpu_finalize')
      IMPORT :: C_INT
    END FUNCTION pmgpu_finalize

    INTEGER(C_INT) FUNCTION pmgpu_last_error(buf, len) BIND(C, NAME='pmgpu_last_error')
      IMPORT :: C_INT, C_CHAR, C_SIZE_T
      CHARACTER(KIND=C_CHAR), DIMENSION(*) :: buf
      INTEGER(C_SIZE_T), VALUE :: len
    END FUNCTION pmgpu_last_error

    INTEGER(C_INT) FUNCTION pmgpu_mesh_extent(g, ext) BIND(C, NAME='pmgpu_mesh_extent')
      IMPORT :: C_INT, C_INT32_T, pmgpu_geom
      TYPE(pmgpu_geom), INTENT(IN) :: g
      INTEGER(C_INT32_T), DIMENSION(3) :: ext
    END FUNCTION pmgpu_mesh_extent

    INTEGER(C_INT) FUNCTION pmgpu_poisson_create(g, p, precision, stream, handle) &
        & BIND(C, NAME='pmgpu_poisson_create')
      IMPORT :: C_INT, C_INT32_T, C_PTR, pmgpu_geom, pmgpu_params
      TYPE(pmgpu_geom), INTENT(IN) :: g
      TYPE(pmgpu_params), INTENT(IN) :: p
      INTEGER(C_INT32_T), VALUE :: precision
      TYPE(C_PTR), VALUE :: stream
      TYPE(C_PTR) :: handle
    END FUNCTION pmgpu_poisson_create

    INTEGER(C_INT) FUNCTION pmgpu_poisson_destroy(handle) BIND(C, NAME='pmgpu_poisson_destroy')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
    END FUNCTION pmgpu_poisson_destroy

    INTEGER(C_INT) FUNCTION pmgpu_poisson_setup(handle) BIND(C, NAME='pmgpu_poisson_setup')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: handle
    END FUNCTION pmgpu_poisson_setup

    INTEGER(C_INT) FUNCTION pmgpu_poisson_set_level(handle, level, parent_level) &
        & BIND(C, NAME='pmgpu_poisson_set_level')
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: handle
      INTEGER(C_INT32_T), VALUE :: level, parent_level
    END FUNCTION pmgpu_poisson_set_level

    INTEGER(C_INT) FUNCTION pmgpu_interp_mesh_mesh(g_from, g_to, from, to, clear) &
        & BIND(C, NAME='pmgpu_interp_mesh_mesh')
      IMPORT :: C_INT, C_INT32_T, C_PTR, pmgpu_geom
      TYPE(pmgpu_geom), INTENT(IN) :: g_from, g_to
      TYPE(C_PTR), VALUE :: from, to
      INTEGER(C_INT32_T), VALUE :: clear
    END FUNCTION pmgpu_interp_mesh_mesh

    INTEGER(C_INT) FUNCTION pmgpu_poisson_solve(handle, vort, vel, vort_out, vecpot, reg_vort, &
        & reproj, calc_vecpot) BIND(C, NAME='pmgpu_poisson_solve')
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: handle, vort, vel, vort_out, vecpot
      INTEGER(C_INT32_T), VALUE :: reg_vort, reproj, calc_vecpot
    END FUNCTION pmgpu_poisson_solve

    INTEGER(C_INT) FUNCTION pmgpu_interp_particle_mesh(g, xpart, part, np, mesh, clear, clear_ghost) &
        & BIND(C, NAME='pmgpu_interp_particle_mesh')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_PTR, pmgpu_geom
      TYPE(pmgpu_geom), INTENT(IN) :: g
      TYPE(C_PTR), VALUE :: xpart, part, mesh
      INTEGER(C_INT64_T), VALUE :: np
      INTEGER(C_INT32_T), VALUE :: clear, clear_ghost
    END FUNCTION pmgpu_interp_particle_mesh

    INTEGER(C_INT) FUNCTION pmgpu_interp_mesh_particle(g, mesh, pos, np, out) &
        & BIND(C, NAME='pmgpu_interp_mesh_particle')
      IMPORT :: C_INT, C_INT64_T, C_PTR, pmgpu_geom
      TYPE(pmgpu_geom), INTENT(IN) :: g
      TYPE(C_PTR), VALUE :: mesh, pos, out
      INTEGER(C_INT64_T), VALUE :: np
    END FUNCTION pmgpu_interp_mesh_particle

    INTEGER(C_INT) FUNCTION pmgpu_remesh(g, vort_m, dvort_m, vel_m, mask, min_vort, min_dvort, &
        & dvort_flag, vel_flag, incl_ghost, pos, vort, vel, dvort, cap, np_out) &
        & BIND(C, NAME='pmgpu_remesh')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR, pmgpu_geom
      TYPE(pmgpu_geom), INTENT(IN) :: g
      TYPE(C_PTR), VALUE :: vort_m, dvort_m, vel_m, mask, pos, vort, vel, dvort
      REAL(C_DOUBLE), VALUE :: min_vort, min_dvort
      INTEGER(C_INT32_T), VALUE :: dvort_flag, vel_flag, incl_ghost
      INTEGER(C_INT64_T), VALUE :: cap
      INTEGER(C_INT64_T) :: np_out
    END FUNCTION pmgpu_remesh

    INTEGER(C_INT) FUNCTION pmgpu_particles_advance(g, rk_order, stage, dt, euler_vort, pos, vel, &
        & vort, dvort, vel_rk, dvort_rk, np, np_out) BIND(C, NAME='pmgpu_particles_advance')
      IMPORT :: C_INT, C_INT32_T, C_INT64_T, C_DOUBLE, C_PTR, pmgpu_geom
      TYPE(pmgpu_geom), INTENT(IN) :: g
      INTEGER(C_INT32_T), VALUE :: rk_order, stage, euler_vort
      REAL(C_DOUBLE), VALUE :: dt
      TYPE(C_PTR), VALUE :: pos, vel, vort, dvort, vel_rk, dvort_rk
      INTEGER(C_INT64_T), VALUE :: np
      INTEGER(C_INT64_T) :: np_out
    END FUNCTION pmgpu_particles_advance

    INTEGER(C_INT) FUNCTION pmgpu_cell_index(g, pos, np, idx) BIND(C, NAME='pmgpu_cell_index')
      IMPORT :: C_INT, C_INT64_T, C_PTR, pmgpu_geom
      TYPE(pmgpu_geom), INTENT(IN) :: g
      TYPE(C_PTR), VALUE :: pos, idx
      INTEGER(C_INT64_T), VALUE :: np
    END FUNCTION pmgpu_cell_index

    INTEGER(C_INT) FUNCTION pmgpu_penalise(g, vel, chi, solid_vel, us_const, relax, fd_order, vort, &
        & diag) BIND(C, NAME='pmgpu_penalise')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR, pmgpu_geom
      TYPE(pmgpu_geom), INTENT(IN) :: g
      TYPE(C_PTR), VALUE :: vel, chi, solid_vel, us_const, vort, diag
      REAL(C_DOUBLE), VALUE :: relax
      INTEGER(C_INT32_T), VALUE :: fd_order
    END FUNCTION pmgpu_penalise
    ! penalisation_solve (vortex_codes/vortex_penalisation/source/
    ! penalisation/penalisation_solve.f:6) on host FP64 meshes
    INTEGER(C_INT) FUNCTION pmgpu_penalisation_solve(h, pp, chi, solid_vel, us_const, vel, vort, &
        & force, npen, residual) BIND(C, NAME='pmgpu_penalisation_solve')
      IMPORT :: C_INT, C_PTR, pmgpu_penal_params
      TYPE(C_PTR), VALUE :: h
      TYPE(pmgpu_penal_params), INTENT(IN) :: pp
      TYPE(C_PTR), VALUE :: chi, solid_vel, us_const, vel, vort, force, npen, residual
    END FUNCTION pmgpu_penalisation_solve
    ! pmlib_output_mesh (source/output/pmlib_output_mesh.f:5): the .pvti /
    ! .vti pair; outdir and tag NUL-terminated
    INTEGER(C_INT) FUNCTION pmgpu_output_mesh(outdir, tag, g, precision, vort, vel, on_device) &
        & BIND(C, NAME='pmgpu_output_mesh')
      IMPORT :: C_INT, C_INT32_T, C_PTR, C_CHAR, pmgpu_geom
      CHARACTER(KIND=C_CHAR), DIMENSION(*), INTENT(IN) :: outdir, tag
      TYPE(pmgpu_geom), INTENT(IN) :: g
      INTEGER(C_INT32_T), VALUE :: precision, on_device
      TYPE(C_PTR), VALUE :: vort, vel
    END FUNCTION pmgpu_output_mesh
    ! part_deriv_vorticity (vortex_codes/*/source/particle_derivatives/
    ! part_deriv_vorticity.f:57-208): dvort from vort and vel, (3, E...) meshes
    INTEGER(C_INT) FUNCTION pmgpu_vorticity_rhs(g, vort, vel, visc, fd_order, dvort) &
        & BIND(C, NAME='pmgpu_vorticity_rhs')
      IMPORT :: C_INT, C_INT32_T, C_DOUBLE, C_PTR, pmgpu_geom
      TYPE(pmgpu_geom), INTENT(IN) :: g
      TYPE(C_PTR), VALUE :: vort, vel, dvort
      REAL(C_DOUBLE), VALUE :: visc
      INTEGER(C_INT32_T), VALUE :: fd_order
    END FUNCTION pmgpu_vorticity_rhs
  END INTERFACE

CONTAINS

  ! the message of the last failed call, as a Fortran string
  FUNCTION pmgpu_error_message() RESULT(msg)
    CHARACTER(LEN=256) :: msg
    CHARACTER(KIND=C_CHAR), DIMENSION(257) :: buf
    INTEGER :: i, rc
    buf = C_NULL_CHAR
    rc = pmgpu_last_error(buf, INT(257, C_SIZE_T))
    msg = ' '
    DO i = 1, 256
      IF (buf(i) .EQ. C_NULL_CHAR) EXIT
      msg(i:i) = buf(i)
    END DO
  END FUNCTION pmgpu_error_message

END MODULE pmgpu_iface
