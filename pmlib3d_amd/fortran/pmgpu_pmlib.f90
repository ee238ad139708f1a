!-------------------------------------------------------------------------------
! pmgpu_pmlib.f90 -- the pmlib3d particle-mesh subroutines on the GPU
!
! Same names, argument lists, optional flags and ierr convention as the
! reference routines they replace, so a client switches by USE-ing this module
! instead of pmlib_mod_poisson / pmlib_mod_interpolation / pmlib_mod_remesh and
! the advance routine of pmlib_mod_particles:
!
!   pmlib_poisson_setup(patch,topo_all,mesh,ierr)         poisson/pmlib_poisson_setup.f:5
!   pmlib_poisson_solve(patch,topo_all,mesh,ierr,reg_vort,reproj,calc_vecpot)
!                                                          poisson/pmlib_poisson_solve.f:15
!   pmlib_interp_particle_mesh(topo,xpart,part,mesh,ierr,clear,clear_ghost)
!                                                          interpolation/pmlib_interp_particle_mesh.f:5
!   pmlib_interp_mesh_particle(topo,mesh,part,ierr,vort,dvort,vel,clear)
!                                                          interpolation/pmlib_interp_mesh_particle.f:5
!   pmlib_remesh(patch,topo,mesh,part,min_vort,min_dvort,ierr,dvort,vel,mask,incl_ghost)
!                                                          remesh/pmlib_remesh.f:5
!   pmlib_particles_advance(patch,topo,part,itimestage,dtime,ierr,euler_vort)
!                                                          particles/pmlib_particles_advance.f:5
! plus one Brinkman penalisation iteration on class_mesh's penalisation
! fields (vortex_codes/vortex_penalisation/source/penalisation/
! penalisation_solve.f:96-111,146-273):
!   pmlib_penalise(topo,mesh,relax,ierr,diag,fd_order)
!
! The derived types and module globals are pmlib3d's own (pmlib_mod_patch,
! pmlib_mod_topology, pmlib_mod_mesh, pmlib_mod_particles,
! pmlib_mod_parameters); the arithmetic runs in libpmgpu.so through
! pmgpu_iface.  One rank per GPU; the library is FP64 (MK = KIND(1D0)).
! The Green's function lives on the device, in a handle kept per patch
! geometry by this module, not in mesh%pkernel_fft.
!-------------------------------------------------------------------------------
MODULE pmgpu_pmlib
  USE, INTRINSIC :: ISO_C_BINDING
  USE pmlib_mod_parameters
  USE pmlib_mod_write
  USE pmlib_mod_patch, ONLY: class_patch
  USE pmlib_mod_topology, ONLY: class_topology, class_topology_all
  USE pmlib_mod_mesh, ONLY: class_mesh
  USE pmlib_mod_particles, ONLY: class_particles
  USE pmgpu_iface
  IMPLICIT NONE
  PRIVATE

  PUBLIC :: pmlib_poisson_setup, pmlib_poisson_solve, pmlib_interp_particle_mesh, &
          & pmlib_interp_mesh_particle, pmlib_remesh, pmlib_particles_advance, pmlib_penalise, &
          & pmgpu_release, pmgpu_patch_solver, mesh_ptr, geom_of, pmlib_output_mesh, pmlib_interp_mesh_mesh

  ! solver handles, one per (geometry, order, kernel, radius, refinement level)
  TYPE solver_entry
    TYPE(pmgpu_geom)   :: g
    TYPE(pmgpu_params) :: p
    INTEGER            :: level = 1, parent = 0
    TYPE(C_PTR)        :: h = C_NULL_PTR
  END TYPE solver_entry
  TYPE(solver_entry), DIMENSION(64), SAVE :: solvers
  INTEGER, SAVE :: nsolvers = 0
  LOGICAL, SAVE :: initialised = .FALSE.

CONTAINS

!------------------------------------------------------------------ helpers
  LOGICAL FUNCTION ensure_init(caller)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    ensure_init = .TRUE.
    IF (initialised) RETURN
    IF (MK .NE. C_DOUBLE) THEN
      CALL pmlib_write(mpi_rank, caller, 'libpmgpu needs MK = KIND(1D0) (build with -D__double)')
      ensure_init = .FALSE.
      RETURN
    END IF
    IF (pmgpu_init(0_C_INT32_T) .NE. 0) THEN
      CALL pmlib_write(mpi_rank, caller, TRIM(pmgpu_error_message()))
      ensure_init = .FALSE.
      RETURN
    END IF
    initialised = .TRUE.
  END FUNCTION ensure_init

  ! the one-rank geometry of a cuboid topology (ghost counts checked against
  ! the library's mesh layout)
  SUBROUTINE geom_of(t, g, ierr, caller)
    TYPE(class_topology), INTENT(IN) :: t
    TYPE(pmgpu_geom), INTENT(OUT)    :: g
    INTEGER, INTENT(OUT)             :: ierr
    CHARACTER(LEN=*), INTENT(IN)     :: caller
    INTEGER(C_INT32_T) :: ext(3)
    INTEGER :: d
    ierr = 0
    g%nghost = 0
    DO d = 1, 3
      g%ncell(d) = t%ncell(d)
      g%bound_cond(d) = t%bound_cond(2*d-1)
      g%xmin(d) = t%xmin(d)
      g%xmax(d) = t%xmax(d)
      g%dx(d) = t%dx(d)
      IF (t%bound_cond(2*d-1) .EQ. 0) g%nghost = t%nghost(2*d-1)
    END DO
    g%interp_order = pmlib_interpolation_order
    IF (g%interp_order .NE. 4) g%interp_order = 3
    IF (pmgpu_mesh_extent(g, ext) .NE. 0) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, TRIM(pmgpu_error_message()))
      RETURN
    END IF
    DO d = 1, 3
      IF (ext(d) .NE. t%ncell(d) + t%nghost(2*d-1) + t%nghost(2*d)) THEN
        ierr = -1
        CALL pmlib_write(mpi_rank, caller, 'ghost layers differ from the one-rank mesh layout')
        RETURN
      END IF
    END DO
  END SUBROUTINE geom_of

  SUBROUTINE fail(caller, ierr)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    INTEGER, INTENT(OUT) :: ierr
    ierr = -1
    CALL pmlib_write(mpi_rank, caller, TRIM(pmgpu_error_message()))
  END SUBROUTINE fail

  LOGICAL FUNCTION flag(f)
    LOGICAL, OPTIONAL, INTENT(IN) :: f
    flag = .FALSE.
    IF (PRESENT(f)) flag = f
  END FUNCTION flag

  INTEGER(C_INT32_T) FUNCTION ci(l)
    LOGICAL, INTENT(IN) :: l
    ci = 0
    IF (l) ci = 1
  END FUNCTION ci

  ! (re)allocate a mesh field with the cuboid's bounds
  SUBROUTINE ensure_mesh(f, t)
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: f
    TYPE(class_topology), INTENT(IN) :: t
    INTEGER :: lo(3), hi(3)
    lo = 1 - t%nghost(1:5:2)
    hi = t%ncell + t%nghost(2:6:2)
    IF (ASSOCIATED(f)) THEN
      IF (SIZE(f, 1) .EQ. 3 .AND. ALL(LBOUND(f) .EQ. (/1, lo/)) .AND. ALL(UBOUND(f) .EQ. (/3, hi/))) RETURN
      DEALLOCATE(f)
    END IF
    ALLOCATE(f(3, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3)))
    f = 0.0_MK
  END SUBROUTINE ensure_mesh

  TYPE(C_PTR) FUNCTION mesh_ptr(f)
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: f
    mesh_ptr = C_NULL_PTR
    IF (ASSOCIATED(f)) mesh_ptr = C_LOC(f(1, LBOUND(f, 2), LBOUND(f, 3), LBOUND(f, 4)))
  END FUNCTION mesh_ptr

  TYPE(C_PTR) FUNCTION part_ptr(a)
    REAL(MK), DIMENSION(:,:), POINTER :: a
    part_ptr = C_NULL_PTR
    IF (ASSOCIATED(a)) THEN
      IF (SIZE(a, 2) .GT. 0) part_ptr = C_LOC(a(1, LBOUND(a, 2)))
    END IF
  END FUNCTION part_ptr

  SUBROUTINE realloc_part(a, n)
    REAL(MK), DIMENSION(:,:), POINTER :: a
    INTEGER, INTENT(IN) :: n
    IF (ASSOCIATED(a)) DEALLOCATE(a)
    ALLOCATE(a(3, n))
  END SUBROUTINE realloc_part

  ! keep the first n columns of a particle array
  SUBROUTINE shrink_part(a, n)
    REAL(MK), DIMENSION(:,:), POINTER :: a
    INTEGER, INTENT(IN) :: n
    REAL(MK), DIMENSION(:,:), POINTER :: b
    IF (.NOT. ASSOCIATED(a)) RETURN
    IF (SIZE(a, 2) .EQ. n) RETURN
    ALLOCATE(b(3, n))
    IF (n .GT. 0) b = a(:, 1:n)
    DEALLOCATE(a)
    a => b
  END SUBROUTINE shrink_part

  INTEGER FUNCTION find_solver(g, p, level, parent)
    TYPE(pmgpu_geom), INTENT(IN)   :: g
    TYPE(pmgpu_params), INTENT(IN) :: p
    INTEGER, INTENT(IN)            :: level, parent
    INTEGER :: i
    find_solver = 0
    DO i = 1, nsolvers
      IF (solvers(i)%level .NE. level .OR. (level .NE. 1 .AND. solvers(i)%parent .NE. parent)) CYCLE
      IF (ALL(solvers(i)%g%ncell .EQ. g%ncell) .AND. solvers(i)%g%nghost .EQ. g%nghost .AND. &
        & ALL(solvers(i)%g%bound_cond .EQ. g%bound_cond) .AND. &
        & solvers(i)%g%interp_order .EQ. g%interp_order .AND. &
        & ALL(solvers(i)%g%xmin .EQ. g%xmin) .AND. ALL(solvers(i)%g%dx .EQ. g%dx) .AND. &
        & solvers(i)%p%poisson_order .EQ. p%poisson_order .AND. &
        & solvers(i)%p%poisson_kernel .EQ. p%poisson_kernel .AND. &
        & solvers(i)%p%regularisation_radius .EQ. p%regularisation_radius) THEN
        find_solver = i
        RETURN
      END IF
    END DO
  END FUNCTION find_solver

  SUBROUTINE current_params(p)
    TYPE(pmgpu_params), INTENT(OUT) :: p
    p%regularisation_radius = pmlib_regularisation_radius
    p%poisson_order = pmlib_poisson_order
    p%poisson_kernel = pmlib_poisson_kernel
  END SUBROUTINE current_params

!------------------------------------------------------------------ Poisson
  SUBROUTINE pmlib_poisson_setup(patch, topo_all, mesh, ierr)
    TYPE(class_patch), INTENT(IN)        :: patch
    TYPE(class_topology_all), INTENT(IN) :: topo_all
    TYPE(class_mesh), INTENT(INOUT)      :: mesh
    INTEGER, INTENT(OUT)                 :: ierr
    CHARACTER(LEN=19) :: caller = 'pmlib_poisson_setup'
    INTEGER :: i
    i = solver_index(patch, topo_all, .TRUE., ierr, caller)
  END SUBROUTINE pmlib_poisson_setup

  ! the solver entry of the patch for the current pmlib parameters: created
  ! on first use; set up (G-hat built) when new or when always_setup
  INTEGER FUNCTION solver_index(patch, topo_all, always_setup, ierr, caller)
    TYPE(class_patch), INTENT(IN)        :: patch
    TYPE(class_topology_all), INTENT(IN) :: topo_all
    LOGICAL, INTENT(IN)                  :: always_setup
    INTEGER, INTENT(OUT)                 :: ierr
    CHARACTER(LEN=*), INTENT(IN)         :: caller
    TYPE(pmgpu_geom)   :: g
    TYPE(pmgpu_params) :: p
    INTEGER :: i
    LOGICAL :: fresh
    ierr = 0
    solver_index = 0
    IF (.NOT. ensure_init(caller)) THEN
      ierr = -1
      RETURN
    END IF
    CALL geom_of(topo_all%cuboid(mpi_rank), g, ierr, caller)
    IF (ierr .NE. 0) RETURN
    g%nghost = patch%nghost
    CALL current_params(p)
    i = find_solver(g, p, patch%level, patch%parent(1))
    fresh = i .EQ. 0
    IF (fresh) THEN
      IF (nsolvers .EQ. SIZE(solvers)) THEN
        ierr = -1
        CALL pmlib_write(mpi_rank, caller, 'too many patch geometries')
        RETURN
      END IF
      nsolvers = nsolvers + 1
      i = nsolvers
      solvers(i)%g = g
      solvers(i)%p = p
      solvers(i)%level = patch%level
      solvers(i)%parent = patch%parent(1)
      IF (pmgpu_poisson_create(g, p, PMGPU_FP64, C_NULL_PTR, solvers(i)%h) .NE. 0) THEN
        nsolvers = nsolvers - 1
        CALL fail(caller, ierr)
        RETURN
      END IF
      ! a refinement patch solves with the difference kernel
      ! (poisson/pmlib_poisson_setup.f:446-528)
      IF (pmgpu_poisson_set_level(solvers(i)%h, INT(patch%level, C_INT32_T), &
        & INT(patch%parent(1), C_INT32_T)) .NE. 0) THEN
        CALL fail(caller, ierr)
        RETURN
      END IF
    END IF
    IF (fresh .OR. always_setup) THEN
      IF (pmgpu_poisson_setup(solvers(i)%h) .NE. 0) THEN
        CALL fail(caller, ierr)
        RETURN
      END IF
    END IF
    solver_index = i
  END FUNCTION solver_index

  ! The FP64 solver handle of a patch (created and set up on first use), for
  ! the drop-in modules that call the library directly.
  TYPE(C_PTR) FUNCTION pmgpu_patch_solver(patch, topo_all, ierr, caller)
    TYPE(class_patch), INTENT(IN)        :: patch
    TYPE(class_topology_all), INTENT(IN) :: topo_all
    INTEGER, INTENT(OUT)                 :: ierr
    CHARACTER(LEN=*), INTENT(IN)         :: caller
    INTEGER :: i
    pmgpu_patch_solver = C_NULL_PTR
    i = solver_index(patch, topo_all, .FALSE., ierr, caller)
    IF (ierr .EQ. 0) pmgpu_patch_solver = solvers(i)%h
  END FUNCTION pmgpu_patch_solver

  SUBROUTINE pmlib_poisson_solve(patch, topo_all, mesh, ierr, reg_vort, reproj, calc_vecpot)
    TYPE(class_patch), INTENT(IN)        :: patch
    TYPE(class_topology_all), INTENT(IN) :: topo_all
    TYPE(class_mesh), INTENT(INOUT)      :: mesh
    INTEGER, INTENT(OUT)                 :: ierr
    LOGICAL, INTENT(IN), OPTIONAL        :: reg_vort, reproj, calc_vecpot
    CHARACTER(LEN=19) :: caller = 'pmlib_poisson_solve'
    TYPE(pmgpu_geom)   :: g
    TYPE(pmgpu_params) :: p
    TYPE(C_PTR) :: vo, ps
    LOGICAL :: rv, rp, cv
    INTEGER :: i
    ierr = 0
    rv = flag(reg_vort)
    rp = flag(reproj)
    cv = flag(calc_vecpot) .AND. pmlib_poisson_kernel .EQ. 1
    CALL geom_of(topo_all%cuboid(mpi_rank), g, ierr, caller)
    IF (ierr .NE. 0) RETURN
    g%nghost = patch%nghost
    CALL current_params(p)
    i = find_solver(g, p, patch%level, patch%parent(1))
    IF (i .EQ. 0) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'pmlib_poisson_setup has not been called for this patch')
      RETURN
    END IF
    IF (.NOT. ASSOCIATED(mesh%vort)) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'mesh%vort is not allocated')
      RETURN
    END IF
    CALL ensure_mesh(mesh%vel, topo_all%cuboid(mpi_rank))
    vo = C_NULL_PTR
    ps = C_NULL_PTR
    ! the filtered / reprojected vorticity overwrites mesh%vort, as in the
    ! reference (fourier/pmlib_fourier_ifft.f:259-289)
    IF (rv .OR. rp) vo = mesh_ptr(mesh%vort)
    IF (cv) THEN
      CALL ensure_mesh(mesh%vec_pot, topo_all%cuboid(mpi_rank))
      ps = mesh_ptr(mesh%vec_pot)
    END IF
    IF (pmgpu_poisson_solve(solvers(i)%h, mesh_ptr(mesh%vort), mesh_ptr(mesh%vel), vo, ps, &
      & ci(rv), ci(rp), ci(cv)) .NE. 0) CALL fail(caller, ierr)
  END SUBROUTINE pmlib_poisson_solve

!------------------------------------------------------------ interpolation
  SUBROUTINE pmlib_interp_particle_mesh(topo, xpart, part, mesh, ierr, clear, clear_ghost)
    TYPE(class_topology), DIMENSION(:), POINTER, INTENT(IN) :: topo
    REAL(MK), DIMENSION(:,:), POINTER, INTENT(IN)           :: xpart
    REAL(MK), DIMENSION(:,:), POINTER, INTENT(IN)           :: part
    REAL(MK), DIMENSION(:,:,:,:), POINTER                   :: mesh
    INTEGER, INTENT(OUT)                                    :: ierr
    LOGICAL, INTENT(IN), OPTIONAL                           :: clear, clear_ghost
    CHARACTER(LEN=26) :: caller = 'pmlib_interp_particle_mesh'
    TYPE(pmgpu_geom) :: g
    INTEGER :: np
    ierr = 0
    IF (.NOT. ensure_init(caller)) THEN
      ierr = -1
      RETURN
    END IF
    IF (SIZE(part, 1) .NE. SIZE(mesh, 1)) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'Dimension of interpolation variables does not correspond.')
      RETURN
    END IF
    IF (SIZE(part, 1) .NE. 3) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'libpmgpu interpolates 3-component fields')
      RETURN
    END IF
    np = topo(mpi_rank)%npart
    IF (np .NE. SIZE(xpart, 2) .OR. np .NE. SIZE(part, 2)) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'Discrepancy in number of particles')
      RETURN
    END IF
    CALL geom_of(topo(mpi_rank), g, ierr, caller)
    IF (ierr .NE. 0) RETURN
    IF (pmgpu_interp_particle_mesh(g, part_ptr(xpart), part_ptr(part), INT(np, C_INT64_T), &
      & mesh_ptr(mesh), ci(flag(clear)), ci(flag(clear_ghost))) .NE. 0) CALL fail(caller, ierr)
  END SUBROUTINE pmlib_interp_particle_mesh

  SUBROUTINE pmlib_interp_mesh_particle(topo, mesh, part, ierr, vort, dvort, vel, clear)
    TYPE(class_topology), DIMENSION(:), POINTER, INTENT(IN) :: topo
    TYPE(class_mesh), INTENT(IN)                            :: mesh
    TYPE(class_particles), INTENT(INOUT)                    :: part
    INTEGER, INTENT(OUT)                                    :: ierr
    LOGICAL, OPTIONAL, INTENT(IN)                           :: vort, dvort, vel, clear
    CHARACTER(LEN=26) :: caller = 'pmlib_interp_mesh_particle'
    TYPE(pmgpu_geom) :: g
    INTEGER(C_INT64_T) :: np
    ierr = 0
    IF (.NOT. ensure_init(caller)) THEN
      ierr = -1
      RETURN
    END IF
    CALL geom_of(topo(mpi_rank), g, ierr, caller)
    IF (ierr .NE. 0) RETURN
    np = topo(mpi_rank)%npart
    ! each selected particle field is overwritten by the interpolated values
    ! (clear only zeroes it first: interpolation/pmlib_interp_mesh_particle.f:64-72)
    IF (flag(vort)) THEN
      IF (pmgpu_interp_mesh_particle(g, mesh_ptr(mesh%vort), part_ptr(part%pos), np, &
        & part_ptr(part%vort)) .NE. 0) THEN
        CALL fail(caller, ierr)
        RETURN
      END IF
    END IF
    IF (flag(dvort)) THEN
      IF (pmgpu_interp_mesh_particle(g, mesh_ptr(mesh%dvort), part_ptr(part%pos), np, &
        & part_ptr(part%dvort)) .NE. 0) THEN
        CALL fail(caller, ierr)
        RETURN
      END IF
    END IF
    IF (flag(vel)) THEN
      IF (pmgpu_interp_mesh_particle(g, mesh_ptr(mesh%vel), part_ptr(part%pos), np, &
        & part_ptr(part%vel)) .NE. 0) THEN
        CALL fail(caller, ierr)
        RETURN
      END IF
    END IF
  END SUBROUTINE pmlib_interp_mesh_particle

!------------------------------------------------------------------ multiresolution
  ! pmlib_interp_mesh_mesh(topo_from, topo_to, mesh_from, mesh_to, ierr, clear)
  ! (interpolation/pmlib_interp_mesh_mesh.f:5-2078) on one rank: between a
  ! patch mesh and its parent's (topologies from pmlib_topology_parent)
  SUBROUTINE pmlib_interp_mesh_mesh(topo_from, topo_to, mesh_from, mesh_to, ierr, clear)
    TYPE(class_topology), DIMENSION(:), POINTER, INTENT(IN) :: topo_from, topo_to
    REAL(MK), DIMENSION(:,:,:,:), POINTER                   :: mesh_from, mesh_to
    INTEGER, INTENT(OUT)                                    :: ierr
    LOGICAL, OPTIONAL, INTENT(IN)                           :: clear
    CHARACTER(LEN=22) :: caller = 'pmlib_interp_mesh_mesh'
    TYPE(pmgpu_geom) :: gf, gt
    ierr = 0
    IF (SIZE(mesh_from, 1) .NE. SIZE(mesh_to, 1)) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'Dimension of interpolation variables does not correspond.')
      RETURN
    END IF
    IF (SIZE(mesh_from, 1) .NE. 3) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'libpmgpu interpolates (3, ...) meshes')
      RETURN
    END IF
    IF (.NOT. ensure_init(caller)) THEN
      ierr = -1
      RETURN
    END IF
    CALL geom_of(topo_from(mpi_rank), gf, ierr, caller)
    IF (ierr .NE. 0) RETURN
    CALL geom_of(topo_to(mpi_rank), gt, ierr, caller)
    IF (ierr .NE. 0) RETURN
    IF (pmgpu_interp_mesh_mesh(gf, gt, mesh_ptr(mesh_from), mesh_ptr(mesh_to), ci(flag(clear))) .NE. 0) &
      & CALL fail(caller, ierr)
  END SUBROUTINE pmlib_interp_mesh_mesh

!------------------------------------------------------------------ remesh
  SUBROUTINE pmlib_remesh(patch, topo, mesh, part, min_vort, min_dvort, ierr, dvort, vel, mask, &
      & incl_ghost)
    TYPE(class_patch), INTENT(IN)                              :: patch
    TYPE(class_topology), DIMENSION(:), POINTER, INTENT(INOUT) :: topo
    TYPE(class_mesh), INTENT(IN)                               :: mesh
    TYPE(class_particles), INTENT(INOUT)                       :: part
    REAL(MK), INTENT(IN)                                       :: min_vort, min_dvort
    INTEGER, INTENT(OUT)                                       :: ierr
    LOGICAL, INTENT(IN), OPTIONAL                              :: dvort, vel, mask, incl_ghost
    CHARACTER(LEN=22) :: caller = 'pmlib_remesh_particles'
    TYPE(pmgpu_geom) :: g
    INTEGER(C_INT8_T), DIMENSION(:,:,:), ALLOCATABLE, TARGET :: mk8
    TYPE(C_PTR) :: mptr
    INTEGER(C_INT64_T) :: n
    INTEGER :: np
    ierr = 0
    IF (.NOT. ensure_init(caller)) THEN
      ierr = -1
      RETURN
    END IF
    CALL geom_of(topo(mpi_rank), g, ierr, caller)
    IF (ierr .NE. 0) RETURN
    mptr = C_NULL_PTR
    IF (flag(mask)) THEN
      IF (.NOT. ASSOCIATED(mesh%patch_mask)) THEN
        ierr = -1
        CALL pmlib_write(mpi_rank, caller, 'mesh%patch_mask is not allocated')
        RETURN
      END IF
      ALLOCATE(mk8(SIZE(mesh%patch_mask, 1), SIZE(mesh%patch_mask, 2), SIZE(mesh%patch_mask, 3)))
      mk8 = MERGE(1_C_INT8_T, 0_C_INT8_T, mesh%patch_mask)
      mptr = C_LOC(mk8)
    END IF
    IF (flag(vel) .AND. .NOT. ASSOCIATED(mesh%vel)) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'mesh%vel is not allocated')
      RETURN
    END IF
    ! count pass, then allocate the particles and create them
    IF (pmgpu_remesh(g, mesh_ptr(mesh%vort), mesh_ptr(mesh%dvort), mesh_ptr(mesh%vel), mptr, &
      & REAL(min_vort, C_DOUBLE), REAL(min_dvort, C_DOUBLE), ci(flag(dvort)), ci(flag(vel)), &
      & ci(flag(incl_ghost)), C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, 0_C_INT64_T, n) .NE. 0) THEN
      CALL fail(caller, ierr)
      RETURN
    END IF
    np = INT(n)
    CALL realloc_part(part%pos, np)
    CALL realloc_part(part%vort, np)
    CALL realloc_part(part%vel, np)
    CALL realloc_part(part%dvort, np)
    CALL realloc_part(part%vel_rk, np)
    CALL realloc_part(part%dvort_rk, np)
    topo(mpi_rank)%npart = np
    IF (np .EQ. 0) RETURN
    part%vel_rk = 0.0_MK
    part%dvort_rk = 0.0_MK
    IF (pmgpu_remesh(g, mesh_ptr(mesh%vort), mesh_ptr(mesh%dvort), mesh_ptr(mesh%vel), mptr, &
      & REAL(min_vort, C_DOUBLE), REAL(min_dvort, C_DOUBLE), ci(flag(dvort)), ci(flag(vel)), &
      & ci(flag(incl_ghost)), part_ptr(part%pos), part_ptr(part%vort), part_ptr(part%vel), &
      & part_ptr(part%dvort), n, n) .NE. 0) CALL fail(caller, ierr)
  END SUBROUTINE pmlib_remesh

!----------------------------------------------------------------- advance
  SUBROUTINE pmlib_particles_advance(patch, topo, part, itimestage, dtime, ierr, euler_vort)
    TYPE(class_patch), INTENT(IN)                            :: patch
    TYPE(class_topology), DIMENSION(:), POINTER, INTENT(INOUT) :: topo
    TYPE(class_particles), INTENT(INOUT)                     :: part
    INTEGER, INTENT(IN)                                      :: itimestage
    REAL(MK), INTENT(IN)                                     :: dtime
    INTEGER, INTENT(OUT)                                     :: ierr
    LOGICAL, INTENT(IN), OPTIONAL                            :: euler_vort
    CHARACTER(LEN=23) :: caller = 'pmlib_particles_advance'
    TYPE(pmgpu_geom) :: g
    INTEGER(C_INT64_T) :: n
    INTEGER :: np
    ierr = 0
    IF (.NOT. ensure_init(caller)) THEN
      ierr = -1
      RETURN
    END IF
    CALL geom_of(topo(mpi_rank), g, ierr, caller)
    IF (ierr .NE. 0) RETURN
    g%nghost = patch%nghost
    np = topo(mpi_rank)%npart
    IF (pmgpu_particles_advance(g, INT(pmlib_time_integration_order, C_INT32_T), &
      & INT(itimestage, C_INT32_T), REAL(dtime, C_DOUBLE), ci(flag(euler_vort)), part_ptr(part%pos), &
      & part_ptr(part%vel), part_ptr(part%vort), part_ptr(part%dvort), part_ptr(part%vel_rk), &
      & part_ptr(part%dvort_rk), INT(np, C_INT64_T), n) .NE. 0) THEN
      CALL fail(caller, ierr)
      RETURN
    END IF
    ! particles that left a free side are gone: keep the compacted prefix
    ! (vel and dvort are not carried by the map; their prefix is the update's)
    np = INT(n)
    CALL shrink_part(part%pos, np)
    CALL shrink_part(part%vel, np)
    CALL shrink_part(part%vort, np)
    CALL shrink_part(part%dvort, np)
    CALL shrink_part(part%vel_rk, np)
    CALL shrink_part(part%dvort_rk, np)
    topo(mpi_rank)%npart = np
  END SUBROUTINE pmlib_particles_advance

!------------------------------------------------------------- penalisation
  ! One iteration on the penalisation fields of class_mesh: mesh%vel is the
  ! flow velocity, mesh%solid(1,:,:,:) the penalisation factor chi, and
  ! mesh%solid_vel (if allocated) the solid velocity; mesh%vort gains
  ! relax curl(chi (solid_vel - vel)).  diag = sum_dforce(1:3), res_enst,
  ! sum_enst_pen of the iteration.
  SUBROUTINE pmlib_penalise(topo, mesh, relax, ierr, diag, fd_order)
    TYPE(class_topology), DIMENSION(:), POINTER, INTENT(IN) :: topo
    TYPE(class_mesh), INTENT(INOUT)                         :: mesh
    REAL(MK), INTENT(IN)                                    :: relax
    INTEGER, INTENT(OUT)                                    :: ierr
    REAL(MK), DIMENSION(5), INTENT(OUT), OPTIONAL, TARGET   :: diag
    INTEGER, INTENT(IN), OPTIONAL                           :: fd_order
    CHARACTER(LEN=14) :: caller = 'pmlib_penalise'
    TYPE(pmgpu_geom) :: g
    REAL(C_DOUBLE), DIMENSION(5), TARGET :: d
    INTEGER(C_INT32_T) :: fd
    ierr = 0
    IF (.NOT. ensure_init(caller)) THEN
      ierr = -1
      RETURN
    END IF
    CALL geom_of(topo(mpi_rank), g, ierr, caller)
    IF (ierr .NE. 0) RETURN
    IF (.NOT. (ASSOCIATED(mesh%vel) .AND. ASSOCIATED(mesh%solid) .AND. ASSOCIATED(mesh%vort))) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'mesh%vel, mesh%solid and mesh%vort must be allocated')
      RETURN
    END IF
    fd = INT(pmlib_fd_order, C_INT32_T)
    IF (PRESENT(fd_order)) fd = INT(fd_order, C_INT32_T)
    IF (pmgpu_penalise(g, mesh_ptr(mesh%vel), mesh_ptr(mesh%solid), mesh_ptr(mesh%solid_vel), &
      & C_NULL_PTR, REAL(relax, C_DOUBLE), fd, mesh_ptr(mesh%vort), C_LOC(d)) .NE. 0) THEN
      CALL fail(caller, ierr)
      RETURN
    END IF
    IF (PRESENT(diag)) diag = d
  END SUBROUTINE pmlib_penalise

!------------------------------------------------------------------ output
  ! The VTK image-data pair of the mesh's interior vorticity and velocity
  ! (Float32 cell data, appended raw), as pmlib_output_mesh writes it; one
  ! rank.
  SUBROUTINE pmlib_output_mesh(outdir, tag, patch, topo, mesh, ierr)
    CHARACTER(LEN=*), INTENT(IN)                            :: outdir
    CHARACTER(LEN=*), INTENT(IN)                            :: tag
    TYPE(class_patch), INTENT(IN)                           :: patch
    TYPE(class_topology), DIMENSION(:), POINTER             :: topo
    TYPE(class_mesh), INTENT(IN)                            :: mesh
    INTEGER, INTENT(OUT)                                    :: ierr
    CHARACTER(LEN=17) :: caller = 'pmlib_output_mesh'
    TYPE(pmgpu_geom) :: g
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: w, u
    ierr = 0
    CALL geom_of(topo(mpi_rank), g, ierr, caller)
    IF (ierr .NE. 0) RETURN
    ! origin and spacing are the patch's, as in the reference
    g%xmin = patch%xmin
    g%dx = patch%dx
    w => mesh%vort
    u => mesh%vel
    IF (.NOT. (ASSOCIATED(w) .AND. ASSOCIATED(u))) THEN
      ierr = -1
      CALL pmlib_write(mpi_rank, caller, 'mesh%vort and mesh%vel must be allocated')
      RETURN
    END IF
    IF (pmgpu_output_mesh(TRIM(outdir)//C_NULL_CHAR, TRIM(tag)//C_NULL_CHAR, g, PMGPU_FP64, mesh_ptr(w), &
      & mesh_ptr(u), 0_C_INT32_T) .NE. 0) CALL fail(caller, ierr)
  END SUBROUTINE pmlib_output_mesh

!----------------------------------------------------------------- release
  SUBROUTINE pmgpu_release(ierr)
    INTEGER, INTENT(OUT) :: ierr
    INTEGER :: i
    ierr = 0
    DO i = 1, nsolvers
      IF (pmgpu_poisson_destroy(solvers(i)%h) .NE. 0) ierr = -1
      solvers(i)%h = C_NULL_PTR
    END DO
    nsolvers = 0
    IF (initialised) THEN
      IF (pmgpu_finalize() .NE. 0) ierr = -1
    END IF
    initialised = .FALSE.
  END SUBROUTINE pmgpu_release

END MODULE pmgpu_pmlib
