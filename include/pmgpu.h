/*
 * pmgpu.h -- C ABI of the MI355X (gfx950) particle-mesh hot path.
 *
 * Each entry point replaces one subroutine of the reference pmlib3d library
 * (Fortran 90 + MPI); the citation names the reference interface.  Arrays use
 * the reference's Fortran layout: meshes are (3, 1-ng:N+ng, ...) with the
 * component fastest and the ghost layers included, particles are (3, np).
 * Host-pointer entry points take FP64 arrays (MK = KIND(1D0)); the *_device
 * variants take device pointers in the handle's precision and run on the
 * handle's stream without host synchronisation.
 *
 * Every function returns 0 on success and -1 on failure; the message of the
 * last failure on the calling thread is read with pmgpu_last_error (the
 * reference's ierr = -1 + pmlib_write convention, pmlib_write.f:11-49).
 */
#ifndef PMGPU_H
#define PMGPU_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Patch geometry on one rank: the class_patch fields the path reads
 * (pmlib_patch.f:15-27) plus the interpolation order that fixes the ghost
 * count on periodic sides (topology/pmlib_topology_cuboid.f:71-80). */
typedef struct pmgpu_geom {
  int32_t ncell[3];      /* interior cells per axis                       */
  int32_t nghost;        /* exterior ghost layers on free-space sides     */
  int32_t bound_cond[3]; /* 0 free space, 1 periodic                      */
  int32_t interp_order;  /* 3 = M'4, 4 = M6*                              */
  double xmin[3];        /* interior minimum (patch%xmin)                 */
  double xmax[3];        /* interior maximum (patch%xmax)                 */
  double dx[3];          /* mesh spacing                                  */
} pmgpu_geom;

/* Library globals of pmlib_parameters.f:46-66 used by the solve. */
typedef struct pmgpu_params {
  double regularisation_radius; /* eps / max(dx)                          */
  int32_t poisson_order;        /* 0, 2, 4, 6, 8, 10                      */
  int32_t poisson_kernel;       /* 1 (G)                                  */
} pmgpu_params;

enum { PMGPU_FP64 = 0, PMGPU_FP32 = 1 };

typedef struct pmgpu_poisson_s *pmgpu_poisson_t;

/* Library state (no reference counterpart; MPI_INIT plays this role). */
int pmgpu_init(int32_t device);
int pmgpu_finalize(void);
int pmgpu_last_error(char *buf, size_t len);

/* Mesh extent (N + ghosts per axis) of a geometry
 * (mesh/pmlib_mesh_allocate.f:46-76). */
int pmgpu_mesh_extent(const pmgpu_geom *g, int32_t ext[3]);

/* Solver handle: owns device buffers, FFT plans and the kernel spectrum.
 * precision is PMGPU_FP64 or PMGPU_FP32 (FFTs and spectral products; the
 * Green's function is always built in FP64).  stream may be NULL. */
int pmgpu_poisson_create(const pmgpu_geom *g, const pmgpu_params *p, int32_t precision,
                         void *stream, pmgpu_poisson_t *out);
int pmgpu_poisson_destroy(pmgpu_poisson_t h);

/* Replaces pmlib_poisson_setup(patch, topo_all, mesh, ierr)
 * (poisson/pmlib_poisson_setup.f:5-19): builds G-hat on the device. */
int pmgpu_poisson_setup(pmgpu_poisson_t h);

/* Replaces pmlib_poisson_solve(patch, topo_all, mesh, ierr, reg_vort, reproj,
 * calc_vecpot) (poisson/pmlib_poisson_solve.f:15-34).  Host FP64 meshes:
 * vort (in), vel (out), vort_out (out when reg_vort or reproj, else may be
 * NULL; it may alias vort, as mesh%vort is overwritten in the reference),
 * vecpot (out when calc_vecpot, else may be NULL). */
int pmgpu_poisson_solve(pmgpu_poisson_t h, const double *vort, double *vel, double *vort_out,
                        double *vecpot, int32_t reg_vort, int32_t reproj, int32_t calc_vecpot);

/* Same, device-resident meshes in the handle's precision, on its stream. */
int pmgpu_poisson_solve_device(pmgpu_poisson_t h, const void *d_vort, void *d_vel,
                               void *d_vort_out, void *d_vecpot, int32_t reg_vort,
                               int32_t reproj, int32_t calc_vecpot);

/* Copy the kernel spectrum octant (FP64 or FP32 per the handle) to host:
 * Ghat[(ky' * Hx + kx) * Hz + kz'], sizes returned in dims (Hx, Hy, Hz). */
int pmgpu_poisson_kernel_octant(pmgpu_poisson_t h, void *host, int64_t cap_bytes, int32_t dims[3]);

/* HIP-event timing on the handle's stream (no reference counterpart): after
 * pmgpu_poisson_set_timing(h, 1) every solve records events around the fused
 * z pass (the dominant kernel) and around the whole solve, without host
 * synchronisation; pmgpu_poisson_last_timing returns their averages in ms
 * over the solves recorded since (up to 512), or -1 if none. */
int pmgpu_poisson_last_timing(pmgpu_poisson_t h, double *zpass_ms, double *total_ms);
int pmgpu_poisson_set_timing(pmgpu_poisson_t h, int32_t enable);

#ifdef __cplusplus
}
#endif
#endif
