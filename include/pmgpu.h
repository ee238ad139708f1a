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
  int32_t poisson_kernel;       /* 1 (G) or 2 (grad G)                    */
} pmgpu_params;

enum { PMGPU_FP64 = 0, PMGPU_FP32 = 1 };

typedef struct pmgpu_poisson_s *pmgpu_poisson_t;
typedef struct pmgpu_poisson_batch_s *pmgpu_poisson_batch_t;

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
 * (poisson/pmlib_poisson_setup.f:5-19): builds G-hat on the device (kernel
 * 1: the real octant of G-hat; kernel 2: the complex half-spectrum of grad G,
 * poisson/pmlib_poisson_setup.f:580-635 and the other orders' kernel-2
 * blocks; nothing on an all-periodic patch, where the z pass forms the
 * spectral kernel). */
int pmgpu_poisson_setup(pmgpu_poisson_t h);

/* The refinement level of the patch a handle solves on (class_patch level and
 * parent(1), pmlib_patch.f:15-27), read by the next pmgpu_poisson_setup.
 * Level 1 is a base patch (the default); on any other level the kernel is
 * the refinement-patch kernel of poisson/pmlib_poisson_setup.f:446-528 (order
 * 2; :640-731 order 4, and the other orders' refinement blocks): the kernel
 * regularised at eps minus the one regularised at a eps, a = 2^(level -
 * parent_level), without periodic images (the reference adds image offsets
 * that no loop sets there; they are taken as 0).  Order 0 then fails in
 * setup, as in the reference (:335-339).  An all-periodic patch keeps its
 * spectral kernel, which the reference's solve uses on every level. */
int pmgpu_poisson_set_level(pmgpu_poisson_t h, int32_t level, int32_t parent_level);

/* Replaces pmlib_poisson_solve(patch, topo_all, mesh, ierr, reg_vort, reproj,
 * calc_vecpot) (poisson/pmlib_poisson_solve.f:15-34).  Host FP64 meshes:
 * vort (in), vel (out), vort_out (out when reg_vort or reproj, else may be
 * NULL; it may alias vort, as mesh%vort is overwritten in the reference),
 * vecpot (out when calc_vecpot with kernel 1, else may be NULL; with kernel
 * 2 the reference computes no stream function and vecpot is left as it is).
 * Kernel 2 forms u-hat = K x w-hat and the reprojection K (2 pi i k . w-hat)
 * (poisson/pmlib_poisson_solve.f:873-913). */
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

/* Single-component solve psi = G * w (the vector potential of one
 * vorticity component, the psi-hat = G w-hat product of
 * poisson/pmlib_poisson_solve.f:854-858 for Cin = Cout = 1).  Meshes are
 * (E...) scalar fields; host FP64, or device in the handle's precision.
 * The z-pass tile width is chosen by timing on first use
 * (pmgpu_poisson_tune_scalar; cached per transform shape, and in the file
 * named by PMGPU_TUNE_CACHE when set). */
int pmgpu_poisson_tune_scalar(pmgpu_poisson_t h);
int pmgpu_poisson_solve_scalar(pmgpu_poisson_t h, const double *w, double *psi);
int pmgpu_poisson_solve_scalar_device(pmgpu_poisson_t h, const void *d_w, void *d_psi);

/* Batched multipatch solve (no reference counterpart: the reference solves
 * one patch per call).  A batch holds n set-up handles of any sizes;
 * pmgpu_poisson_batch_solve_scalar runs all n single-component solves as one
 * HIP graph (captured on the first call and whenever a pointer changes,
 * replayed otherwise), each patch on its own branch, on the batch stream. */
int pmgpu_poisson_batch_create(const pmgpu_poisson_t *handles, int32_t n, void *stream,
                               pmgpu_poisson_batch_t *out);
int pmgpu_poisson_batch_solve_scalar(pmgpu_poisson_batch_t b, const void *const *d_in, void *const *d_out);
int pmgpu_poisson_batch_destroy(pmgpu_poisson_batch_t b);

/* ------------------------------------------------------------------------
 * Particles.  Particle arrays are (3, np) FP64 everywhere (positions in FP64
 * keep the particle -> cell index exact).  The host entry points use FP64
 * meshes; the *_device variants take device pointers, a hipStream_t (NULL:
 * the default stream) and the mesh precision: PMGPU_FP64 reproduces the
 * reference's arithmetic, PMGPU_FP32 uses an FP32 mesh with FP32 kernel
 * weights, compensated FP32 sums in deposition (each stencil row's terms of
 * a node fused into one FP32 partial, the row partials Kahan-added) and FP64
 * sums of the weighted FP32 node values in the gather.
 * ------------------------------------------------------------------------ */

/* Replaces pmlib_interp_particle_mesh(topo, xpart, part, mesh, ierr, clear,
 * clear_ghost) (interpolation/pmlib_interp_particle_mesh.f:5-6): deposits
 * part at xpart onto the (3, E...) mesh with the M'4 or M6* kernel of
 * g->interp_order, then adds the periodic ghost layers back onto their
 * images.  clear zeroes the whole mesh first; otherwise clear_ghost zeroes the
 * ghost layers (:81-135).  Deterministic: no atomics touch the mesh. */
int pmgpu_interp_particle_mesh(const pmgpu_geom *g, const double *xpart, const double *part, int64_t np,
                               double *mesh, int32_t clear, int32_t clear_ghost);
int pmgpu_interp_particle_mesh_device(const pmgpu_geom *g, int32_t precision, const void *d_xpart,
                                      const void *d_part, int64_t np, void *d_mesh, int32_t clear,
                                      int32_t clear_ghost, void *stream);

/* Replaces pmlib_topology_parent(topo, topo_parent, ierr)
 * (topology/pmlib_topology_parent.f:5-100) on one rank: the parent cuboid of
 * a patch on `level` whose parent is on `parent_level`: the same box and
 * boundary conditions, cells reldx = 2^(level - parent_level) times larger
 * (ncell must be a multiple of reldx), 2 (M'4) or 3 (M6*) ghost layers. */
int pmgpu_topology_parent(const pmgpu_geom *g, int32_t level, int32_t parent_level, pmgpu_geom *out);

/* Replaces pmlib_interp_mesh_mesh(topo_from, topo_to, mesh_from, mesh_to,
 * ierr, clear) (interpolation/pmlib_interp_mesh_mesh.f:5-2078) for (3, E...)
 * meshes on one rank, with the M'4 or M6* kernel of the geometries'
 * interp_order.  reldx = dx_from(1) / dx_to(1).  reldx <= 0.9999 (patch to
 * parent, :93-1392): every source cell, ghosts included, deposits reldx W
 * per axis onto the target stencil (indices beyond a free side clamped with
 * zero weight; on a periodic side a cell outside the box deposits nothing),
 * then the periodic ghost layers are added onto their images.  Otherwise
 * (parent to patch, :1395-2062): every target cell (interior cells only on
 * periodic axes) adds the interpolated source value, then the periodic ghost
 * layers are copied from their images.  clear zeroes the target first.
 * Both directions are tensor products of per-axis operators and run as three
 * 1D passes (the reference's term-by-term sums regrouped: agreement to
 * rounding, ~1e-15 relative); with PMGPU_M2M_EXACT=1 in the environment they
 * run as gathers in the reference's term order instead, bit-identical to it.
 * Deterministic either way (no atomics).  Cell ratios up to 8 are supported.
 * The passes use a library workspace (kept until pmgpu_finalize): calls that
 * overlap in time race. */
int pmgpu_interp_mesh_mesh(const pmgpu_geom *g_from, const pmgpu_geom *g_to, const double *from, double *to,
                           int32_t clear);
int pmgpu_interp_mesh_mesh_device(const pmgpu_geom *g_from, const pmgpu_geom *g_to, int32_t precision,
                                  const void *d_from, void *d_to, int32_t clear, void *stream);

/* Replaces one field of pmlib_interp_mesh_particle(topo, mesh, part, ierr,
 * vort, dvort, vel, clear) (interpolation/pmlib_interp_mesh_particle.f:5-6):
 * out(:, q) = sum of kernel weights times mesh around pos(:, q), in the
 * reference's term order (bit-identical to it). */
int pmgpu_interp_mesh_particle(const pmgpu_geom *g, const double *mesh, const double *pos, int64_t np,
                               double *out);
int pmgpu_interp_mesh_particle_device(const pmgpu_geom *g, int32_t precision, const void *d_mesh,
                                      const void *d_pos, int64_t np, void *d_out, void *stream);

/* Replaces pmlib_remesh(patch, topo, mesh, part, min_vort, min_dvort, ierr,
 * dvort, vel, mask, incl_ghost) (remesh/pmlib_remesh.f:5-6): one particle at
 * each cell centre whose |vort| > min_vort or |dvort| > min_dvort (and whose
 * mask byte is non-zero when mask is given), created in (k, j, i) order.
 * dvort_m may be NULL (a zero field); vel_m may be NULL unless vel_flag.
 * dvort_flag / vel_flag copy the mesh dvort / vel onto the particles (else
 * zero).  mask is one byte per mesh cell (E...) or NULL.  The count is
 * returned in *np_out; the particle arrays (3, cap) are written only when pos
 * is non-NULL and the count fits in cap (call once with pos = NULL to size
 * them). */
int pmgpu_remesh(const pmgpu_geom *g, const double *vort_m, const double *dvort_m, const double *vel_m,
                 const uint8_t *mask, double min_vort, double min_dvort, int32_t dvort_flag,
                 int32_t vel_flag, int32_t incl_ghost, double *pos, double *vort, double *vel,
                 double *dvort, int64_t cap, int64_t *np_out);
/* Device variant (meshes in the given precision, particles FP64); also
 * zeroes d_vel_rk and d_dvort_rk; d_vel, d_dvort, d_vel_rk and d_dvort_rk
 * may each be NULL (not written). */
int pmgpu_remesh_device(const pmgpu_geom *g, int32_t precision, const void *d_vort_m, const void *d_dvort_m,
                        const void *d_vel_m, const void *d_mask, double min_vort, double min_dvort,
                        int32_t dvort_flag, int32_t vel_flag, int32_t incl_ghost, void *d_pos,
                        void *d_vort, void *d_vel, void *d_dvort, void *d_vel_rk, void *d_dvort_rk,
                        int64_t cap, int64_t *np_out, void *stream);

/* Replaces pmlib_particles_advance(patch, topo, part, itimestage, dtime, ierr,
 * euler_vort) (particles/pmlib_particles_advance.f:5-6) on one rank: the
 * low-storage RK stage update of order rk_order (1-3) followed by the
 * periodic wrap / free-side removal of pmlib_particles_map.  The arrays the
 * map carries (particles/pmlib_particles_advance.f:121-166: pos and vort,
 * and above first order vel_rk and dvort_rk) are reordered and compacted in
 * place; vel and dvort are not carried, as in the reference, and keep the
 * values the update left in them.  The new count is returned in *np_out. */
int pmgpu_particles_advance(const pmgpu_geom *g, int32_t rk_order, int32_t stage, double dt,
                            int32_t euler_vort, double *pos, double *vel, double *vort, double *dvort,
                            double *vel_rk, double *dvort_rk, int64_t np, int64_t *np_out);
/* Device variant; d_arrays = {pos, vel, vort, dvort, vel_rk, dvort_rk}.
 * A NULL dvort means a zero right-hand side (dvort = dvort_rk = +0; no
 * reference counterpart, for callers without one): the update then neither
 * reads dvort nor reads, writes or carries dvort_rk (which may be NULL as
 * well), with the same results as zero arrays. */
int pmgpu_particles_advance_device(const pmgpu_geom *g, int32_t rk_order, int32_t stage, double dt,
                                   int32_t euler_vort, void *const d_arrays[6], int64_t np,
                                   int64_t *np_out, void *stream);
/* The same with a second set of arrays, d_spare, of at least np particles
 * each and distinct from d_arrays (vel and dvort may be NULL, and dvort_rk
 * with a zero right-hand side).  No reference
 * counterpart: it is for time loops that own two particle sets and swap
 * them.  When the map reorders particles, the carried arrays' new contents
 * are written to d_spare instead of back into d_arrays and *in_spare = 1; the
 * caller then continues with d_spare (its vel and dvort are not written).
 * Otherwise the update is in place in d_arrays and *in_spare = 0. */
int pmgpu_particles_advance_device_spare(const pmgpu_geom *g, int32_t rk_order, int32_t stage, double dt,
                                         int32_t euler_vort, void *const d_arrays[6],
                                         void *const d_spare[6], int64_t np, int64_t *np_out,
                                         int32_t *in_spare, void *stream);

/* The particle -> cell map of every interpolation and mapping routine:
 * idx(d, q) = NINT((pos(d, q) - xmin(d)) * (1 / dx(d)))
 * (interpolation/pmlib_interp_particle_mesh.f:153-155). */
int pmgpu_cell_index(const pmgpu_geom *g, const double *pos, int64_t np, int32_t *idx);

/* ------------------------------------------------------------------------
 * Brinkman penalisation
 * ------------------------------------------------------------------------ */

/* One iteration of penalisation_solve(patch, topo_all, mesh, ierr)
 * (vortex_codes/vortex_penalisation/source/penalisation/penalisation_solve.f:6,
 * the loop body :96-111 and :146-273): f = chi (us - vel) on every cell, then
 * vort += relax curl(f) on the interior cells with central differences of
 * order fd_order (2 or 4; pmlib_fd_order) over 2 dx.  chi is one value per
 * mesh cell (the mask, or lambda dt chi_s / (1 + lambda dt chi_s) for a
 * finite penalisation factor); us is solid_vel (3, E...) if non-NULL, else
 * the constant us_const (NULL: at rest).  diag (may be NULL) receives
 * sum_dforce(1:3), res_enst and sum_enst_pen of the iteration (:177-190). */
int pmgpu_penalise(const pmgpu_geom *g, const double *vel, const double *chi, const double *solid_vel,
                   const double us_const[3], double relax, int32_t fd_order, double *vort,
                   double diag[5]);
/* Device variant in FP64 or FP32 (precision); the sums are FP64 either way
 * and, when diag is non-NULL, the call synchronises the stream to return
 * them. */
int pmgpu_penalise_device(const pmgpu_geom *g, int32_t precision, const void *d_vel, const void *d_chi,
                          const void *d_solid_vel, const double us_const[3], double relax,
                          int32_t fd_order, void *d_vort, double *diag, void *stream);
/* The penalisation iteration's module globals (vortex_codes/
 * vortex_penalisation/source/mod_parameters.f:43-74, pmlib_fd_order). */
typedef struct pmgpu_penal_params {
  double relaxation;       /* penal_relaxation                                */
  double tolerance;        /* penal_tolerance (relative force increment)      */
  int32_t max_iter;        /* penal_max                                       */
  int32_t fd_order;        /* pmlib_fd_order: 2 or 4                          */
  double dtime;            /* dtime                                           */
  double vel_inf[3];       /* vel_inf, added to the velocity at the end       */
  int32_t reproject;       /* reproject: the final solve reprojects vort      */
  int32_t fix_free_stream; /* 0: as the reference (vel(2) += vel_inf(3), and
                              vel(3) gets nothing); 1: vel(3) += vel_inf(3)   */
} pmgpu_penal_params;

/* penalisation_solve(patch, topo_all, mesh, ierr)
 * (vortex_codes/vortex_penalisation/source/penalisation/penalisation_solve.f:6,
 * set-up :61-111, iteration :133-343, map back :348-365, final solve
 * :372-398, free stream :405-414) with the penalisation mesh on the solver's
 * own patch: vel_pen = us - vel; then, until the relative force increment
 * falls to the tolerance (at least once, at most max_iter times),
 * vort_pen += relax curl(chi (vel_pen - u_pen)) and u_pen = the velocity of
 * vort_pen (pmgpu_poisson_solve_device on h); then vort += vort_pen (periodic
 * ghosts mapped), vel = the velocity of vort (reprojected vorticity into vort
 * when reproject), and vel += vel_inf.  chi, solid_vel (NULL: us_const, NULL
 * for rest) and the meshes are device arrays in h's precision; h must be set
 * up.  force, npen and residual (each may be NULL) receive the accumulated
 * force, the iterations and the last relative force increment.  The solve
 * of the penalisation vorticity uses the library's free-space solver where
 * the reference calls an external one (poisson_solver_solve3d, not
 * vendored; the reference's commented-out alternative is
 * pmlib_poisson_solve, :295-296).  Synchronises the stream once per
 * iteration (the stopping test). */
int pmgpu_penalisation_solve_device(pmgpu_poisson_t h, const pmgpu_penal_params *pp, const void *d_chi,
                                    const void *d_solid_vel, const double us_const[3], void *d_vel,
                                    void *d_vort, double force[3], int32_t *npen, double *residual);
/* Same on host FP64 meshes (an FP64 handle); copies over PCIe. */
int pmgpu_penalisation_solve(pmgpu_poisson_t h, const pmgpu_penal_params *pp, const double *chi,
                             const double *solid_vel, const double us_const[3], double *vel, double *vort,
                             double force[3], int32_t *npen, double *residual);

/* Allocation-free variant of pmgpu_penalise_device.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It calls no hipMalloc/hipFree and never synchronises: the
 * sums go to d_diag, a DEVICE array of 5 doubles written on the stream (NULL:
 * no sums), and the per-block partial sums live in a fixed scratch array in
 * the library, one per GPU (10 MiB, held until the process ends), enough for
 * ceil(N1/64) * ceil(N2/4) * N3 <= 262144 blocks (e.g. 384^3); larger meshes
 * are refused with an error.  Calls that overlap on one GPU share that
 * scratch and race: order them on one stream. */
int pmgpu_penalise_device_noalloc(const pmgpu_geom *g, int32_t precision, const void *d_vel,
                                  const void *d_chi, const void *d_solid_vel,
                                  const double us_const[3], double relax, int32_t fd_order,
                                  void *d_vort, double *d_diag, void *stream);

/* ------------------------------------------------------------------------
 * Vorticity right-hand side (stretching and diffusion)
 * ------------------------------------------------------------------------ */

/* part_deriv_vorticity(topo, mesh, ierr)
 * (vortex_codes/vortex_penalisation/source/particle_derivatives/
 * part_deriv_vorticity.f:6, body :57-208): on the interior cells
 * dvort_l = visc (d2 w_l/dx2 + d2 w_l/dy2 + d2 w_l/dz2) + w1 du_l/dx +
 * w2 du_l/dy + w3 du_l/dz with central differences of order fd_order
 * (2 or 4; pmlib_fd_order), visc being the client's viscosity; then the
 * periodic ghost map of dvort.  vort, vel and dvort are (3, E...) meshes;
 * dvort's ghost cells on free sides keep their values.  Needs fd_order / 2
 * ghost layers on every side. */
int pmgpu_vorticity_rhs(const pmgpu_geom *g, const double *vort, const double *vel, double visc,
                        int32_t fd_order, double *dvort);
/* Device variant in FP64 or FP32 (precision), stream-ordered; dvort must not
 * alias vort or vel. */
int pmgpu_vorticity_rhs_device(const pmgpu_geom *g, int32_t precision, const void *d_vort, const void *d_vel,
                               double visc, int32_t fd_order, void *d_dvort, void *stream);

/* ------------------------------------------------------------------------
 * Global diagnostics (FP64)
 * ------------------------------------------------------------------------ */

/* diagnostics_ring(topo, part, ierr)
 * (vortex_codes/vortex_structures/source/diagnostics/diagnostics_ring.f:5,
 * sums :72-158, normalisations :166-226) on one rank, ring axis x: out =
 * impulse(3), angular impulse(3), kinetic energy, enstrophy, helicity,
 * circulation, vortex centroid(3), ring radius, core radius, max |vort|
 * (16 doubles), each particle weighted by dx(1) dx(2) dx(3).  d_pos, d_vort,
 * d_vel are (3, np) FP64 device arrays.  The totals go to d_out (device, 16
 * doubles, stream-ordered) and/or h_out (host; the call then synchronises);
 * either may be NULL.  Sums are reduced in a fixed order (run-to-run
 * identical), not in the reference's serial order. */
int pmgpu_diagnostics_particles(const double dx[3], const double *d_pos, const double *d_vort,
                                const double *d_vel, int64_t np, double *d_out, double *h_out,
                                void *stream);
/* diagnostics_energy_mesh(topo, mesh, ierr)
 * (vortex_codes/vortex_structures/source/diagnostics/diagnostics_energy_mesh.f:5,
 * :72-134) on one rank: kinetic energy, enstrophy, helicity energy, Lamb
 * energy and max |vort| over the interior cells (5 doubles), accumulated in
 * FP64 from FP64 or FP32 (precision) meshes (3, E...).  As in the reference,
 * the helicity and Lamb terms divide by the cell's enstrophy. */
int pmgpu_diagnostics_mesh(const pmgpu_geom *g, int32_t precision, const void *d_vel, const void *d_vort,
                           double *d_out, double *h_out, void *stream);

/* ------------------------------------------------------------------------
 * Device-resident vortex time step, recorded once as a HIP graph
 * ------------------------------------------------------------------------ */

/* One time step of the vortex client's loop
 * (vortex_codes/vortex_structures/source/main.f:453-505: per RK stage
 * pmlib_interp_particle_mesh, pmlib_poisson_solve, pmlib_interp_mesh_particle,
 * pmlib_particles_advance; then pmlib_interp_particle_mesh and pmlib_remesh
 * with the stage's velocity mesh, no vorticity right-hand side) on a
 * particle set of fixed capacity held by the handle.  The particle count
 * lives in device memory: every launch covers the capacity and reads the
 * live count, so a step needs no host synchronisation; with use_graph the
 * step is recorded once as a HIP graph (again when dt or min_vort change)
 * and replayed.  The arithmetic and the particle order are those of the
 * eager entry points (pmgpu_interp_*_device, pmgpu_particles_advance_device,
 * pmgpu_remesh_device).  Migration makes one pass per axis: a particle still
 * outside a periodic axis afterwards (it moved more than a period in one
 * stage) sets status bit 2; a remesh above the capacity keeps the first cap
 * particles and sets status bit 1 (read with pmgpu_vstep_sync).  No
 * reference counterpart: it is the graph-replayed time step of the build. */
typedef struct pmgpu_vstep_s *pmgpu_vstep_t;
int pmgpu_vstep_create(const pmgpu_geom *g, const pmgpu_params *p, int32_t precision, int32_t rk_order,
                       int64_t capacity, pmgpu_vstep_t *out);
int pmgpu_vstep_destroy(pmgpu_vstep_t h);
/* (re)starts from device (3, np) FP64 positions and vorticities; the other
 * particle fields start at zero (synchronises). */
int pmgpu_vstep_load(pmgpu_vstep_t h, const void *d_pos, const void *d_vort, int64_t np);
/* enqueues one step on the handle's stream (no synchronisation) */
int pmgpu_vstep_step(pmgpu_vstep_t h, double dt, double min_vort, int32_t use_graph);
/* waits for the stream; the particle count and the status bits */
int pmgpu_vstep_sync(pmgpu_vstep_t h, int64_t *np, int32_t *status);
/* device pointers of pos, vel, vort, dvort, vel_rk, dvort_rk ((3, cap)
 * FP64 each) and the handle's stream */
int pmgpu_vstep_arrays(pmgpu_vstep_t h, void *d_arrays[6], void **stream);
/* copies the live particles' pos, vort, vel (each may be NULL) to host
 * (3, np) arrays of capacity cap (synchronises) */
int pmgpu_vstep_fetch(pmgpu_vstep_t h, double *pos, double *vort, double *vel, int64_t cap, int64_t *np);

/* ------------------------------------------------------------------------
 * On-disk formats (host I/O from device-resident or host arrays)
 * ------------------------------------------------------------------------ */

/* pmlib_output_mesh(outdir, tag, patch, topo, mesh, ierr)
 * (source/output/pmlib_output_mesh.f:5-188) on one rank: writes
 * outdir//tag//'.pvti' and outdir//tag//'_P00000.vti' (VTK image data, cell
 * data vort and vel of the interior cells as Float32, appended raw),
 * byte for byte as the reference writes them.  vort, vel: (3, E...) meshes
 * in precision, device pointers when on_device, else host. */
int pmgpu_output_mesh(const char *outdir, const char *tag, const pmgpu_geom *g, int32_t precision,
                      const void *vort, const void *vel, int32_t on_device);
/* output_particles(tag, topo, part, ierr)
 * (vortex_codes/vortex_structures/source/output/output_particles.f:50-69) on
 * one rank: the unformatted sequential file of records itime + 1,
 * time + dtime, npart, then one record pos(1:3), vort(1:3) per particle
 * (FP64, 4-byte record markers).  pos, vort: (3, np) FP64, device pointers
 * when on_device. */
int pmgpu_output_particles(const char *path, int32_t itime, double time, double dtime, const double *pos,
                           const double *vort, int64_t np, int32_t on_device);
/* The particle reading of input_particles(infile, patch, topo_all, mesh,
 * part, trunc, ierr) (vortex_codes/vortex_structures/source/input/
 * input_particles.f:81-135, 200-230) on one rank: reads the header (itime,
 * time) and keeps, in file order, the particles with |vort| >= trunc; np
 * receives their count and vort_max the largest kept |vort|.  With pos and
 * vort NULL only the header and count are returned; otherwise they receive
 * (3, np) FP64 arrays (device pointers when on_device, capacity cap
 * particles).  No particle above trunc is an error, as in the reference.
 * The reference then re-patches and distributes (out of this path). */
int pmgpu_input_particles(const char *path, double trunc, int32_t *itime, double *time, int64_t *np,
                          double *vort_max, double *pos, double *vort, int64_t cap, int32_t on_device);

#ifdef __cplusplus
}
#endif
#endif
