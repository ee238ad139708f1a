/*
 * pmoracle.c -- CPU restatement of the pmlib3d particle-mesh hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg call it to check or
 * time the HIP path.  Nothing in the product (pmlib3d_amd/) links or calls it.
 *
 * It restates, in plain FP64 C with floating-point contraction disabled, the
 * algorithms of the reference Fortran library (paths relative to the
 * reference tree):
 *   or_glassman        fourier/pmlib_fourier_fft.f:296-431 (and ifft :315-450)
 *   or_green_real      poisson/pmlib_poisson_setup.f:185-1519 (kernel 1 and 2,
 *                      orders 0,2,4,6,8,10 on the base level)
 *   or_poisson_setup   poisson/pmlib_poisson_setup.f:1533 + pmlib_fft with
 *                      fft_shift (fourier/pmlib_fourier_fft.f:139-150)
 *   or_poisson_solve   poisson/pmlib_poisson_solve.f:119-233, :697-954,
 *                      fourier/pmlib_fourier_fft.f, fourier/pmlib_fourier_ifft.f,
 *                      mesh/pmlib_mesh_map_ghost.f (forward ghost fill)
 *   or_p2m             interpolation/pmlib_interp_particle_mesh.f:137-1134
 *   or_m2p             interpolation/pmlib_interp_mesh_particle.f:76-1266
 *   or_remesh          remesh/pmlib_remesh.f:84-226 (incl. mask, dvort, vel)
 *   or_advance         particles/pmlib_particles_advance.f:49-113 +
 *                      particles/pmlib_particles_map.f:57-133 (one rank)
 *   or_penalise        vortex_codes/vortex_penalisation/source/penalisation/
 *                      penalisation_solve.f:96-111,146-273 (one iteration,
 *                      2nd/4th-order curl, Brinkman factor per cell)
 *   or_vort_rhs        vortex_codes/vortex_penalisation/source/
 *                      particle_derivatives/part_deriv_vorticity.f:57-208
 *                      (stretching + diffusion, 2nd/4th-order FD, ghost map)
 *   or_diag_particles  vortex_codes/vortex_structures/source/diagnostics/
 *                      diagnostics_ring.f:72-226 (impulse, angular impulse,
 *                      energy, enstrophy, helicity, circulation, centroid,
 *                      ring and core radius)
 *   or_diag_mesh       .../diagnostics/diagnostics_energy_mesh.f:72-134
 *
 * Array layout is the reference's: meshes are (3, i, j, k) with the component
 * fastest and ghost layers included, particles are (3, np).  Fixed-size fp64
 * complex is C99 double complex.
 */
#include <complex.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#pragma STDC FP_CONTRACT OFF

typedef double complex cplx;

/* Geometry of one patch on one rank (class_patch + class_topology fields). */
typedef struct {
  int32_t ncell[3];
  int32_t nghost;       /* patch%nghost (exterior free-space ghosts)        */
  int32_t bound_cond[3];/* 0 free, 1 periodic                               */
  int32_t interp_order; /* 3 = M'4, 4 = M6*; sets periodic-side ghost count  */
  double xmin[3], xmax[3], dx[3];
} or_geom;

typedef struct {
  double regularisation_radius;
  int32_t poisson_order, poisson_kernel;
  double refine;        /* a = 2^(level - parent level) of a refinement patch; 0: base level */
} or_params;

/* ghost layers on the low/high side of axis d on a single rank:
 * exterior free sides get patch%nghost, periodic sides get the interior
 * count 2 (M'4) or 3 (M6*) (topology/pmlib_topology_cuboid.f:71-80,192-270) */
static int ng_side(const or_geom *g, int d) {
  if (g->bound_cond[d] == 0) return g->nghost;
  return g->interp_order == 4 ? 3 : 2;
}
static int mesh_ext(const or_geom *g, int d) { return g->ncell[d] + 2 * ng_side(g, d); }

int or_mesh_extent(const or_geom *g, int32_t *ext) {
  for (int d = 0; d < 3; ++d) ext[d] = mesh_ext(g, d);
  return 0;
}

/* solve-domain size M and FFT length L per axis
 * (topology/pmlib_topology_pencil.f:77-103, poisson/pmlib_poisson_solve.f:176-182) */
static void solve_dims(const or_geom *g, int *M, int *L) {
  for (int d = 0; d < 3; ++d) {
    if (g->bound_cond[d] == 0) { M[d] = g->ncell[d] + 2 * g->nghost; L[d] = 2 * M[d]; }
    else { M[d] = g->ncell[d]; L[d] = M[d]; }
  }
}

/* ------------------------------------------------------------------------ */
/* Glassman general-N FFT, unnormalised; sign = -1 forward, +1 inverse.      */
/* u holds the input and receives the output, w is scratch of length n.      */
/* Twiddles advance by repeated multiplication as in the reference.          */
void or_glassman(cplx *u, cplx *w, int n, int sign) {
  const double pi = acos(-1.0);
  int a = 1, b = n, c = 1;
  int in_u = 1;
  while (b > 1) {
    a = c * a;
    for (c = 2; c <= b; ++c)
      if (b % c == 0) break;
    b = b / c;
    double angle = 2.0 * pi / (double)(a * c);
    cplx delta = CMPLX(cos(angle), sign * sin(angle));
    cplx omega = CMPLX(1.0, 0.0);
    cplx *src = in_u ? u : w, *dst = in_u ? w : u;
    for (int ic = 1; ic <= c; ++ic) {
      for (int ia = 1; ia <= a; ++ia) {
        for (int ib = 1; ib <= b; ++ib) {
          cplx s = src[(ib + b * (c - 1) + b * c * (ia - 1)) - 1];
          for (int jrc = 2; jrc <= c; ++jrc) {
            int jc = c + 1 - jrc;
            s = src[(ib + b * (jc - 1) + b * c * (ia - 1)) - 1] + omega * s;
          }
          dst[(ib + b * (ia - 1) + b * a * (ic - 1)) - 1] = s;
        }
        omega = delta * omega;
      }
    }
    in_u = !in_u;
  }
  if (!in_u) memcpy(u, w, sizeof(cplx) * (size_t)n);
}

/* ------------------------------------------------------------------------ */
/* Regularised Green's function on the kernel grid                          */

/* kernel-grid size per axis (topology/pmlib_topology_kernel.f:82-111):
 * free axes 2(N+2ng), periodic axes N */
static void kernel_dims(const or_geom *g, int *L) {
  for (int d = 0; d < 3; ++d)
    L[d] = g->bound_cond[d] == 0 ? 2 * (g->ncell[d] + 2 * g->nghost) : g->ncell[d];
}

/* Fills G (nvar x Lx x Ly x Lz, Fortran order, nvar = 1 for kernel 1 and 3
 * for kernel 2) with the real-space kernel on the kernel grid, images summed
 * over +-1 patch length on periodic axes.  Returns nonzero on a bad order. */
/* The sine integral Si(x) as poisson/pmlib_poisson_setup.f:1739-1841
 * evaluates it: the published cisi method (Numerical Recipes, sec. 6.9):
 * for |x| > 2 the continued fraction of E1(ix) by the modified Lentz method,
 * Si = pi/2 + Im(e^{-it} h); otherwise the power series, odd and even terms
 * accumulated alternately; both to a relative 1e-10, with the reference's
 * constants, so that the truncation matches. */
double or_si(double x) {
  const double pi = acos(-1.0), eps = 1.0e-10, fpmin = 1.0e-30, fpmax = 1.0e30, tmin = 2.0;
  const int maxit = 1000;
  const double t = fabs(x);
  if (t == 0.0) return 0.0;
  double si;
  if (t > tmin) {
    cplx b = 1.0 + t * I, c = fpmax, d = 1.0 / b, h = d;
    for (int i = 2; i <= maxit; ++i) {
      const double a = -(double)((i - 1) * (i - 1));
      b = b + 2.0;
      d = 1.0 / (a * d + b);
      c = b + a / c;
      const cplx del = c * d;
      h = h * del;
      if (fabs(creal(del - 1.0)) + fabs(cimag(del - 1.0)) <= eps) break;
    }
    h = (cos(t) - sin(t) * I) * h;
    si = 0.5 * pi + cimag(h);
  } else if (t < sqrt(fpmin)) {
    si = t;
  } else {
    double sum = 0.0, sums = 0.0, sumc = 0.0, sign = 1.0, fact = 1.0;
    int odd = 1;
    for (int k = 1; k <= maxit; ++k) {
      fact = fact * t / (double)k;
      const double term = fact / (double)k;
      sum = sum + sign * term;
      const double err = term / fabs(sum);
      if (odd) {
        sign = -sign;
        sums = sum;
        sum = sumc;
      } else {
        sumc = sum;
        sum = sums;
      }
      if (err < eps) break;
      odd = !odd;
    }
    si = sums;
  }
  return x < 0.0 ? -si : si;
}

int or_green_real(const or_geom *g, const or_params *p, double *G) {
  int L[3];
  kernel_dims(g, L);
  const double pi = acos(-1.0);
  double xk[3], len[3];
  for (int d = 0; d < 3; ++d) {
    xk[d] = -(0.5 * (double)L[d] + 0.5) * g->dx[d];
    len[d] = g->bound_cond[d] == 0 ? g->dx[d] * (double)(2 * (g->ncell[d] + 2 * g->nghost))
                                   : g->dx[d] * (double)g->ncell[d];
  }
  double dmax = g->dx[0];
  if (g->dx[1] > dmax) dmax = g->dx[1];
  if (g->dx[2] > dmax) dmax = g->dx[2];
  const double eps = p->regularisation_radius * dmax;
  const double c_1_4pi = 0.25 / pi;
  const double c_1_sqrt2 = 1.0 / sqrt(2.0);
  const double c_1_sqrt2pi = 1.0 / sqrt(2.0 * pi);
  const double c_sq2_4pieps = sqrt(2.0) / (4.0 * pow(pi, 1.5) * eps);
  const int ord = p->poisson_order, ker = p->poisson_kernel;
  const int nvar = ker == 1 ? 1 : 3;
  if (ker != 1 && ker != 2) return -1;
  if (ord != 0 && ord != 2 && ord != 4 && ord != 6 && ord != 8 && ord != 10 && ord != 100) return -1;
  const size_t n = (size_t)L[0] * L[1] * L[2];
  memset(G, 0, sizeof(double) * n * nvar);
  if (p->refine > 0.0) {
    /* refinement patch (level > 1, poisson/pmlib_poisson_setup.f:446-528,
     * :640-731 and the other orders' refinement blocks): the difference of
     * the kernels regularised at eps and at a eps, G(rho) - G(rho / a),
     * without periodic images (the reference's image indices are unset
     * there; they are taken as 0).  Order 0 has no refinement kernel. */
    if (ord == 0 || ord == 100) return -1;
    const double ca = 1.0 / p->refine;
    for (int k = 1; k <= L[2]; ++k) {
      double pz = xk[2] + ((double)(k - 1) + 0.5) * g->dx[2];
      for (int j = 1; j <= L[1]; ++j) {
        double py = xk[1] + ((double)(j - 1) + 0.5) * g->dx[1];
        for (int i = 1; i <= L[0]; ++i) {
          double px = xk[0] + ((double)(i - 1) + 0.5) * g->dx[0];
          double r = sqrt(px * px + py * py + pz * pz);
          size_t o = ((size_t)(k - 1) * L[1] + (j - 1)) * L[0] + (i - 1);
          if (r == 0.0) {
            if (ker == 1) {
              double c = ord == 2 ? 1.0 : ord == 4 ? 1.5 : ord == 6 ? 1.875 : ord == 8 ? 2.1875 : 2.4609375;
              G[o] = c * c_sq2_4pieps * (1.0 - ca);
            }
            continue;
          }
          double part[2];
          for (int q = 0; q < 2; ++q) {
            /* q = 0: regularised at eps; q = 1: at a eps (rho / a) */
            double rho = r / eps * (q ? ca : 1.0);
            double E = exp(-0.5 * (rho * rho)) * c_1_sqrt2pi, F = erf(c_1_sqrt2 * rho);
            double r2 = rho * rho, r3 = r2 * rho, r5 = r3 * r2, r7 = r5 * r2, r9 = r7 * r2, poly;
            if (ker == 1) {
              switch (ord) {
                case 2: poly = 0.0; break;
                case 4: poly = rho; break;
                case 6: poly = 1.75 * rho - 0.25 * r3; break;
                case 8: poly = 2.375 * rho + (-2.0 / 3.0) * r3 + (1.0 / 24.0) * r5; break;
                default:
                  poly = 2.921875 * rho + (-233.0 / 192.0) * r3 + (29.0 / 192.0) * r5 + (-1.0 / 192.0) * r7;
                  break;
              }
            } else {
              switch (ord) {
                case 2: poly = -2.0 * rho; break;
                case 4: poly = -2.0 * rho + r3; break;
                case 6: poly = -2.0 * rho + 2.25 * r3 - 0.25 * r5; break;
                case 8: poly = -2.0 * rho + (89.0 / 24.0) * r3 + (-20.0 / 24.0) * r5 + (1.0 / 24.0) * r7; break;
                default:
                  poly = -2.0 * rho + (1027.0 / 192.0) * r3 + (-349.0 / 192.0) * r5 + (35.0 / 192.0) * r7 +
                         (-1.0 / 192.0) * r9;
                  break;
              }
            }
            part[q] = poly * E + F;
          }
          if (ker == 1) {
            G[o] = c_1_4pi / r * (part[0] - part[1]);
          } else {
            double s = -c_1_4pi / (r * r * r) * (part[0] - part[1]);
            G[o * 3 + 0] = px * s;
            G[o * 3 + 1] = py * s;
            G[o * 3 + 2] = pz * s;
          }
        }
      }
    }
    return 0;
  }
  const int b0 = g->bound_cond[0], b1 = g->bound_cond[1], b2 = g->bound_cond[2];
  for (int kk = -b2; kk <= b2; ++kk)
    for (int jj = -b1; jj <= b1; ++jj)
      for (int ii = -b0; ii <= b0; ++ii)
        for (int k = 1; k <= L[2]; ++k) {
          double pz = xk[2] + ((double)(k - 1) + 0.5) * g->dx[2] + len[2] * (double)kk;
          for (int j = 1; j <= L[1]; ++j) {
            double py = xk[1] + ((double)(j - 1) + 0.5) * g->dx[1] + len[1] * (double)jj;
            for (int i = 1; i <= L[0]; ++i) {
              double px = xk[0] + ((double)(i - 1) + 0.5) * g->dx[0] + len[0] * (double)ii;
              double r = sqrt(px * px + py * py + pz * pz);
              double rho = r / eps;
              size_t o = ((size_t)(k - 1) * L[1] + (j - 1)) * L[0] + (i - 1);
              if (ord == 100) {
                /* the band-limited kernel (:1419-1505): rho = pi r / h,
                 * G = Si(rho) / (2 pi^2 r), G(0) = 1 / (2 pi h); grad G =
                 * x (sin rho - Si(rho)) / (2 pi^2 r^3), 0 at r = 0 */
                const double rb = pi * r / dmax;
                if (ker == 1) {
                  G[o] += r == 0.0 ? 1.0 / (2.0 * pi * dmax) : or_si(rb) / (2.0 * (pi * pi) * r);
                } else if (r != 0.0) {
                  const double f = (sin(rb) - or_si(rb)) / (2.0 * (pi * pi) * (r * r * r));
                  G[o * 3 + 0] += px * f;
                  G[o * 3 + 1] += py * f;
                  G[o * 3 + 2] += pz * f;
                }
                continue;
              }
              double E = exp(-0.5 * (rho * rho)) * c_1_sqrt2pi;
              double F = erf(c_1_sqrt2 * rho);
              double r2 = rho * rho, r3 = r2 * rho, r5 = r3 * r2, r7 = r5 * r2, r9 = r7 * r2;
              if (ker == 1) {
                double v;
                if (r == 0.0) {
                  switch (ord) {
                    case 0: v = 1.0; break;
                    case 2: v = c_sq2_4pieps; break;
                    case 4: v = 1.5 * c_sq2_4pieps; break;
                    case 6: v = 1.875 * c_sq2_4pieps; break;
                    case 8: v = 2.1875 * c_sq2_4pieps; break;
                    default: v = 2.4609375 * c_sq2_4pieps; break;
                  }
                } else {
                  double poly;
                  switch (ord) {
                    case 0: v = c_1_4pi / r; G[o] += v; continue;
                    case 2: poly = 0.0; break;
                    case 4: poly = rho; break;
                    case 6: poly = 1.75 * rho - 0.25 * r3; break;
                    case 8: poly = 2.375 * rho + (-2.0 / 3.0) * r3 + (1.0 / 24.0) * r5; break;
                    default:
                      poly = 2.921875 * rho + (-233.0 / 192.0) * r3 + (29.0 / 192.0) * r5 +
                             (-1.0 / 192.0) * r7;
                      break;
                  }
                  v = c_1_4pi / r * (poly * E + F);
                }
                G[o] += v;
              } else {
                if (r == 0.0) continue;
                double poly;
                switch (ord) {
                  case 0: poly = 0.0; F = 1.0; break;
                  case 2: poly = -2.0 * rho; break;
                  case 4: poly = -2.0 * rho + r3; break;
                  case 6: poly = -2.0 * rho + 2.25 * r3 - 0.25 * r5; break;
                  case 8:
                    poly = -2.0 * rho + (89.0 / 24.0) * r3 + (-20.0 / 24.0) * r5 + (1.0 / 24.0) * r7;
                    break;
                  default:
                    poly = -2.0 * rho + (1027.0 / 192.0) * r3 + (-349.0 / 192.0) * r5 +
                           (35.0 / 192.0) * r7 + (-1.0 / 192.0) * r9;
                    break;
                }
                double s = -c_1_4pi / (r * r * r) * (poly * E + F);
                G[o * 3 + 0] += px * s;
                G[o * 3 + 1] += py * s;
                G[o * 3 + 2] += pz * s;
              }
            }
          }
        }
  return 0;
}

/* One 1D pass along axis `ax` of a complex (nvar, n0, n1, n2) array, in
 * place; `shift` applies the fft_shift half rotation before the transform. */
static void pass_1d(cplx *a, int nvar, const int *n, int ax, int sign, int shift, double scale,
                    int len_used, int len_keep) {
  int len = n[ax];
  cplx *u = (cplx *)malloc(sizeof(cplx) * (size_t)len);
  cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)len);
  size_t st[3] = {(size_t)nvar, (size_t)nvar * n[0], (size_t)nvar * n[0] * n[1]};
  int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
  for (int b = 0; b < n[o2]; ++b)
    for (int c = 0; c < n[o1]; ++c)
      for (int v = 0; v < nvar; ++v) {
        size_t base = (size_t)v + st[o1] * c + st[o2] * b;
        for (int i = 0; i < len; ++i) {
          int src = shift ? (i < len / 2 ? i + len / 2 : i - len / 2) : i;
          u[i] = src < len_used ? a[base + st[ax] * src] : 0.0;
        }
        or_glassman(u, w, len, sign);
        for (int i = 0; i < len_keep; ++i) a[base + st[ax] * i] = scale == 1.0 ? u[i] : u[i] / scale;
      }
  free(u);
  free(w);
}

/* Size (complex elements, nvar included) of the kernel spectrum. */
int64_t or_kernel_size(const or_geom *g, const or_params *p) {
  int L[3];
  kernel_dims(g, L);
  return (int64_t)L[0] * L[1] * L[2] * (p->poisson_kernel == 1 ? 1 : 3);
}

/* pmlib_poisson_setup: G on the kernel grid, then the shifted forward 3D FFT
 * (x, y, z passes) into Ghat (nvar, Lx, Ly, Lz) complex. */
int or_poisson_setup(const or_geom *g, const or_params *p, cplx *Ghat) {
  int L[3];
  kernel_dims(g, L);
  int nvar = p->poisson_kernel == 1 ? 1 : 3;
  size_t n = (size_t)L[0] * L[1] * L[2] * nvar;
  double *G = (double *)malloc(sizeof(double) * n);
  int e = or_green_real(g, p, G);
  if (e) { free(G); return e; }
  for (size_t i = 0; i < n; ++i) Ghat[i] = G[i];
  free(G);
  for (int ax = 0; ax < 3; ++ax) pass_1d(Ghat, nvar, L, ax, -1, 1, 1.0, L[ax], L[ax]);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* pmlib_poisson_solve with fft_seq = (1,2,3).  Input/outputs are meshes in  */
/* the reference layout including ghosts.  vort_out receives the filtered    */
/* (and reprojected) vorticity when reg_vort or reproj, vecpot the stream    */
/* function when calc_vecpot and kernel 1.  Returns nonzero on bad input.    */
int or_poisson_solve(const or_geom *g, const or_params *p, const cplx *Ghat, const double *vort,
                     double *vel, double *vort_out, double *vecpot, int reg_vort, int reproj,
                     int calc_vecpot) {
  int M[3], L[3], E[3], lo[3];
  solve_dims(g, M, L);
  for (int d = 0; d < 3; ++d) {
    E[d] = mesh_ext(g, d);
    lo[d] = g->bound_cond[d] == 0 ? 0 : ng_side(g, d); /* first pencil cell in mesh */
    if (L[d] % 2) return -2;                           /* fourier/pmlib_fourier_fft.f:126 */
  }
  const int ifft_vort = reg_vort || reproj;
  /* order 100 has no vorticity filter (poisson/pmlib_poisson_solve.f:311-329
   * sets none for it) */
  if (ifft_vort && p->poisson_order == 100) return -3;
  const int ifft_psi = calc_vecpot && p->poisson_kernel == 1;
  const double pi = acos(-1.0);
  /* z-pencil array (3, Lx, Ly, Mz); the x and y passes fill it progressively */
  int nz[3] = {L[0], L[1], M[2]};
  size_t nzp = (size_t)L[0] * L[1] * M[2];
  cplx *W = (cplx *)calloc(3 * nzp, sizeof(cplx));
  cplx *V = (cplx *)calloc(3 * nzp, sizeof(cplx));
  cplx *Wf = ifft_vort ? (cplx *)calloc(3 * nzp, sizeof(cplx)) : NULL;
  cplx *P = ifft_psi ? (cplx *)calloc(3 * nzp, sizeof(cplx)) : NULL;
  for (int k = 0; k < M[2]; ++k)
    for (int j = 0; j < M[1]; ++j)
      for (int i = 0; i < M[0]; ++i)
        for (int c = 0; c < 3; ++c) {
          size_t s = (size_t)c + 3 * ((size_t)(i + lo[0]) + (size_t)E[0] * ((j + lo[1]) + (size_t)E[1] * (k + lo[2])));
          W[c + 3 * (i + (size_t)L[0] * (j + (size_t)L[1] * k))] = vort[s];
        }
  pass_1d(W, 3, nz, 0, -1, 0, 1.0, M[0], L[0]);
  pass_1d(W, 3, nz, 1, -1, 0, 1.0, M[1], L[1]);
  /* wavenumbers (poisson/pmlib_poisson_solve.f:184-223, single rank) */
  double dk[3], ksmp[3], kmin[3];
  for (int d = 0; d < 3; ++d) {
    double lp = (double)L[d] * g->dx[d];
    dk[d] = 1.0 / lp;
    ksmp[d] = 1.0 / (2.0 * g->dx[d]);
    kmin[d] = -ksmp[d] + 0.0 * dk[d];
  }
  double dmax = g->dx[0];
  if (g->dx[1] > dmax) dmax = g->dx[1];
  if (g->dx[2] > dmax) dmax = g->dx[2];
  const double eps = p->regularisation_radius * dmax;
  const int allper = g->bound_cond[0] == 1 && g->bound_cond[1] == 1 && g->bound_cond[2] == 1;
  const int nfft = L[2];
  cplx *f[3], *o[9], *tmp;
  for (int q = 0; q < 3; ++q) f[q] = (cplx *)malloc(sizeof(cplx) * nfft);
  for (int q = 0; q < 9; ++q) o[q] = (cplx *)malloc(sizeof(cplx) * nfft);
  tmp = (cplx *)malloc(sizeof(cplx) * nfft);
  const double h3 = g->dx[0] * g->dx[1] * g->dx[2];
  for (int j = 1; j <= L[1]; ++j) {
    double kj = (j <= L[1] / 2) ? kmin[1] + (double)(j - 1) * dk[1] + ksmp[1]
                                : kmin[1] + (double)(j - 1) * dk[1] - ksmp[1];
    for (int i = 1; i <= L[0]; ++i) {
      double ki = (i <= L[0] / 2) ? kmin[0] + (double)(i - 1) * dk[0] + ksmp[0]
                                  : kmin[0] + (double)(i - 1) * dk[0] - ksmp[0];
      for (int q = 0; q < 3; ++q) {
        for (int k = 0; k < nfft; ++k)
          f[q][k] = k < M[2] ? W[q + 3 * ((i - 1) + (size_t)L[0] * ((j - 1) + (size_t)L[1] * k))] : 0.0;
        or_glassman(f[q], tmp, nfft, -1);
      }
      for (int k = 1; k <= nfft; ++k) {
        double kk = (k <= L[2] / 2) ? kmin[2] + (double)(k - 1) * dk[2] + ksmp[2]
                                    : kmin[2] + (double)(k - 1) * dk[2] - ksmp[2];
        cplx f1 = f[0][k - 1], f2 = f[1][k - 1], f3 = f[2][k - 1];
        double zeta = 1.0;
        if (ifft_vort) {
          double s = eps * 2.0 * pi * sqrt(ki * ki + kj * kj + kk * kk);
          double s2 = s * s, e = exp(-0.5 * s2);
          switch (p->poisson_order) {
            case 0: zeta = 1.0; break;
            case 2: zeta = 1.0 * e; break;
            case 4: zeta = (1.0 + 0.5 * s2) * e; break;
            case 6: zeta = (1.0 + 0.5 * s2 + 0.125 * (s2 * s2)) * e; break;
            case 8: zeta = (1.0 + 0.5 * s2 + 0.125 * (s2 * s2) + (1.0 / 48.0) * (s2 * s2 * s2)) * e; break;
            default:
              zeta = (1.0 + 0.5 * s2 + 0.125 * (s2 * s2) + (1.0 / 48.0) * (s2 * s2 * s2) +
                      (1.0 / 384.0) * (s2 * s2 * s2 * s2)) * e;
              break;
          }
          o[3][k - 1] = f1 * zeta;
          o[4][k - 1] = f2 * zeta;
          o[5][k - 1] = f3 * zeta;
        }
        size_t gi = (size_t)(i - 1) + (size_t)L[0] * ((j - 1) + (size_t)L[1] * (k - 1));
        if (p->poisson_kernel == 1) {
          cplx G;
          if (allper) {
            double kmag = ki * ki + kj * kj + kk * kk;
            G = kmag > 0.0 ? zeta / (4.0 * (pi * pi) * kmag) : 0.0;
          } else {
            G = Ghat[gi] * g->dx[0] * g->dx[1] * g->dx[2];
          }
          cplx tp = G * 2.0 * pi;
          o[0][k - 1] = tp * (CMPLX(0.0, kj) * f3 - CMPLX(0.0, kk) * f2);
          o[1][k - 1] = tp * (CMPLX(0.0, kk) * f1 - CMPLX(0.0, ki) * f3);
          o[2][k - 1] = tp * (CMPLX(0.0, ki) * f2 - CMPLX(0.0, kj) * f1);
          if (ifft_psi) {
            o[6][k - 1] = G * f1;
            o[7][k - 1] = G * f2;
            o[8][k - 1] = G * f3;
          }
          if (reproj) {
            cplx dp = tp * (CMPLX(0.0, ki) * f1 + CMPLX(0.0, kj) * f2 + CMPLX(0.0, kk) * f3);
            o[3][k - 1] += 2.0 * pi * CMPLX(0.0, ki) * dp;
            o[4][k - 1] += 2.0 * pi * CMPLX(0.0, kj) * dp;
            o[5][k - 1] += 2.0 * pi * CMPLX(0.0, kk) * dp;
          }
        } else {
          cplx Kx, Ky, Kz;
          if (allper) {
            double kmag = ki * ki + kj * kj + kk * kk;
            if (kmag > 0.0) {
              Kx = zeta * CMPLX(0.0, ki) / (2.0 * pi * kmag);
              Ky = zeta * CMPLX(0.0, kj) / (2.0 * pi * kmag);
              Kz = zeta * CMPLX(0.0, kk) / (2.0 * pi * kmag);
            } else {
              Kx = Ky = Kz = 0.0;
            }
          } else {
            Kx = Ghat[gi * 3 + 0] * g->dx[0] * g->dx[1] * g->dx[2];
            Ky = Ghat[gi * 3 + 1] * g->dx[0] * g->dx[1] * g->dx[2];
            Kz = Ghat[gi * 3 + 2] * g->dx[0] * g->dx[1] * g->dx[2];
          }
          o[0][k - 1] = Ky * f3 - Kz * f2;
          o[1][k - 1] = Kz * f1 - Kx * f3;
          o[2][k - 1] = Kx * f2 - Ky * f1;
          if (reproj) {
            cplx dv = 2.0 * pi * (CMPLX(0.0, ki) * f1 + CMPLX(0.0, kj) * f2 + CMPLX(0.0, kk) * f3);
            o[3][k - 1] += Kx * dv;
            o[4][k - 1] += Ky * dv;
            o[5][k - 1] += Kz * dv;
          }
        }
      }
      (void)h3;
      int nout = 3 + (ifft_vort ? 3 : 0);
      for (int q = 0; q < 9; ++q) {
        if (q >= 3 && q < 6 && !ifft_vort) continue;
        if (q >= 6 && !ifft_psi) continue;
        or_glassman(o[q], tmp, nfft, 1);
        cplx *dst = q < 3 ? V : (q < 6 ? Wf : P);
        for (int k = 0; k < M[2]; ++k)
          dst[(q % 3) + 3 * ((i - 1) + (size_t)L[0] * ((j - 1) + (size_t)L[1] * k))] =
              o[q][k] / (double)nfft;
      }
      (void)nout;
    }
  }
  for (int q = 0; q < 3; ++q) free(f[q]);
  for (int q = 0; q < 9; ++q) free(o[q]);
  free(tmp);
  free(W);
  /* inverse y then x passes, each divided by its length, then the real part */
  cplx *outs[3] = {V, Wf, P};
  double *dsts[3] = {vel, vort_out, vecpot};
  for (int q = 0; q < 3; ++q) {
    if (!outs[q]) continue;
    pass_1d(outs[q], 3, nz, 1, 1, 0, (double)L[1], L[1], L[1]);
    pass_1d(outs[q], 3, nz, 0, 1, 0, (double)L[0], L[0], L[0]);
    double *dst = dsts[q];
    if (!dst) continue;
    for (int k = 0; k < M[2]; ++k)
      for (int j = 0; j < M[1]; ++j)
        for (int i = 0; i < M[0]; ++i)
          for (int c = 0; c < 3; ++c)
            dst[(size_t)c + 3 * ((size_t)(i + lo[0]) + (size_t)E[0] * ((j + lo[1]) + (size_t)E[1] * (k + lo[2])))] =
                creal(outs[q][c + 3 * (i + (size_t)L[0] * (j + (size_t)L[1] * k))]);
    /* periodic ghost fill, x then y then z (mesh/pmlib_mesh_map_ghost.f,
     * incl_edges = .FALSE., op = replace) */
    for (int d = 0; d < 3; ++d) {
      if (g->bound_cond[d] == 0) continue;
      int ng = ng_side(g, d), N = g->ncell[d];
      size_t st[3] = {3, 3 * (size_t)E[0], 3 * (size_t)E[0] * E[1]};
      int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
      for (int b = 0; b < E[o2]; ++b)
        for (int a = 0; a < E[o1]; ++a)
          for (int c = 0; c < 3; ++c) {
            size_t base = (size_t)c + st[o1] * a + st[o2] * b;
            for (int t = 0; t < ng; ++t) {
              /* low ghost t <- interior N-ng+t ; high ghost ng+N+t <- interior ng+t */
              dst[base + st[d] * t] = dst[base + st[d] * (N + t)];
              dst[base + st[d] * (ng + N + t)] = dst[base + st[d] * (ng + t)];
            }
          }
    }
  }
  free(V);
  free(Wf);
  free(P);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Interpolation weights                                                     */

/* M'4: W(d) for the 4 stencil nodes i2-1..i2+2 given the offset d2 in [0,1)
 * (interpolation/pmlib_interp_particle_mesh.f:182-214). */
static void w_m4(double d2, double *a) {
  double d1 = d2 + 1.0, d3 = 1.0 - d2, d4 = 2.0 - d2;
  a[0] = 0.5 * ((2.0 - d1) * (2.0 - d1)) * (1.0 - d1);
  a[1] = 1.0 - 2.5 * (d2 * d2) + 1.5 * (d2 * d2 * d2);
  a[2] = 1.0 - 2.5 * (d3 * d3) + 1.5 * (d3 * d3 * d3);
  a[3] = 0.5 * ((2.0 - d4) * (2.0 - d4)) * (1.0 - d4);
}

/* M6*: the 6 nodes i3-2..i3+3 given d3 (pmlib_interp_particle_mesh.f:565-602). */
static void w_m6(double d3, double *a) {
  const double c1 = -1.0 / 24.0, c2 = 1.0 / 24.0, c3 = -1.0 / 12.0;
  double d[6] = {d3 + 2.0, d3 + 1.0, d3, 1.0 - d3, 2.0 - d3, 3.0 - d3};
  for (int q = 0; q < 6; ++q) {
    double x = d[q], x2 = x * x, x3 = x2 * x, x4 = x3 * x;
    if (q == 0 || q == 5) {
      double t = x - 3.0;
      a[q] = c1 * (x - 2.0) * (t * t * t) * (5.0 * x - 8.0);
    } else if (q == 1 || q == 4) {
      a[q] = c2 * (x - 1.0) * (x - 2.0) * (25.0 * x3 - 114.0 * x2 + 153.0 * x - 48.0);
    } else {
      a[q] = c3 * (x - 1.0) * (25.0 * x4 - 38.0 * x3 - 3.0 * x2 + 12.0 * x + 12.0);
    }
  }
}

/* The particle -> cell index NINT((x - xmin) * (1/h)): a correctly rounded
 * subtract, a correctly rounded multiply, round half away from zero. */
int32_t or_cell_index(double x, double xmin, double c_1_dx) {
  double t = (x - xmin) * c_1_dx;
  return (int32_t)round(t);
}

void or_cell_indices(const or_geom *g, const double *pos, int64_t np, int32_t *idx) {
  for (int64_t q = 0; q < np; ++q)
    for (int d = 0; d < 3; ++d) idx[3 * q + d] = or_cell_index(pos[3 * q + d], g->xmin[d], 1.0 / g->dx[d]);
}

/* Forward/reverse periodic halo of a (3, E...) mesh along axis d.
 * add = 1: reverse-halo add (incl_edges, op = add) used after deposition. */
static void halo_periodic(const or_geom *g, double *m, int d, int add) {
  int E[3];
  for (int q = 0; q < 3; ++q) E[q] = mesh_ext(g, q);
  int ng = ng_side(g, d), N = g->ncell[d];
  size_t st[3] = {3, 3 * (size_t)E[0], 3 * (size_t)E[0] * E[1]};
  int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
  double *buf = (double *)malloc(sizeof(double) * 4 * ng);
  for (int b = 0; b < E[o2]; ++b)
    for (int a = 0; a < E[o1]; ++a)
      for (int c = 0; c < 3; ++c) {
        size_t base = (size_t)c + st[o1] * a + st[o2] * b;
        if (add) {
          /* cells [1-ng, ng] (mesh 0..2ng-1) and [N-ng+1, N+ng] (mesh N..N+2ng-1)
           * exchange and add, both packed before either is unpacked */
          for (int t = 0; t < 2 * ng; ++t) {
            buf[t] = m[base + st[d] * t];
            buf[2 * ng + t] = m[base + st[d] * (N + t)];
          }
          for (int t = 0; t < 2 * ng; ++t) {
            m[base + st[d] * (N + t)] += buf[t];
            m[base + st[d] * t] += buf[2 * ng + t];
          }
        } else {
          for (int t = 0; t < ng; ++t) {
            m[base + st[d] * t] = m[base + st[d] * (N + t)];
            m[base + st[d] * (ng + N + t)] = m[base + st[d] * (ng + t)];
          }
        }
      }
  free(buf);
}

/* pmlib_interp_particle_mesh: deposit part (3,np) at xpart (3,np) onto mesh.
 * clear = 1 zeroes the whole mesh first. */
int or_p2m(const or_geom *g, const double *xpart, const double *part, int64_t np, double *mesh,
           int clear) {
  int E[3], lo[3], hi[3], ngl[3];
  for (int d = 0; d < 3; ++d) {
    E[d] = mesh_ext(g, d);
    ngl[d] = ng_side(g, d);
    lo[d] = 1 - ngl[d];
    hi[d] = g->ncell[d] + ngl[d];
  }
  if (clear) memset(mesh, 0, sizeof(double) * 3 * (size_t)E[0] * E[1] * E[2]);
  const int m6 = g->interp_order == 4;
  const int ns = m6 ? 6 : 4, off = m6 ? 2 : 1;
  double c1d[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
  for (int64_t q = 0; q < np; ++q) {
    int base[3], idx[3][6];
    double w[3][6];
    int skip = 0;
    for (int d = 0; d < 3; ++d) {
      base[d] = or_cell_index(xpart[3 * q + d], g->xmin[d], c1d[d]);
      if (base[d] - off + ns - 1 < lo[d] || base[d] - off > hi[d]) skip = 1;
    }
    if (skip) continue;
    for (int d = 0; d < 3; ++d) {
      double dd = (xpart[3 * q + d] - g->xmin[d] - g->dx[d] * ((double)base[d] - 0.5)) * c1d[d];
      if (m6) w_m6(dd, w[d]); else w_m4(dd, w[d]);
      for (int s = 0; s < ns; ++s) idx[d][s] = base[d] - off + s;
    }
    for (int d = 0; d < 3; ++d) {
      if (g->bound_cond[d] == 0) {
        for (int s = 0; s < ns; ++s) {
          if (idx[d][s] < lo[d]) { w[d][s] = 0.0; idx[d][s] = lo[d]; }
          if (idx[d][s] > hi[d]) { w[d][s] = 0.0; idx[d][s] = hi[d]; }
        }
      } else {
        if (xpart[3 * q + d] < g->xmin[d]) {
          for (int s = 0; s < ns; ++s) { w[0][s] = w[1][s] = w[2][s] = 0.0; idx[d][s] = lo[d]; }
        }
        if (xpart[3 * q + d] > g->xmax[d]) {
          for (int s = 0; s < ns; ++s) { w[0][s] = w[1][s] = w[2][s] = 0.0; idx[d][s] = hi[d]; }
        }
      }
    }
    for (int c3 = 0; c3 < ns; ++c3)
      for (int c2 = 0; c2 < ns; ++c2)
        for (int c1 = 0; c1 < ns; ++c1) {
          double wt = w[0][c1] * w[1][c2] * w[2][c3];
          size_t o = 3 * ((size_t)(idx[0][c1] - lo[0]) +
                          (size_t)E[0] * ((idx[1][c2] - lo[1]) + (size_t)E[1] * (idx[2][c3] - lo[2])));
          for (int c = 0; c < 3; ++c) mesh[o + c] += wt * part[3 * q + c];
        }
  }
  for (int d = 0; d < 3; ++d)
    if (g->bound_cond[d] == 1) halo_periodic(g, mesh, d, 1);
  return 0;
}

/* pmlib_interp_mesh_particle for one mesh field: out(3,np) = sum W * mesh.
 * The M'4 branch normalises the y and z offsets by 1/dx(1) as the reference
 * does (interpolation/pmlib_interp_mesh_particle.f:114-116). */
int or_m2p(const or_geom *g, const double *mesh, const double *pos, int64_t np, double *out) {
  int E[3], lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    E[d] = mesh_ext(g, d);
    lo[d] = 1 - ng_side(g, d);
    hi[d] = g->ncell[d] + ng_side(g, d);
  }
  const int m6 = g->interp_order == 4;
  const int ns = m6 ? 6 : 4, off = m6 ? 2 : 1;
  double c1d[3] = {1.0 / g->dx[0], 1.0 / g->dx[1], 1.0 / g->dx[2]};
  for (int64_t q = 0; q < np; ++q) {
    int idx[3][6];
    double w[3][6];
    for (int d = 0; d < 3; ++d) {
      int b = or_cell_index(pos[3 * q + d], g->xmin[d], c1d[d]);
      double nrm = m6 ? c1d[d] : c1d[0];
      double dd = (pos[3 * q + d] - g->xmin[d] - g->dx[d] * ((double)b - 0.5)) * nrm;
      if (m6) w_m6(dd, w[d]); else w_m4(dd, w[d]);
      for (int s = 0; s < ns; ++s) idx[d][s] = b - off + s;
    }
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < ns; ++s) {
        if (idx[d][s] < lo[d]) {
          /* M6* z-side quirk (interpolation/pmlib_interp_mesh_particle.f:587-590):
           * the k5/k6 low-side test zeroes ay5/ay6 and sets j5/j6 to 1-nghost(5);
           * az5/az6 are left alone there and the mesh is read below its lower
           * bound.  The y entries are reproduced; the out-of-range z read is
           * replaced by a zero weight (all other z weights are already zero
           * whenever k5 < 1-nghost(5), so only undefined terms change). */
          if (m6 && d == 2 && s >= 4) { w[1][s] = 0.0; idx[1][s] = lo[2]; }
          w[d][s] = 0.0;
          idx[d][s] = lo[d];
        }
        if (idx[d][s] > hi[d]) { w[d][s] = 0.0; idx[d][s] = hi[d]; }
      }
    /* j5/j6 = 1-nghost(5) can lie outside the y range when the z axis has more
     * ghost layers; their weight is zero, so the read is kept inside the mesh
     * (the reference reads a neighbouring plane there, a finite value times 0). */
    for (int s = 0; s < ns; ++s) {
      if (idx[1][s] < lo[1]) idx[1][s] = lo[1];
      if (idx[1][s] > hi[1]) idx[1][s] = hi[1];
    }
    double acc[3] = {0.0, 0.0, 0.0};
    int first = 1;
    for (int c3 = 0; c3 < ns; ++c3)
      for (int c2 = 0; c2 < ns; ++c2)
        for (int c1 = 0; c1 < ns; ++c1) {
          double wt = w[0][c1] * w[1][c2] * w[2][c3];
          size_t o = 3 * ((size_t)(idx[0][c1] - lo[0]) +
                          (size_t)E[0] * ((idx[1][c2] - lo[1]) + (size_t)E[1] * (idx[2][c3] - lo[2])));
          for (int c = 0; c < 3; ++c) acc[c] = first ? wt * mesh[o + c] : acc[c] + wt * mesh[o + c];
          first = 0;
        }
    for (int c = 0; c < 3; ++c) out[3 * q + c] = acc[c];
  }
  return 0;
}

/* pmlib_remesh (remesh/pmlib_remesh.f:5-233): count pass then create pass
 * over (k,j,i).  Returns the particle count; pos/vort/vel/dvort (3,cap) are
 * written when non-NULL and the count fits in cap.  dvort_m may be NULL
 * (treated as zero; the reference always reads mesh%dvort for the threshold,
 * :102-105); mask (1 byte per cell) may be NULL (the mask=.FALSE. case,
 * :107-113).  copy_dvort / copy_vel are the dvort / vel flags (:208-219). */
int64_t or_remesh(const or_geom *g, const double *vort_m, const double *dvort_m,
                  const double *vel_m, const uint8_t *mask, double min_vort, double min_dvort,
                  int copy_dvort, int copy_vel, int incl_ghost,
                  double *pos, double *vort, double *vel, double *dvort, int64_t cap) {
  int E[3], ngl[3];
  for (int d = 0; d < 3; ++d) { E[d] = mesh_ext(g, d); ngl[d] = ng_side(g, d); }
  int64_t n = 0;
  for (int k = 1 - ngl[2]; k <= g->ncell[2] + ngl[2]; ++k) {
    double pz = g->xmin[2] + ((double)(k - 1) + 0.5) * g->dx[2];
    for (int j = 1 - ngl[1]; j <= g->ncell[1] + ngl[1]; ++j) {
      double py = g->xmin[1] + ((double)(j - 1) + 0.5) * g->dx[1];
      for (int i = 1 - ngl[0]; i <= g->ncell[0] + ngl[0]; ++i) {
        double px = g->xmin[0] + ((double)(i - 1) + 0.5) * g->dx[0];
        int ii[3] = {i, j, k}, place = 1;
        for (int d = 0; d < 3; ++d) {
          int in = ii[d] >= 1 && ii[d] <= g->ncell[d];
          int gh = incl_ghost && g->bound_cond[d] == 0;
          if (!(in || gh)) place = 0;
        }
        size_t o = 3 * ((size_t)(i - 1 + ngl[0]) + (size_t)E[0] * ((j - 1 + ngl[1]) + (size_t)E[1] * (k - 1 + ngl[2])));
        double s = sqrt(vort_m[o] * vort_m[o] + vort_m[o + 1] * vort_m[o + 1] + vort_m[o + 2] * vort_m[o + 2]);
        double ds = 0.0;
        if (dvort_m)
          ds = sqrt(dvort_m[o] * dvort_m[o] + dvort_m[o + 1] * dvort_m[o + 1] + dvort_m[o + 2] * dvort_m[o + 2]);
        if (!(place && (s > min_vort || ds > min_dvort))) continue;
        if (mask && !mask[o / 3]) continue;
        if (pos && n < cap) {
          pos[3 * n] = px; pos[3 * n + 1] = py; pos[3 * n + 2] = pz;
          for (int c = 0; c < 3; ++c) {
            vort[3 * n + c] = vort_m[o + c];
            if (vel) vel[3 * n + c] = (copy_vel && vel_m) ? vel_m[o + c] : 0.0;
            if (dvort) dvort[3 * n + c] = (copy_dvort && dvort_m) ? dvort_m[o + c] : 0.0;
          }
        }
        ++n;
      }
    }
  }
  return n;
}

/* Low-storage RK coefficients (particles/pmlib_particles_advance.f:49-88). */
static void rk_coeffs(int order, int stage, int euler_vort, double *a1, double *b1, double *c1,
                      double *a2, double *b2, double *c2) {
  double A1[3] = {0, 0, 0}, B1[3] = {1, 0, 0}, C1[3] = {1, 0, 0};
  double A2[3] = {0, 0, 0}, B2[3] = {1, 0, 0}, C2[3] = {1, 0, 0};
  if (order == 2) {
    A1[1] = -0.5; B1[1] = 1.0; C1[0] = 0.5; C1[1] = 1.0;
    if (euler_vort) { A2[1] = 1.0; B2[1] = 0.0; C2[0] = 0.5; C2[1] = 0.5; }
    else { A2[1] = -0.5; B2[1] = 1.0; C2[0] = 0.5; C2[1] = 1.0; }
  } else if (order == 3) {
    A1[1] = -5.0 / 9.0; A1[2] = -153.0 / 128.0; B1[1] = B1[2] = 1.0;
    C1[0] = 1.0 / 3.0; C1[1] = 15.0 / 16.0; C1[2] = 8.0 / 15.0;
    if (euler_vort) {
      A2[1] = A2[2] = 1.0; B2[1] = B2[2] = 0.0;
      C2[0] = C2[1] = C2[2] = 1.0 / 3.0;
    } else {
      A2[1] = -5.0 / 9.0; A2[2] = -153.0 / 128.0; B2[1] = B2[2] = 1.0;
      C2[0] = 1.0 / 3.0; C2[1] = 15.0 / 16.0; C2[2] = 8.0 / 15.0;
    }
  }
  int s = stage - 1;
  *a1 = A1[s]; *b1 = B1[s]; *c1 = C1[s]; *a2 = A2[s]; *b2 = B2[s]; *c2 = C2[s];
}

/* pmlib_particles_advance on one rank: the RK stage update, then per
 * direction the stable three-way partition of pmlib_particles_map
 * (moved west & wrapped, moved east & wrapped, staying; free sides drop)
 * applied to the arrays the map carries.  Arrays are (3, np) in/out; the
 * carried ones are compacted to the new count, which is returned. */
int64_t or_advance(const or_geom *g, int rk_order, int stage, double dt, int euler_vort,
                   double *pos, double *vel, double *vort, double *dvort, double *vel_rk,
                   double *dvort_rk, int64_t np) {
  double a1, b1, c1, a2, b2, c2;
  rk_coeffs(rk_order, stage, euler_vort, &a1, &b1, &c1, &a2, &b2, &c2);
  for (int64_t q = 0; q < np; ++q)
    for (int c = 0; c < 3; ++c) {
      size_t o = 3 * (size_t)q + c;
      vel_rk[o] = a1 * vel_rk[o] + b1 * dt * vel[o];
      dvort_rk[o] = a2 * dvort_rk[o] + b2 * dt * dvort[o];
      pos[o] = pos[o] + c1 * vel_rk[o];
      vort[o] = vort[o] + c2 * dvort_rk[o];
    }
  /* the arrays the map carries (pmlib_particles_advance.f:121-166): pos and
   * vort, and above first order vel_rk and dvort_rk; vel and dvort are not
   * carried and keep what the update left in them */
  double *arrs[4] = {pos, vort, vel_rk, dvort_rk};
  const int narr = rk_order > 1 ? 4 : 2;
  int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (np > 0 ? np : 1));
  double *tmp = (double *)malloc(sizeof(double) * 3 * (np > 0 ? np : 1));
  int8_t *cls = (int8_t *)malloc((size_t)(np > 0 ? np : 1));
  for (int d = 0; d < 3; ++d) {
    double lo = g->xmin[d], hi = g->xmax[d], len = g->xmax[d] - g->xmin[d];
    int per = g->bound_cond[d] != 0;
    /* repeated while a wrapped particle is still outside (the map_done loop
     * of particles/pmlib_particles_advance.f:145-166) */
    for (int again = 1; again;) {
      again = 0;
      for (int64_t q = 0; q < np; ++q) {
        double x = pos[3 * q + d];
        cls[q] = x < lo ? 1 : (x >= hi ? 2 : 0);
      }
      int64_t n = 0;
      for (int pass = 0; pass < 3; ++pass) {
        int want = pass == 0 ? 1 : (pass == 1 ? 2 : 0);
        for (int64_t q = 0; q < np; ++q) {
          if (cls[q] != want) continue;
          if (want != 0 && !per) continue;
          if (want == 1) pos[3 * q + d] += len;
          if (want == 2) pos[3 * q + d] -= len;
          perm[n++] = q;
        }
      }
      for (int a = 0; a < narr; ++a) {
        for (int64_t q = 0; q < n; ++q)
          for (int c = 0; c < 3; ++c) tmp[3 * q + c] = arrs[a][3 * perm[q] + c];
        memcpy(arrs[a], tmp, sizeof(double) * 3 * (size_t)n);
      }
      np = n;
      if (per)
        for (int64_t q = 0; q < np; ++q)
          if (pos[3 * q + d] < lo || pos[3 * q + d] >= hi) again = 1;
    }
  }
  free(perm);
  free(tmp);
  free(cls);
  return np;
}

/* One Brinkman penalisation iteration on a (3, E...) mesh, restating
 * penalisation_solve.f:96-111 and the FD branches :146-193 (2nd order) and
 * :195-273 (4th order):
 *   vel_pen = solid_vel - vel on all cells, f = chi * vel_pen,
 *   vort += relax * curl(f) on interior cells (central differences over 2h;
 *   4th order with c1 = 1/6, c2 = 4/3), accumulating the force increment and
 *   enstrophy sums.  chi (1, E...) is the per-cell factor (the sharp mask,
 *   or lambda dt chi_s / (1 + lambda dt chi_s) for the finite-lambda form);
 *   solid_vel may be NULL (solid at rest).  diag[0..4] = sum_dforce(3),
 *   res_enst, sum_enst_pen. */
int or_penalise(const or_geom *g, const double *vel, const double *chi, const double *solid_vel,
                double relax, int fd_order, double *vort, double *diag) {
  int E[3], ngl[3];
  for (int d = 0; d < 3; ++d) { E[d] = mesh_ext(g, d); ngl[d] = ng_side(g, d); }
  const int need = fd_order == 2 ? 1 : 2;
  if (fd_order != 2 && fd_order != 4) return -1;
  if (ngl[0] < need || ngl[1] < need || ngl[2] < need) return -1;
  size_t n = (size_t)E[0] * E[1] * E[2];
  double *f = (double *)malloc(sizeof(double) * 3 * n);
  for (size_t q = 0; q < n; ++q)
    for (int c = 0; c < 3; ++c) {
      double us = solid_vel ? solid_vel[3 * q + c] : 0.0;
      f[3 * q + c] = chi[q] * (us - vel[3 * q + c]);
    }
  const double c1 = 1.0 / 6.0, c2 = 4.0 / 3.0;
  double cx = 1.0 / (2.0 * g->dx[0]), cy = 1.0 / (2.0 * g->dx[1]), cz = 1.0 / (2.0 * g->dx[2]);
  size_t sx = 3, sy = 3 * (size_t)E[0], sz = 3 * (size_t)E[0] * E[1];
  double F0 = 0, F1 = 0, F2 = 0, re = 0, se = 0;
  for (int k = 1; k <= g->ncell[2]; ++k) {
    double pz = g->xmin[2] + ((double)(k - 1) + 0.5) * g->dx[2];
    for (int j = 1; j <= g->ncell[1]; ++j) {
      double py = g->xmin[1] + ((double)(j - 1) + 0.5) * g->dx[1];
      for (int i = 1; i <= g->ncell[0]; ++i) {
        double px = g->xmin[0] + ((double)(i - 1) + 0.5) * g->dx[0];
        size_t o = 3 * ((size_t)(i - 1 + ngl[0]) + (size_t)E[0] * ((j - 1 + ngl[1]) + (size_t)E[1] * (k - 1 + ngl[2])));
#define D4(c, s, cc) ((c1 * f[o - 2 * (s) + (c)] - c2 * f[o - (s) + (c)] + c2 * f[o + (s) + (c)] - c1 * f[o + 2 * (s) + (c)]) * (cc))
#define D2(c, s, cc) ((-f[o - (s) + (c)] + f[o + (s) + (c)]) * (cc))
        double dvdx, dwdx, dudy, dwdy, dudz, dvdz;
        if (fd_order == 4) {
          dvdx = D4(1, sx, cx); dwdx = D4(2, sx, cx);
          dudy = D4(0, sy, cy); dwdy = D4(2, sy, cy);
          dudz = D4(0, sz, cz); dvdz = D4(1, sz, cz);
        } else {
          dvdx = D2(1, sx, cx); dwdx = D2(2, sx, cx);
          dudy = D2(0, sy, cy); dwdy = D2(2, sy, cy);
          dudz = D2(0, sz, cz); dvdz = D2(1, sz, cz);
        }
#undef D4
#undef D2
        double w0 = relax * (dwdy - dvdz), w1 = relax * (dudz - dwdx), w2 = relax * (dvdx - dudy);
        vort[o] += w0;
        vort[o + 1] += w1;
        vort[o + 2] += w2;
        F0 += py * w2 - pz * w1;
        F1 += pz * w0 - px * w2;
        F2 += px * w1 - py * w0;
        re += w0 * w0 + w1 * w1 + w2 * w2;
        se += vort[o] * vort[o] + vort[o + 1] * vort[o + 1] + vort[o + 2] * vort[o + 2];
      }
    }
  }
  free(f);
  diag[0] = F0; diag[1] = F1; diag[2] = F2; diag[3] = re; diag[4] = se;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Vorticity right-hand side                                                 */

/* periodic forward ghost fill of a (3, E...) mesh, x then y then z
 * (mesh/pmlib_mesh_map_ghost.f, incl_edges = .FALSE., op = replace) */
static void ghost_fill_periodic(const or_geom *g, const int *E, double *m) {
  for (int d = 0; d < 3; ++d) {
    if (g->bound_cond[d] == 0) continue;
    int ng = ng_side(g, d), N = g->ncell[d];
    size_t st[3] = {3, 3 * (size_t)E[0], 3 * (size_t)E[0] * E[1]};
    int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
    for (int b = 0; b < E[o2]; ++b)
      for (int a = 0; a < E[o1]; ++a)
        for (int c = 0; c < 3; ++c) {
          size_t base = (size_t)c + st[o1] * a + st[o2] * b;
          for (int t = 0; t < ng; ++t) {
            m[base + st[d] * t] = m[base + st[d] * (N + t)];
            m[base + st[d] * (ng + N + t)] = m[base + st[d] * (ng + t)];
          }
        }
  }
}

/* Right-hand side of the vorticity equation on the interior cells
 * (part_deriv_vorticity.f:57-196), for each component l:
 *   dvort_l = visc (d2w_l/dx2 + d2w_l/dy2 + d2w_l/dz2)
 *             + w1 du_l/dx + w2 du_l/dy + w3 du_l/dz
 * fd_order 2: (1, -2, 1) / dx^2 and (-1, 1) / (2 dx) (:57-115);
 * fd_order 4: (c1, c2, -2.5, c2, c1) / dx^2 with c1 = -1/12, c2 = 4/3 and
 * (c3, -c4, c4, -c3) / (2 dx) with c3 = 1/6, c4 = 4/3 (:117-196); the
 * reference's operation order, left to right.  Then the periodic ghost map of
 * dvort (:201-208); free-side ghost cells of dvort are left unchanged.
 * Needs fd_order / 2 ghost layers on every side. */
int or_vort_rhs(const or_geom *g, const double *vort, const double *vel, double visc, int fd_order,
                double *dvort) {
  int E[3], ngl[3];
  for (int d = 0; d < 3; ++d) { E[d] = mesh_ext(g, d); ngl[d] = ng_side(g, d); }
  if (fd_order != 2 && fd_order != 4) return -1;
  const int need = fd_order / 2;
  if (ngl[0] < need || ngl[1] < need || ngl[2] < need) return -1;
  double cdx2[3], c2dx[3];
  for (int d = 0; d < 3; ++d) {
    cdx2[d] = 1.0 / (g->dx[d] * g->dx[d]);
    c2dx[d] = 1.0 / (2.0 * g->dx[d]);
  }
  const double c1 = -1.0 / 12.0, c2 = 4.0 / 3.0, c3 = 1.0 / 6.0, c4 = 4.0 / 3.0;
  const size_t st[3] = {3, 3 * (size_t)E[0], 3 * (size_t)E[0] * E[1]};
  for (int k = 0; k < g->ncell[2]; ++k)
    for (int j = 0; j < g->ncell[1]; ++j)
      for (int i = 0; i < g->ncell[0]; ++i) {
        size_t o = 3 * ((size_t)(i + ngl[0]) + (size_t)E[0] * ((j + ngl[1]) + (size_t)E[1] * (k + ngl[2])));
        double w1 = vort[o], w2 = vort[o + 1], w3 = vort[o + 2];
        for (int l = 0; l < 3; ++l) {
          double dd[3], du[3];
          for (int d = 0; d < 3; ++d) {
            const double *a = vort + o + l, *u = vel + o + l;
            size_t s = st[d];
            if (fd_order == 2) {
              dd[d] = (1.0 * a[-(ptrdiff_t)s] - 2.0 * a[0] + 1.0 * a[s]) * cdx2[d];
              du[d] = (-1.0 * u[-(ptrdiff_t)s] + 1.0 * u[s]) * c2dx[d];
            } else {
              dd[d] = (c1 * a[-2 * (ptrdiff_t)s] + c2 * a[-(ptrdiff_t)s] - 2.5 * a[0] + c2 * a[s] + c1 * a[2 * s]) *
                      cdx2[d];
              du[d] = (c3 * u[-2 * (ptrdiff_t)s] - c4 * u[-(ptrdiff_t)s] + c4 * u[s] - c3 * u[2 * s]) * c2dx[d];
            }
          }
          dvort[o + l] = visc * (dd[0] + dd[1] + dd[2]) + w1 * du[0] + w2 * du[1] + w3 * du[2];
        }
      }
  ghost_fill_periodic(g, E, dvort);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Global diagnostics                                                        */

/* Particle diagnostics of a vortex ring (ring axis x), in FP64
 * (vortex_codes/vortex_structures/source/diagnostics/diagnostics_ring.f:
 * first loop :72-158, centroid :166-168, radius and circulation :174-207,
 * core radius :211-226), one rank.  dx is the mesh spacing whose cell volume
 * weights every particle.  out[0..15] = impulse(3), angular impulse(3),
 * kinetic energy, enstrophy, helicity, circulation, vortex centroid(3),
 * vortex radius, core radius, max |vort|.  Each term is formed in the
 * reference's operation order; the sums run over particles in index order. */
void or_diag_particles(const double *dx, const double *pos, const double *vort, const double *vel, int64_t np,
                       double *out) {
  const double PI = acos(-1.0), c_1_3 = 1.0 / 3.0;
  double imp[3] = {0, 0, 0}, ang[3] = {0, 0, 0}, kin = 0, enst = 0, hel = 0, cen[3] = {0, 0, 0}, vmax = 0;
  for (int64_t i = 0; i < np; ++i) {
    const double *x = pos + 3 * i, *w = vort + 3 * i, *u = vel + 3 * i;
    double m1[3], m2[3];
    m1[0] = x[1] * w[2] - x[2] * w[1];
    m1[1] = x[2] * w[0] - x[0] * w[2];
    m1[2] = x[0] * w[1] - x[1] * w[0];
    m2[0] = x[1] * m1[2] - x[2] * m1[1];
    m2[1] = x[2] * m1[0] - x[0] * m1[2];
    m2[2] = x[0] * m1[1] - x[1] * m1[0];
    for (int d = 0; d < 3; ++d) imp[d] = imp[d] + 0.5 * m1[d] * dx[0] * dx[1] * dx[2];
    for (int d = 0; d < 3; ++d) ang[d] = ang[d] + c_1_3 * m2[d] * dx[0] * dx[1] * dx[2];
    kin = kin + (u[0] * m1[0] + u[1] * m1[1] + u[2] * m1[2]) * dx[0] * dx[1] * dx[2];
    const double e = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    enst = enst + e * dx[0] * dx[1] * dx[2];
    hel = hel + (u[0] * w[0] + u[1] * w[1] + u[2] * w[2]) * dx[0] * dx[1] * dx[2];
    for (int d = 0; d < 3; ++d) cen[d] = cen[d] + e * x[d] * dx[0] * dx[1] * dx[2];
    const double a = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (a > vmax) vmax = a;
  }
  for (int d = 0; d < 3; ++d) cen[d] = cen[d] / enst;
  double rad = 0, circ = 0;
  for (int64_t i = 0; i < np; ++i) {
    const double *x = pos + 3 * i, *w = vort + 3 * i;
    const double e = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double dy = x[1] - cen[1], dz = x[2] - cen[2];
    rad = rad + e * sqrt(dy * dy + dz * dz) * dx[0] * dx[1] * dx[2];
    circ = circ - dz / (dy * dy + dz * dz) * w[1] + dy / (dy * dy + dz * dz) * w[2];
  }
  rad = rad / enst;
  circ = circ * dx[0] * dx[1] * dx[2] / (2.0 * PI);
  double core = 0;
  for (int64_t i = 0; i < np; ++i) {
    const double *x = pos + 3 * i, *w = vort + 3 * i;
    const double e = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double dy = x[1] - cen[1], dz = x[2] - cen[2], dxx = x[0] - cen[0];
    const double s = sqrt(dy * dy + dz * dz) - rad;
    core = core + e * sqrt(s * s + dxx * dxx) * dx[0] * dx[1] * dx[2];
  }
  core = core / enst;
  for (int d = 0; d < 3; ++d) {
    out[d] = imp[d];
    out[3 + d] = ang[d];
    out[10 + d] = cen[d];
  }
  out[6] = kin;
  out[7] = enst;
  out[8] = hel;
  out[9] = circ;
  out[13] = rad;
  out[14] = core;
  out[15] = vmax;
}

/* Mesh energy diagnostics over the interior cells, in FP64
 * (vortex_codes/vortex_structures/source/diagnostics/diagnostics_energy_mesh.f:
 * :72-134), one rank.  out[0..4] = kinetic energy, enstrophy, helicity
 * energy, Lamb energy, max |vort|.  The helicity and Lamb terms divide by the
 * local enstrophy as the reference does (a cell with zero vorticity makes
 * them NaN there too).  Sums over k, j, i in that loop order. */
int or_diag_mesh(const or_geom *g, const double *vel, const double *vort, double *out) {
  int E[3], ngl[3];
  for (int d = 0; d < 3; ++d) { E[d] = mesh_ext(g, d); ngl[d] = ng_side(g, d); }
  const double *dx = g->dx;
  double kin = 0, enst = 0, hel = 0, lamb = 0, vmax = 0;
  for (int k = 0; k < g->ncell[2]; ++k)
    for (int j = 0; j < g->ncell[1]; ++j)
      for (int i = 0; i < g->ncell[0]; ++i) {
        size_t o = 3 * ((size_t)(i + ngl[0]) + (size_t)E[0] * ((j + ngl[1]) + (size_t)E[1] * (k + ngl[2])));
        const double *w = vort + o, *u = vel + o;
        const double e = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        kin = kin + 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * dx[0] * dx[1] * dx[2];
        enst = enst + e * dx[0] * dx[1] * dx[2];
        const double h = w[0] * u[0] + w[1] * u[1] + w[2] * u[2];
        hel = hel + h * h / e * dx[0] * dx[1] * dx[2];
        const double l0 = w[1] * u[2] - w[2] * u[1], l1 = w[2] * u[0] - w[0] * u[2], l2 = w[0] * u[1] - w[1] * u[0];
        lamb = lamb + (l0 * l0 + l1 * l1 + l2 * l2) / e * dx[0] * dx[1] * dx[2];
        const double a = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (a > vmax) vmax = a;
      }
  out[0] = kin;
  out[1] = enst;
  out[2] = hel;
  out[3] = lamb;
  out[4] = vmax;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Multiresolution: parent topology and mesh-to-mesh interpolation           */

/* pmlib_topology_parent (topology/pmlib_topology_parent.f:5-100) on one
 * rank: the parent cuboid over the same box with cells a = 2^(level -
 * parent level) times larger, 2 (M'4) or 3 (M6*) ghost layers, the same
 * boundary conditions.  Returns -1 if ncell is not a multiple of a. */
int or_topology_parent(const or_geom *g, int level, int parent_level, or_geom *out) {
  if (parent_level == 0) return -1;
  const double reldx = pow(2.0, (double)(level - parent_level));
  const int r = (int)lround(reldx);
  *out = *g;
  for (int d = 0; d < 3; ++d) {
    if (g->ncell[d] % r != 0) return -1;
    out->ncell[d] = (int)lround((double)g->ncell[d] / reldx);
    out->dx[d] = reldx * g->dx[d];
  }
  out->nghost = g->interp_order == 4 ? 3 : 2;
  return 0;
}

/* pmlib_interp_mesh_mesh (interpolation/pmlib_interp_mesh_mesh.f:5-2078) on
 * one rank for (3, E...) meshes, M'4 or M6* (interp_order of the
 * geometries): reldx = dx_from / dx_to.  Below 1 (patch to parent, :93-1392)
 * every cell of the source mesh, ghosts included, is deposited onto the
 * target with weights reldx W per axis (stencils cancelled as in P2M), then
 * the periodic ghost layers are added to their images; above 1 (parent to
 * patch, :1395-2062) every target cell (interior cells only on periodic
 * sides) adds the interpolated source value, stencil indices clamped into
 * the source mesh with zero weight, then the periodic ghosts are copied.
 * clear zeroes the target first.  Loop orders and term orders are the
 * reference's. */
int or_interp_mesh_mesh(const or_geom *gf, const or_geom *gt, const double *from, double *to, int clear) {
  const int m6 = gt->interp_order == 4, ns = m6 ? 6 : 4, off = m6 ? 2 : 1;
  int Ef[3], Et[3], nf[3], nt[3];
  for (int d = 0; d < 3; ++d) {
    Ef[d] = mesh_ext(gf, d); Et[d] = mesh_ext(gt, d);
    nf[d] = ng_side(gf, d); nt[d] = ng_side(gt, d);
  }
  const size_t Nt = (size_t)Et[0] * Et[1] * Et[2];
  if (clear) memset(to, 0, sizeof(double) * 3 * Nt);
  const double reldx = gf->dx[0] / gt->dx[0];
#define TO(c, i, j, k) to[(c) + 3 * ((size_t)((i) - 1 + nt[0]) + (size_t)Et[0] * (((j) - 1 + nt[1]) + (size_t)Et[1] * ((k) - 1 + nt[2])))]
#define FR(c, i, j, k) from[(c) + 3 * ((size_t)((i) - 1 + nf[0]) + (size_t)Ef[0] * (((j) - 1 + nf[1]) + (size_t)Ef[1] * ((k) - 1 + nf[2])))]
  if (reldx <= 0.9999) {
    double c1d[3];
    for (int d = 0; d < 3; ++d) c1d[d] = 1.0 / gt->dx[d];
    for (int k = 1 - nf[2]; k <= gf->ncell[2] + nf[2]; ++k)
      for (int j = 1 - nf[1]; j <= gf->ncell[1] + nf[1]; ++j)
        for (int i = 1 - nf[0]; i <= gf->ncell[0] + nf[0]; ++i) {
          const int ii[3] = {i, j, k};
          double p[3], w[3][6];
          int idx[3][6], b[3];
          for (int d = 0; d < 3; ++d) {
            p[d] = gf->xmin[d] + ((double)(ii[d] - 1) + 0.5) * gf->dx[d];
            b[d] = or_cell_index(p[d], gt->xmin[d], c1d[d]);
          }
          int inside = 1;
          for (int d = 0; d < 3; ++d)
            if (b[d] - off + ns - 1 < 1 - nt[d] || b[d] - off > gt->ncell[d] + nt[d]) inside = 0;
          if (!inside) continue;
          for (int d = 0; d < 3; ++d) {
            double dd = (p[d] - gt->xmin[d] - gt->dx[d] * ((double)b[d] - 0.5)) * c1d[d], a[6];
            if (m6) w_m6(dd, a);
            else w_m4(dd, a);
            for (int s = 0; s < ns; ++s) {
              idx[d][s] = b[d] - off + s;
              w[d][s] = reldx * a[s];
            }
          }
          /* cancellation (:159-290): free sides clamp each index, periodic
           * sides drop the whole cell when it lies outside the box */
          for (int d = 0; d < 3; ++d) {
            if (gt->bound_cond[d] == 0) {
              for (int s = 0; s < ns; ++s) {
                if (idx[d][s] < 1 - nt[d]) { w[d][s] = 0.0; idx[d][s] = 1 - nt[d]; }
                if (idx[d][s] > gt->ncell[d] + nt[d]) { w[d][s] = 0.0; idx[d][s] = gt->ncell[d] + nt[d]; }
              }
            } else if (p[d] < gt->xmin[d] || p[d] > gt->xmax[d]) {
              for (int e = 0; e < 3; ++e)
                for (int s = 0; s < ns; ++s) w[e][s] = 0.0;
              for (int s = 0; s < ns; ++s) idx[d][s] = p[d] < gt->xmin[d] ? 1 - nt[d] : gt->ncell[d] + nt[d];
            }
          }
          for (int sz = 0; sz < ns; ++sz)
            for (int sy = 0; sy < ns; ++sy)
              for (int sx = 0; sx < ns; ++sx) {
                const double ww = w[0][sx] * w[1][sy] * w[2][sz];
                for (int c = 0; c < 3; ++c)
                  TO(c, idx[0][sx], idx[1][sy], idx[2][sz]) = TO(c, idx[0][sx], idx[1][sy], idx[2][sz]) + ww * FR(c, i, j, k);
              }
        }
    /* ghost layers added to their periodic images (incl_edges, op add) */
    for (int d = 0; d < 3; ++d)
      if (gt->bound_cond[d] == 1) halo_periodic(gt, to, d, 1);
  } else if (reldx > 0.9999) {
    double c1d[3];
    for (int d = 0; d < 3; ++d) c1d[d] = 1.0 / gf->dx[d];
    for (int k = 1 - nt[2]; k <= gt->ncell[2] + nt[2]; ++k)
      for (int j = 1 - nt[1]; j <= gt->ncell[1] + nt[1]; ++j)
        for (int i = 1 - nt[0]; i <= gt->ncell[0] + nt[0]; ++i) {
          const int ii[3] = {i, j, k};
          int skip = 0;
          for (int d = 0; d < 3; ++d)
            if (gt->bound_cond[d] != 0 && (ii[d] < 1 || ii[d] > gt->ncell[d])) skip = 1;
          if (skip) continue;
          double w[3][6];
          int idx[3][6];
          for (int d = 0; d < 3; ++d) {
            double p = gt->xmin[d] + ((double)(ii[d] - 1) + 0.5) * gt->dx[d];
            int b = or_cell_index(p, gf->xmin[d], c1d[d]);
            double dd = (p - gf->xmin[d] - gf->dx[d] * ((double)b - 0.5)) * c1d[d], a[6];
            if (m6) w_m6(dd, a);
            else w_m4(dd, a);
            for (int s = 0; s < ns; ++s) {
              idx[d][s] = b - off + s;
              w[d][s] = a[s];
              if (idx[d][s] < 1 - nf[d]) { w[d][s] = 0.0; idx[d][s] = 1 - nf[d]; }
              if (idx[d][s] > gf->ncell[d] + nf[d]) { w[d][s] = 0.0; idx[d][s] = gf->ncell[d] + nf[d]; }
            }
          }
          for (int c = 0; c < 3; ++c) {
            double acc = TO(c, i, j, k);
            for (int sz = 0; sz < ns; ++sz)
              for (int sy = 0; sy < ns; ++sy)
                for (int sx = 0; sx < ns; ++sx)
                  acc = acc + w[0][sx] * w[1][sy] * w[2][sz] * FR(c, idx[0][sx], idx[1][sy], idx[2][sz]);
            TO(c, i, j, k) = acc;
          }
        }
    /* periodic ghosts copied from their images (op replace) */
    int E[3] = {Et[0], Et[1], Et[2]};
    ghost_fill_periodic(gt, E, to);
  }
#undef TO
#undef FR
  return 0;
}
