// Particle <-> mesh kernels (HIP, gfx950): M'4 / M6* deposition and gather,
// remeshing and the low-storage RK update with one-rank migration.
//
// Deposition is a gather, not a scatter: particles are binned by the cell of
// their stencil base index (bit-exact NINT), a stable radix sort orders
// them by bin and, inside a bin, by particle id, each sorted particle is
// staged once as a record of its kernel weights and strength, and every mesh
// node sums the records of the ns^2 stencil rows that cover it (each row one
// contiguous range of records), in a fixed order.  No atomics touch the
// mesh, so the result is deterministic run to run.
#pragma once
#include "pm_common.h"

namespace pmgpu {

struct PGeom {
  int ncell[3], ngl[3], bc[3];
  int E[3];           // mesh extent incl. ghosts
  int lo[3], hi[3];   // node index range 1-ng .. N+ng
  double xmin[3], xmax[3], dx[3], c1dx[3];
  int ns, off;        // stencil width and base offset (M'4: 4,1; M6*: 6,2)
  int m6;
  int nb[3];          // bins per axis: base index b in [lo + off - ns + 1, hi + off]
  int blo[3];         // b of bin 0
};

// NINT((x - xmin) * c_1_dx): correctly rounded subtract and multiply (no
// contraction), then round half away from zero, done by hand: trunc is exact
// and so is t - trunc(t), so the half-way test is exact too.
__device__ __forceinline__ int cell_index(double x, double xmin, double c1) {
  const double t = __dmul_rn(__dsub_rn(x, xmin), c1);
  if (!(fabs(t) < 1073741824.0)) return t > 0 ? 1073741824 : (t < 0 ? -1073741824 : 0);  // far off the mesh
  double r = trunc(t);
  const double f = __dsub_rn(t, r);
  if (f >= 0.5) r = __dadd_rn(r, 1.0);
  else if (f <= -0.5) r = __dsub_rn(r, 1.0);
  return (int)r;
}

// normalised offset d = (x - xmin - h (b - 1/2)) * nrm
__device__ __forceinline__ double stencil_offset(double x, double xmin, double h, int b, double nrm) {
  double t = __dmul_rn(h, __dsub_rn((double)b, 0.5));
  return __dmul_rn(__dsub_rn(__dsub_rn(x, xmin), t), nrm);
}

// weight of stencil slot s for offset d (interp_particle_mesh.f:182-214, :565-602)
__device__ __forceinline__ double weight(int m6, int s, double d) {
  if (!m6) {
    switch (s) {
      case 0: { double u = __dadd_rn(d, 1.0), a = __dsub_rn(2.0, u);
                return __dmul_rn(__dmul_rn(0.5, __dmul_rn(a, a)), __dsub_rn(1.0, u)); }
      case 1: { double d2 = __dmul_rn(d, d), d3 = __dmul_rn(d2, d);
                return __dadd_rn(__dsub_rn(1.0, __dmul_rn(2.5, d2)), __dmul_rn(1.5, d3)); }
      case 2: { double u = __dsub_rn(1.0, d), u2 = __dmul_rn(u, u), u3 = __dmul_rn(u2, u);
                return __dadd_rn(__dsub_rn(1.0, __dmul_rn(2.5, u2)), __dmul_rn(1.5, u3)); }
      default: { double u = __dsub_rn(2.0, d), a = __dsub_rn(2.0, u);
                 return __dmul_rn(__dmul_rn(0.5, __dmul_rn(a, a)), __dsub_rn(1.0, u)); }
    }
  }
  const double c1 = -1.0 / 24.0, c2 = 1.0 / 24.0, c3 = -1.0 / 12.0;
  double x;
  switch (s) {
    case 0: x = __dadd_rn(d, 2.0); break;
    case 1: x = __dadd_rn(d, 1.0); break;
    case 2: x = d; break;
    case 3: x = __dsub_rn(1.0, d); break;
    case 4: x = __dsub_rn(2.0, d); break;
    default: x = __dsub_rn(3.0, d); break;
  }
  double x2 = __dmul_rn(x, x), x3 = __dmul_rn(x2, x), x4 = __dmul_rn(x3, x);
  if (s == 0 || s == 5) {
    double t = __dsub_rn(x, 3.0);
    double t3 = __dmul_rn(__dmul_rn(t, t), t);
    return __dmul_rn(__dmul_rn(__dmul_rn(c1, __dsub_rn(x, 2.0)), t3), __dsub_rn(__dmul_rn(5.0, x), 8.0));
  }
  if (s == 1 || s == 4) {
    double p = __dsub_rn(__dadd_rn(__dsub_rn(__dmul_rn(25.0, x3), __dmul_rn(114.0, x2)), __dmul_rn(153.0, x)), 48.0);
    return __dmul_rn(__dmul_rn(__dmul_rn(c2, __dsub_rn(x, 1.0)), __dsub_rn(x, 2.0)), p);
  }
  double p = __dadd_rn(__dadd_rn(__dsub_rn(__dsub_rn(__dmul_rn(25.0, x4), __dmul_rn(38.0, x3)), __dmul_rn(3.0, x2)),
                                 __dmul_rn(12.0, x)), 12.0);
  return __dmul_rn(__dmul_rn(c3, __dsub_rn(x, 1.0)), p);
}

// FP32 form of the kernel weights (same polynomials, single precision)
__device__ __forceinline__ float weight_f(int m6, int s, float d) {
  if (!m6) {
    switch (s) {
      case 0: { float u = d + 1.0f, a = 2.0f - u; return 0.5f * (a * a) * (1.0f - u); }
      case 1: return 1.0f - 2.5f * (d * d) + 1.5f * (d * d * d);
      case 2: { float u = 1.0f - d; return 1.0f - 2.5f * (u * u) + 1.5f * (u * u * u); }
      default: { float u = 2.0f - d, a = 2.0f - u; return 0.5f * (a * a) * (1.0f - u); }
    }
  }
  const float c1 = -1.0f / 24.0f, c2 = 1.0f / 24.0f, c3 = -1.0f / 12.0f;
  const float x = s == 0 ? d + 2.0f : (s == 1 ? d + 1.0f : (s == 2 ? d : (s == 3 ? 1.0f - d : (s == 4 ? 2.0f - d
                                                                                                      : 3.0f - d))));
  const float x2 = x * x, x3 = x2 * x, x4 = x3 * x;
  if (s == 0 || s == 5) {
    const float t = x - 3.0f;
    return c1 * (x - 2.0f) * (t * t * t) * (5.0f * x - 8.0f);
  }
  if (s == 1 || s == 4) return c2 * (x - 1.0f) * (x - 2.0f) * (25.0f * x3 - 114.0f * x2 + 153.0f * x - 48.0f);
  return c3 * (x - 1.0f) * (25.0f * x4 - 38.0f * x3 - 3.0f * x2 + 12.0f * x + 12.0f);
}

// compensated (Kahan) accumulation: sum + c carries the low-order bits the
// float sum drops (this file is compiled without contraction, so the
// compensation survives)
struct Kahan {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float v) {
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// bin key of a particle (nbins: no stencil node touches the mesh, or the
// particle lies outside a periodic axis, where the reference zeroes every
// weight); ids[q] = q feeds the stable key sort
__global__ void k_p2m_keys(PGeom g, const double *pos, int64_t np, int nbins, int *key, int *ids, int *count) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    int b[3];
    bool ok = true;
    for (int d = 0; d < 3; ++d) {
      double x = pos[3 * q + d];
      b[d] = cell_index(x, g.xmin[d], g.c1dx[d]);
      const int64_t bb = b[d];
      if (bb - g.off + g.ns - 1 < g.lo[d] || bb - g.off > g.hi[d]) ok = false;
      if (g.bc[d] != 0 && (x < g.xmin[d] || x > g.xmax[d])) ok = false;
    }
    int k = nbins;
    if (ok) {
      k = (b[0] - g.blo[0]) + g.nb[0] * ((b[1] - g.blo[1]) + g.nb[1] * (b[2] - g.blo[2]));
      atomicAdd(&count[k], 1);
    }
    key[q] = k;
    ids[q] = (int)q;
  }
}

// Staged particle record, in bin order: the 3 x ns kernel weights, the 3
// strength components and the base cell along x (a T-typed copy of the int).
template <typename T> struct PRec {
  static constexpr int kSlots = 24;  // 3 * 6 weights + 3 values + base x, padded
};

// Records of the sorted particles t < start[nbins]: weights from the FP64
// offsets (FP64: the reference's arithmetic; FP32: weights in single
// precision from the FP64 offset).
template <typename T>
__global__ void k_p2m_stage(PGeom g, const double *pos, const double *val, const int *ids, const int *start,
                            int nbins, T *rec) {
  const int nvalid = start[nbins];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nvalid; t += gridDim.x * blockDim.x) {
    const int q = ids[t];
    T *r = rec + (int64_t)t * PRec<T>::kSlots;
    int bx = 0;
    for (int d = 0; d < 3; ++d) {
      const double x = pos[3 * q + d];
      const int b = cell_index(x, g.xmin[d], g.c1dx[d]);
      if (d == 0) bx = b;
      const double dd = stencil_offset(x, g.xmin[d], g.dx[d], b, g.c1dx[d]);
      for (int s2 = 0; s2 < g.ns; ++s2) {
        if constexpr (sizeof(T) == sizeof(double))
          r[d * g.ns + s2] = weight(g.m6, s2, dd);
        else
          r[d * g.ns + s2] = weight_f(g.m6, s2, (float)dd);
      }
    }
    r[18] = (T)val[3 * q];
    r[19] = (T)val[3 * q + 1];
    r[20] = (T)val[3 * q + 2];
    r[21] = (T)bx;
  }
}

// One thread per mesh node: for each (sz, sy) stencil row the bins whose
// x stencils cover the node are consecutive, so their particles are one
// contiguous range of records.  Summation order: sz and sy descending, then
// bins ascending (sx descending), particles by id within a bin.  FP64 sums
// with the reference's rounding per term; FP32 sums are Kahan-compensated.
template <typename T>
__global__ void k_p2m_gather_staged(PGeom g, const int *start, const T *rec, T *mesh, int accumulate) {
  const int64_t nn = (int64_t)g.E[0] * g.E[1] * g.E[2];
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nn;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int n[3];
    n[0] = (int)(idx % g.E[0]) + g.lo[0];
    n[1] = (int)((idx / g.E[0]) % g.E[1]) + g.lo[1];
    n[2] = (int)(idx / ((int64_t)g.E[0] * g.E[1])) + g.lo[2];
    const int ns = g.ns;
    int kx_lo = n[0] + g.off - ns + 1 - g.blo[0], kx_hi = n[0] + g.off - g.blo[0];
    if (kx_lo < 0) kx_lo = 0;
    if (kx_hi > g.nb[0] - 1) kx_hi = g.nb[0] - 1;
    double d0 = 0.0, d1 = 0.0, d2 = 0.0;
    Kahan f0, f1, f2;
    if (kx_lo <= kx_hi) {
      for (int sz = ns - 1; sz >= 0; --sz) {
        const int kz = n[2] - sz + g.off - g.blo[2];
        if (kz < 0 || kz >= g.nb[2]) continue;
        for (int sy = ns - 1; sy >= 0; --sy) {
          const int ky = n[1] - sy + g.off - g.blo[1];
          if (ky < 0 || ky >= g.nb[1]) continue;
          const int krow = g.nb[0] * (ky + g.nb[1] * kz);
          const int t0 = start[krow + kx_lo], t1 = start[krow + kx_hi + 1];
          for (int t = t0; t < t1; ++t) {
            const T *r = rec + (int64_t)t * PRec<T>::kSlots;
            const int sx = n[0] - (int)r[21] + g.off;
            if constexpr (sizeof(T) == sizeof(double)) {
              const double w = __dmul_rn(__dmul_rn(r[sx], r[ns + sy]), r[2 * ns + sz]);
              d0 = __dadd_rn(d0, __dmul_rn(w, r[18]));
              d1 = __dadd_rn(d1, __dmul_rn(w, r[19]));
              d2 = __dadd_rn(d2, __dmul_rn(w, r[20]));
            } else {
              const float w = r[sx] * r[ns + sy] * r[2 * ns + sz];
              f0.add(w * r[18]);
              f1.add(w * r[19]);
              f2.add(w * r[20]);
            }
          }
        }
      }
    }
    T *m = mesh + 3 * idx;
    const T a0 = sizeof(T) == sizeof(double) ? (T)d0 : (T)f0.s;
    const T a1 = sizeof(T) == sizeof(double) ? (T)d1 : (T)f1.s;
    const T a2 = sizeof(T) == sizeof(double) ? (T)d2 : (T)f2.s;
    if (accumulate) {
      m[0] += a0;
      m[1] += a1;
      m[2] += a2;
    } else {
      m[0] = a0;
      m[1] = a1;
      m[2] = a2;
    }
  }
}

// Reverse periodic halo along axis d (mesh/pmlib_mesh_map_ghost.f with
// incl_edges = .TRUE., op = add): cells [0, 2ng) and [N, N + 2ng) exchange
// and add, both read before either is written.
template <typename T> __global__ void k_halo_add(T *m, PGeom g, int d) {
  const int E0 = g.E[0], E1 = g.E[1];
  int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
  int64_t st[3] = {3, 3 * (int64_t)E0, 3 * (int64_t)E0 * E1};
  const int ng = g.ngl[d], N = g.ncell[d];
  int64_t n = (int64_t)g.E[o1] * g.E[o2] * 3 * 2 * ng;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int c = idx % 3;
    int64_t r = idx / 3;
    int t = r % (2 * ng);
    r /= 2 * ng;
    int a1 = r % g.E[o1];
    int a2 = r / g.E[o1];
    int64_t base = c + st[o1] * a1 + st[o2] * a2;
    T lo = m[base + st[d] * t], hi = m[base + st[d] * (N + t)];
    m[base + st[d] * t] = lo + hi;
    m[base + st[d] * (N + t)] = hi + lo;
  }
}

// M2P: one thread per particle, the reference's term order
// (interp_mesh_particle.f:87-409; the M'4 branch scales the y and z offsets
// by 1/dx(1), the M6* branch reproduces the k5/k6 quirk of :587-590).
__global__ void k_m2p(PGeom g, const double *mesh, const double *pos, int64_t np, double *out) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    int idx[3][6];
    double w[3][6];
    for (int d = 0; d < 3; ++d) {
      double x = pos[3 * q + d];
      int b = cell_index(x, g.xmin[d], g.c1dx[d]);
      double nrm = g.m6 ? g.c1dx[d] : g.c1dx[0];
      double dd = stencil_offset(x, g.xmin[d], g.dx[d], b, nrm);
      for (int s = 0; s < g.ns; ++s) {
        w[d][s] = weight(g.m6, s, dd);
        idx[d][s] = b - g.off + s;
      }
    }
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < g.ns; ++s) {
        if (idx[d][s] < g.lo[d]) {
          if (g.m6 && d == 2 && s >= 4) {
            w[1][s] = 0.0;
            idx[1][s] = g.lo[2];
          }
          w[d][s] = 0.0;
          idx[d][s] = g.lo[d];
        }
        if (idx[d][s] > g.hi[d]) {
          w[d][s] = 0.0;
          idx[d][s] = g.hi[d];
        }
      }
    // The quirk can leave a y index outside the y range (when the z axis has
    // more ghost layers); its weight is zero, so the read is clamped into the
    // mesh instead of leaving the allocation.
    for (int s = 0; s < g.ns; ++s) idx[1][s] = min(max(idx[1][s], g.lo[1]), g.hi[1]);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    bool first = true;
    for (int sz = 0; sz < g.ns; ++sz)
      for (int sy = 0; sy < g.ns; ++sy)
        for (int sx = 0; sx < g.ns; ++sx) {
          double wt = __dmul_rn(__dmul_rn(w[0][sx], w[1][sy]), w[2][sz]);
          int64_t o = 3 * ((int64_t)(idx[0][sx] - g.lo[0]) +
                           (int64_t)g.E[0] * ((idx[1][sy] - g.lo[1]) + (int64_t)g.E[1] * (idx[2][sz] - g.lo[2])));
          double t0 = __dmul_rn(wt, mesh[o]), t1 = __dmul_rn(wt, mesh[o + 1]), t2 = __dmul_rn(wt, mesh[o + 2]);
          if (first) {
            a0 = t0;
            a1 = t1;
            a2 = t2;
            first = false;
          } else {
            a0 = __dadd_rn(a0, t0);
            a1 = __dadd_rn(a1, t1);
            a2 = __dadd_rn(a2, t2);
          }
        }
    out[3 * q] = a0;
    out[3 * q + 1] = a1;
    out[3 * q + 2] = a2;
  }
}

// FP32 form of M2P: same stencil and cancellation, FP32 weights, Kahan sums
__global__ void k_m2p_f32(PGeom g, const float *mesh, const double *pos, int64_t np, double *out) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    int idx[3][6];
    float w[3][6];
    for (int d = 0; d < 3; ++d) {
      double x = pos[3 * q + d];
      int b = cell_index(x, g.xmin[d], g.c1dx[d]);
      double nrm = g.m6 ? g.c1dx[d] : g.c1dx[0];
      float dd = (float)stencil_offset(x, g.xmin[d], g.dx[d], b, nrm);
      for (int s = 0; s < g.ns; ++s) {
        w[d][s] = weight_f(g.m6, s, dd);
        idx[d][s] = b - g.off + s;
      }
    }
    for (int d = 0; d < 3; ++d)
      for (int s = 0; s < g.ns; ++s) {
        if (idx[d][s] < g.lo[d]) {
          if (g.m6 && d == 2 && s >= 4) {
            w[1][s] = 0.0f;
            idx[1][s] = g.lo[2];
          }
          w[d][s] = 0.0f;
          idx[d][s] = g.lo[d];
        }
        if (idx[d][s] > g.hi[d]) {
          w[d][s] = 0.0f;
          idx[d][s] = g.hi[d];
        }
      }
    for (int s = 0; s < g.ns; ++s) idx[1][s] = min(max(idx[1][s], g.lo[1]), g.hi[1]);
    Kahan a0, a1, a2;
    for (int sz = 0; sz < g.ns; ++sz)
      for (int sy = 0; sy < g.ns; ++sy)
        for (int sx = 0; sx < g.ns; ++sx) {
          float wt = w[0][sx] * w[1][sy] * w[2][sz];
          int64_t o = 3 * ((int64_t)(idx[0][sx] - g.lo[0]) +
                           (int64_t)g.E[0] * ((idx[1][sy] - g.lo[1]) + (int64_t)g.E[1] * (idx[2][sz] - g.lo[2])));
          a0.add(wt * mesh[o]);
          a1.add(wt * mesh[o + 1]);
          a2.add(wt * mesh[o + 2]);
        }
    out[3 * q] = a0.s;
    out[3 * q + 1] = a1.s;
    out[3 * q + 2] = a2.s;
  }
}

__global__ void k_cell_index(PGeom g, const double *pos, int64_t np, int *out) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < 3 * np; q += (int64_t)gridDim.x * blockDim.x) {
    int d = q % 3;
    out[q] = cell_index(pos[q], g.xmin[d], g.c1dx[d]);
  }
}

// ------------------------------------------------------------ remesh
// flag of mesh cell idx (remesh/pmlib_remesh.f:104-125)
template <typename T>
__device__ __forceinline__ int remesh_flag(const PGeom &g, int64_t idx, const T *vort, const T *dvort,
                                           const uint8_t *mask, double min_vort, double min_dvort,
                                           int incl_ghost) {
  int ii[3];
  ii[0] = (int)(idx % g.E[0]) + g.lo[0];
  ii[1] = (int)((idx / g.E[0]) % g.E[1]) + g.lo[1];
  ii[2] = (int)(idx / ((int64_t)g.E[0] * g.E[1])) + g.lo[2];
  for (int d = 0; d < 3; ++d) {
    bool in = ii[d] >= 1 && ii[d] <= g.ncell[d];
    bool gh = incl_ghost && g.bc[d] == 0;
    if (!(in || gh)) return 0;
  }
  const T *v = vort + 3 * idx;
  T s = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  T ds = 0;
  if (dvort) {
    const T *u = dvort + 3 * idx;
    ds = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
  }
  if (!(s > min_vort || ds > min_dvort)) return 0;
  return mask ? (mask[idx] ? 1 : 0) : 1;
}

template <typename T>
__global__ void k_remesh_flags(PGeom g, const T *vort, const T *dvort, const uint8_t *mask,
                               double min_vort, double min_dvort, int incl_ghost, int *flag) {
  const int64_t nn = (int64_t)g.E[0] * g.E[1] * g.E[2];
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nn;
       idx += (int64_t)gridDim.x * blockDim.x)
    flag[idx] = remesh_flag(g, idx, vort, dvort, mask, min_vort, min_dvort, incl_ghost);
}

// create particles at cell centres in (k, j, i) order from the scanned flags;
// vel / dvort are the fields copied (NULL: zero), per the vel / dvort flags
template <typename T>
__global__ void k_remesh_create(PGeom g, const int *flag, const int *offs, const T *vort, const T *dvort,
                                const T *vel, double *pos, double *pvort, double *pvel, double *pdvort,
                                double *vel_rk, double *dvort_rk) {
  const int64_t nn = (int64_t)g.E[0] * g.E[1] * g.E[2];
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nn;
       idx += (int64_t)gridDim.x * blockDim.x) {
    if (!flag[idx]) continue;
    int64_t p = offs[idx];
    int ii[3];
    ii[0] = (int)(idx % g.E[0]) + g.lo[0];
    ii[1] = (int)((idx / g.E[0]) % g.E[1]) + g.lo[1];
    ii[2] = (int)(idx / ((int64_t)g.E[0] * g.E[1])) + g.lo[2];
    for (int d = 0; d < 3; ++d)
      pos[3 * p + d] = __dadd_rn(g.xmin[d], __dmul_rn(__dadd_rn((double)(ii[d] - 1), 0.5), g.dx[d]));
    for (int c = 0; c < 3; ++c) {
      pvort[3 * p + c] = (double)vort[3 * idx + c];
      if (pvel) pvel[3 * p + c] = vel ? (double)vel[3 * idx + c] : 0.0;
      if (pdvort) pdvort[3 * p + c] = dvort ? (double)dvort[3 * idx + c] : 0.0;
      if (vel_rk) vel_rk[3 * p + c] = 0.0;
      if (dvort_rk) dvort_rk[3 * p + c] = 0.0;
    }
  }
}

// ------------------------------------------------------------ advance
struct RkCoef {
  double a1, b1, c1, a2, b2, c2;
};

__global__ void k_rk_update(RkCoef k, double dt, int64_t np, double *pos, const double *vel, double *vort,
                            const double *dvort, double *vel_rk, double *dvort_rk) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 3 * np; i += (int64_t)gridDim.x * blockDim.x) {
    double vr = __dadd_rn(__dmul_rn(k.a1, vel_rk[i]), __dmul_rn(__dmul_rn(k.b1, dt), vel[i]));
    double dr = __dadd_rn(__dmul_rn(k.a2, dvort_rk[i]), __dmul_rn(__dmul_rn(k.b2, dt), dvort[i]));
    vel_rk[i] = vr;
    dvort_rk[i] = dr;
    pos[i] = __dadd_rn(pos[i], __dmul_rn(k.c1, vr));
    vort[i] = __dadd_rn(vort[i], __dmul_rn(k.c2, dr));
  }
}

// migration class along axis d: 0 stay, 1 west (wrap), 2 east (wrap), 3 drop
__global__ void k_classify(int64_t np, const double *pos, int d, double lo, double hi, int per, int *cls,
                           int *counts) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    double x = pos[3 * q + d];
    int c = x < lo ? 1 : (x >= hi ? 2 : 0);
    if (c != 0 && !per) c = 3;
    cls[q] = c;
    if (c != 0) atomicAdd(&counts[c], 1);
  }
}

}  // namespace pmgpu
