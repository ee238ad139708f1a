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

// Particle count of a launch: a host value, or, in a step recorded once
// and replayed (HIP graph), the count in device memory that earlier kernels
// of the same step wrote; the grid then covers the capacity and threads past
// the count do nothing.
struct NCount {
  int64_t n;
  const int *dev;
  __host__ __device__ NCount(int64_t v = 0) : n(v), dev(nullptr) {}
  __host__ __device__ NCount(int64_t cap, const int *d) : n(cap), dev(d) {}
  __device__ __forceinline__ int64_t get() const { return dev ? (int64_t)*dev : n; }
};

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
// weight) and its arrival rank in the bin (an atomic count: the order within a
// bin is restored by particle id when the bin is staged)
static __global__ void k_p2m_keys(PGeom g, const double *pos, NCount npc, int nbins, int *key, int *rank, int *count) {
  const int64_t np = npc.get();
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
    int k = nbins, rk = 0;
    if (ok) {
      k = (b[0] - g.blo[0]) + g.nb[0] * ((b[1] - g.blo[1]) + g.nb[1] * (b[2] - g.blo[2]));
      rk = atomicAdd(&count[k], 1);
    }
    key[q] = k;
    rank[q] = rk;
  }
}

// counting-sort scatter: particle q to slot start[key] + rank of its bin
static __global__ void k_p2m_scatter(NCount npc, int nbins, const int *key, const int *rank, const int *start, int *ids) {
  const int64_t np = npc.get();
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    const int k = key[q];
    if (k < nbins) ids[start[k] + rank[q]] = (int)q;
  }
}

// ascending order of ids[0, n) in place: insertion sort for the usual bins of
// a few particles, heap sort (n log n, no extra memory) for crowded ones
__device__ inline void sort_ids(int *ids, int n) {
  if (n <= 16) {
    for (int i = 1; i < n; ++i) {
      const int v = ids[i];
      int j = i - 1;
      while (j >= 0 && ids[j] > v) {
        ids[j + 1] = ids[j];
        --j;
      }
      ids[j + 1] = v;
    }
    return;
  }
  auto sift = [&](int root, int end) {
    while (2 * root + 1 < end) {
      int c = 2 * root + 1;
      if (c + 1 < end && ids[c + 1] > ids[c]) ++c;
      if (ids[root] >= ids[c]) return;
      const int t = ids[root];
      ids[root] = ids[c];
      ids[c] = t;
      root = c;
    }
  };
  for (int i = n / 2 - 1; i >= 0; --i) sift(i, n);
  for (int end = n - 1; end > 0; --end) {
    const int t = ids[0];
    ids[0] = ids[end];
    ids[end] = t;
    sift(0, end);
  }
}

// Staged particle record, in bin order: the 3 x ns kernel weights in slots
// [d * ns + s], the 3 strength components in slots 3 ns .. 3 ns + 2 and the
// base cell along x (a T-typed copy of the int) in slot 3 ns + 3.  For M'4
// (ns = 4) a record is 16 slots: one 128-byte line in FP64; for M6* 24.
template <typename T, int NS> struct PRec {
  static constexpr int kSlots = NS == 4 ? 16 : 24;
  static constexpr int kVal = 3 * NS, kBx = 3 * NS + 3;
};

// One thread per bin: order the bin's particles by id (the order a stable
// sort by bin would give); bins of one particle, the usual case, do nothing.
static __global__ void __launch_bounds__(256) k_p2m_binsort(const int *start, int nbins, int *ids) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nbins; k += gridDim.x * blockDim.x) {
    const int t0 = start[k], n = start[k + 1] - t0;
    if (n > 1) sort_ids(ids + t0, n);
  }
}

// One thread per sorted particle t < start[nbins]: its record at slot t
// (weights from the FP64 offsets; FP64: the reference's arithmetic; FP32:
// weights in single precision from the FP64 offset).
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_p2m_stage(PGeom g, const double *pos, const double *val, const int *ids,
                                                   const int *start, int nbins, T *rec, int64_t rs) {
  const int nvalid = start[nbins];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nvalid; t += gridDim.x * blockDim.x) {
    const int q = ids[t];
    T wt[3 * NS];
    int bx = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double x = pos[3 * q + d];
      const int b = cell_index(x, g.xmin[d], g.c1dx[d]);
      if (d == 0) bx = b;
      const double dd = stencil_offset(x, g.xmin[d], g.dx[d], b, g.c1dx[d]);
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) {
        if constexpr (sizeof(T) == sizeof(double))
          wt[d * NS + s2] = weight(NS == 6, s2, dd);
        else
          wt[d * NS + s2] = weight_f(NS == 6, s2, (float)dd);
      }
    }
    // planar records: field i of record t at rec[i * rs + t], so that the
    // stores of consecutive threads, and the gather's loads of one field of
    // consecutive records, are contiguous
    T *r = rec + t;
#pragma unroll
    for (int i = 0; i < 3 * NS; ++i) r[i * rs] = wt[i];
    r[PRec<T, NS>::kVal * rs] = (T)val[3 * q];
    r[(PRec<T, NS>::kVal + 1) * rs] = (T)val[3 * q + 1];
    r[(PRec<T, NS>::kVal + 2) * rs] = (T)val[3 * q + 2];
    r[PRec<T, NS>::kBx * rs] = (T)bx;
  }
}

// Row-staged gather.  A workgroup is one wavefront and owns a line of 64
// consecutive mesh nodes along x (node ix = 64 blockIdx.x + lane at node
// row iy = blockIdx.y, iz = blockIdx.z); each lane sums its own node.  The
// node line is reached by the ns x ns stencil rows (ky, kz) = (iy + a, iz + b),
// a, b < ns, and within a row by the bins [64 blockIdx.x, 64 blockIdx.x + 63
// + ns - 1], one contiguous range of records.  Row by row the wavefront
// copies those records into LDS with coalesced loads (only the values the
// row needs: the ns x weights, the row's y and z weights, the strength and
// the bin), and each lane then sums the records of its own bins [ix, ix + ns)
// from LDS.  The next row's records are loaded into registers while the
// current row is summed.  Per node the terms arrive in the order kz
// ascending (sz descending), ky ascending (sy descending), records
// ascending (bins ascending, particles by id within a bin); FP64 terms are
// ((wx wy) wz) v rounded per operation as the reference does; FP32 terms
// are wx (wy wz) v, fused into one partial sum per stencil row, and the ns^2
// row partials are Kahan-added.
constexpr int kRowCap = 128;   // records staged per row pass (two per lane)

template <typename T, int NS> struct RowStage {
  static constexpr int NV = NS + 5;   // wx[ns], wy, wz, v0, v1, v2
  T val[kRowCap / 64][NV];
  int xs[kRowCap / 64];               // x slot base
};

template <typename T, int NS>
__device__ __forceinline__ void row_fetch(RowStage<T, NS> &st, const T *rec, int64_t rs, int t0, int n, int lane,
                                          int sy, int sz, int x0, int blo0) {
  constexpr int KV = PRec<T, NS>::kVal, KB = PRec<T, NS>::kBx;
#pragma unroll
  for (int h = 0; h < kRowCap / 64; ++h) {
    const int i = lane + 64 * h;
    // unconditional loads (no branch for the wait counters to merge over):
    // an empty row reads record t0 <= np, inside the allocation of np + 1.
    // FP32: field f of record t0 + li is read as (rec + f rs + t0)[li], a
    // wave-uniform base (t0 is) and a 32-bit byte offset, so that no load
    // needs per-lane 64-bit address arithmetic (gather 193 -> 188 us at 2.1M
    // particles); FP64 keeps per-lane addresses (measured faster there)
    const int li = i < n ? i : (n > 0 ? n - 1 : 0);
    const T *rb = rec + t0, *r = rb + li;
    const uint32_t lb = (uint32_t)li * (uint32_t)sizeof(T);
    auto fld = [&](int f) -> T {
      if constexpr (sizeof(T) == sizeof(float))
        return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(rb + (int64_t)f * rs) + lb);
      else
        return r[f * rs];
    };
#pragma unroll
    for (int s = 0; s < NS; ++s) st.val[h][s] = fld(s);
    if constexpr (sizeof(T) == sizeof(double)) {
      st.val[h][NS] = fld(NS + sy);
      st.val[h][NS + 1] = fld(2 * NS + sz);
    } else {
      // FP32: the record's y-z weight product, formed once for its ns lanes
      st.val[h][NS] = fld(NS + sy) * fld(2 * NS + sz);
      st.val[h][NS + 1] = (T)0;
    }
    st.val[h][NS + 2] = fld(KV);
    st.val[h][NS + 3] = fld(KV + 1);
    st.val[h][NS + 4] = fld(KV + 2);
    const T bx = fld(KB);
    // x slot of lane 0: sx = lane + (x0 + ns - 1 - kx)
    st.xs[h] = x0 + NS - 1 - ((int)bx - blo0);
  }
}

template <typename T, int NS>
__global__ void __launch_bounds__(64) k_p2m_rows(PGeom g, const int *start, const T *rec, int64_t rs, T *mesh,
                                                 int accumulate) {
  constexpr int NV = RowStage<T, NS>::NV, NR = NS * NS;
  constexpr bool F64 = sizeof(T) == sizeof(double);
  __shared__ T lds[NV][kRowCap];
  __shared__ int ldsx[kRowCap];
  const int lane = threadIdx.x;
  // XCD-aware line order: the workgroups of one XCD take consecutive lines
  // (x fastest, then y, then z), so that the record rows neighbouring lines
  // share are fetched into that XCD's L2 once rather than into every XCD's
  const int64_t lt = xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y * gridDim.z);
  const int x0 = (int)(lt % gridDim.x) * 64, iy = (int)((lt / gridDim.x) % gridDim.y),
            iz = (int)(lt / ((int64_t)gridDim.x * gridDim.y));
  const int ix = x0 + lane;
  const bool node_ok = ix < g.E[0];
  const int x1 = min(x0 + 64 + NS - 2, g.nb[0] - 1);
  const int blo0 = g.blo[0];
  // start[] offset of stencil row r: (ky, kz) = (iy + r % ns, iz + r / ns)
  auto krow = [&](int r) { return g.nb[0] * ((iy + r % NS) + g.nb[1] * (iz + r / NS)); };
  double d0 = 0.0, d1 = 0.0, d2 = 0.0;
  Kahan f0, f1, f2;
  float r0 = 0.0f, r1 = 0.0f, r2 = 0.0f;
  auto to_lds = [&](const RowStage<T, NS> &st, int m) {
#pragma unroll
    for (int h = 0; h < kRowCap / 64; ++h) {
      const int i = lane + 64 * h;
      if (i < m) {
#pragma unroll
        for (int v = 0; v < NV; ++v) lds[v][i] = st.val[h][v];
        ldsx[i] = st.xs[h];
      }
    }
  };
  // sum the lane's records t in [lo, hi) staged from record `base` on: the
  // first 8 as straight-line predicated code (no loop for the wait counters
  // to drain pending loads in front of), the rest in a loop
  auto term = [&](int i) {
    const int sx = lane + ldsx[i];   // in [0, ns) for the lane's bins
    const T wx = lds[sx][i];
    if constexpr (F64) {
      const double w = __dmul_rn(__dmul_rn(wx, lds[NS][i]), lds[NS + 1][i]);
      d0 = __dadd_rn(d0, __dmul_rn(w, lds[NS + 2][i]));
      d1 = __dadd_rn(d1, __dmul_rn(w, lds[NS + 3][i]));
      d2 = __dadd_rn(d2, __dmul_rn(w, lds[NS + 4][i]));
    } else {
      const float w = wx * lds[NS][i];
      r0 = __fmaf_rn(w, lds[NS + 2][i], r0);
      r1 = __fmaf_rn(w, lds[NS + 3][i], r1);
      r2 = __fmaf_rn(w, lds[NS + 4][i], r2);
    }
  };
  // FP32: a stencil row's terms (at most ns records per lane on a remeshed
  // set) go into one partial sum, and the row partials are Kahan-added
  auto row_done = [&]() {
    if constexpr (!F64) {
      f0.add(r0);
      f1.add(r1);
      f2.add(r2);
      r0 = r1 = r2 = 0.0f;
    }
  };
  auto sum = [&](int lo, int hi, int base) {
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (lo + u < hi) term(lo + u - base);
    for (int t = lo + 8; t < hi; ++t) term(t - base);
  };
  // every row's record range, loaded up front: uniform [rt0, rt1) over the
  // line's bins, per lane [ra0, ra1) over the lane's bins [ix, ix + ns)
  int rt0[NR], rt1[NR], ra0[NR], ra1[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int k = krow(r);
    // the line's range is the same for every lane: keep it in scalar registers
    rt0[r] = __builtin_amdgcn_readfirstlane(start[k + x0]);
    rt1[r] = __builtin_amdgcn_readfirstlane(start[k + x1 + 1]);
    ra0[r] = node_ok ? start[k + ix] : 0;
    ra1[r] = node_ok ? start[k + ix + NS] : 0;
  }
  int total = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) total += rt1[r] - rt0[r];
  if (total > 0) {
    // rows in order (the loop is unrolled: every index is static); row r + 1's
    // first pass loads while row r is summed
    RowStage<T, NS> cur, nxt;
    row_fetch<T, NS>(cur, rec, rs, rt0[0], min(rt1[0] - rt0[0], kRowCap), lane, NS - 1, NS - 1, x0, blo0);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const int T0 = rt0[r], n = rt1[r] - rt0[r], a0 = ra0[r], a1 = ra1[r];
      const int sy = NS - 1 - r % NS, sz = NS - 1 - r / NS;
      if (n > 0) {
        to_lds(cur, min(n, kRowCap));
        __syncthreads();
      }
      if (r + 1 < NR)
        row_fetch<T, NS>(nxt, rec, rs, rt0[min(r + 1, NR - 1)], min(rt1[min(r + 1, NR - 1)] - rt0[min(r + 1, NR - 1)], kRowCap),
                         lane, NS - 1 - (r + 1) % NS, NS - 1 - (r + 1) / NS, x0, blo0);
      if (n > 0) {
        sum(a0, min(a1, T0 + kRowCap), T0);
        __syncthreads();
        for (int p = kRowCap; p < n; p += kRowCap) {   // rows of more than kRowCap records
          const int m = min(n - p, kRowCap);
          row_fetch<T, NS>(cur, rec, rs, T0 + p, m, lane, sy, sz, x0, blo0);
          to_lds(cur, m);
          __syncthreads();
          sum(max(a0, T0 + p), min(a1, T0 + p + m), T0 + p);
          __syncthreads();
        }
        row_done();
      }
      cur = nxt;
    }
  }
  if (node_ok) {
    T *m = mesh + 3 * ((int64_t)ix + (int64_t)g.E[0] * (iy + (int64_t)g.E[1] * iz));
    const T a[3] = {F64 ? (T)d0 : (T)f0.s, F64 ? (T)d1 : (T)f1.s, F64 ? (T)d2 : (T)f2.s};
#pragma unroll
    for (int c = 0; c < 3; ++c) m[c] = accumulate ? m[c] + a[c] : a[c];
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
// (interp_mesh_particle.f:87-409: sz, sy, sx ascending; the M'4 branch scales
// the y and z offsets by 1/dx(1), the M6* branch reproduces the k5/k6 quirk of
// :587-590), FP64 rounded per operation as the reference does; the FP32 form
// (FP32 mesh) takes FP32 weights, forms their product in FP64 and sums the
// weighted FP32 node values in FP64 (the output is FP64; one fused
// multiply-add per term and component, against four FP32 operations of a
// compensated FP32 sum).
// Stencil weights and indices live in registers (the stencil width is a
// template parameter).
template <typename T, int NS>
__global__ void __launch_bounds__(256) k_m2p(PGeom g, const T *mesh, const double *pos, NCount npc, double *out) {
  const int64_t np = npc.get();
  constexpr bool F64 = sizeof(T) == sizeof(double);
  constexpr bool M6 = NS == 6;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    int idx[3][NS];
    T w[3][NS];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double x = pos[3 * q + d];
      const int b = cell_index(x, g.xmin[d], g.c1dx[d]);
      // the reference normalises every axis by 1/dx of x for M'4
      const double nrm = M6 ? g.c1dx[d] : g.c1dx[0];
      const double dd = stencil_offset(x, g.xmin[d], g.dx[d], b, nrm);
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) {
        if constexpr (F64)
          w[d][s2] = weight(M6, s2, dd);
        else
          w[d][s2] = weight_f(M6, s2, (float)dd);
        idx[d][s2] = b - g.off + s2;
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int s2 = 0; s2 < NS; ++s2) {
        if (idx[d][s2] < g.lo[d]) {
          if (M6 && d == 2 && s2 >= 4) {   // the reference's M6* quirk on the y slots
            w[1][s2] = (T)0;
            idx[1][s2] = g.lo[2];
          }
          w[d][s2] = (T)0;
          idx[d][s2] = g.lo[d];
        }
        if (idx[d][s2] > g.hi[d]) {
          w[d][s2] = (T)0;
          idx[d][s2] = g.hi[d];
        }
      }
    // The quirk can leave a y index outside the y range (when the z axis has
    // more ghost layers); its weight is zero, so the read is clamped into the
    // mesh instead of leaving the allocation.
#pragma unroll
    for (int s2 = 0; s2 < NS; ++s2) idx[1][s2] = min(max(idx[1][s2], g.lo[1]), g.hi[1]);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    // node (sx, sy, sz) at row base (sy, sz) + x offset sx: one 64-bit base
    // per stencil row, 32-bit offsets along x
    int xo[NS];
#pragma unroll
    for (int sx = 0; sx < NS; ++sx) xo[sx] = 3 * (idx[0][sx] - g.lo[0]);
    // stencil rows r = (sz, sy) in order, G rows at a time: a group's node
    // triples are all loaded before any is summed (G ns gathers in flight per
    // wavefront rather than one), and the scheduling barriers keep the
    // compiler from moving loads back beside their uses or hoisting the next
    // group's loads over this group's sums (register use stays bounded)
    constexpr int G = F64 ? (M6 ? 1 : 2) : (M6 ? 2 : 4);
#pragma unroll
    for (int r0 = 0; r0 < NS * NS; r0 += G) {
      T mv[G][NS][3];
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const int sy = (r0 + k) % NS, sz = (r0 + k) / NS;
        const int64_t row = (int64_t)g.E[0] * ((idx[1][sy] - g.lo[1]) + (int64_t)g.E[1] * (idx[2][sz] - g.lo[2]));
        const T *mrow = mesh + 3 * row;
#pragma unroll
        for (int sx = 0; sx < NS; ++sx)
#pragma unroll
          for (int c = 0; c < 3; ++c) mv[k][sx][c] = mrow[xo[sx] + c];
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const int sy = (r0 + k) % NS, sz = (r0 + k) / NS;
        // FP32 form: the FP32 weights' product in FP64, y and z once per row
        const double wyz = F64 ? 0.0 : (double)w[1][sy] * (double)w[2][sz];
#pragma unroll
        for (int sx = 0; sx < NS; ++sx) {
          const T *m = mv[k][sx];
          if constexpr (F64) {
            const double wt = __dmul_rn(__dmul_rn(w[0][sx], w[1][sy]), w[2][sz]);
            const double t0 = __dmul_rn(wt, m[0]), t1 = __dmul_rn(wt, m[1]), t2 = __dmul_rn(wt, m[2]);
            if (sz == 0 && sy == 0 && sx == 0) {   // the first term starts the sums
              a0 = t0;
              a1 = t1;
              a2 = t2;
            } else {
              a0 = __dadd_rn(a0, t0);
              a1 = __dadd_rn(a1, t1);
              a2 = __dadd_rn(a2, t2);
            }
          } else {
            // weight times FP32 node value fused into the FP64 sum
            const double wt = (double)w[0][sx] * wyz;
            a0 = __fma_rn(wt, (double)m[0], a0);
            a1 = __fma_rn(wt, (double)m[1], a1);
            a2 = __fma_rn(wt, (double)m[2], a2);
          }
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    out[3 * q] = a0;
    out[3 * q + 1] = a1;
    out[3 * q + 2] = a2;
  }
}

static __global__ void k_cell_index(PGeom g, const double *pos, int64_t np, int *out) {
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
// vel / dvort are the fields copied (NULL: zero), per the vel / dvort flags.
// A workgroup takes 256 consecutive cells at a time: their particles are one
// contiguous range of the outputs, staged in LDS and then written with
// consecutive threads on consecutive words.  Launched with 256 threads.
constexpr int kRemeshTile = 256;
template <typename T>
__global__ void __launch_bounds__(kRemeshTile) k_remesh_create(PGeom g, const int *flag, const int *offs,
                                                               const T *vort, const T *dvort, const T *vel,
                                                               double *pos, double *pvort, double *pvel,
                                                               double *pdvort, double *vel_rk, double *dvort_rk,
                                                               int64_t cap) {
  __shared__ double stage[4][3 * kRemeshTile];   // pos, vort, vel, dvort of the tile's particles
  const int64_t nn = (int64_t)g.E[0] * g.E[1] * g.E[2];
  const int64_t ntiles = (nn + kRemeshTile - 1) / kRemeshTile;
  const int tid = threadIdx.x;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t c0 = tile * kRemeshTile, idx = c0 + tid;
    const int64_t p0 = offs[c0], p1 = offs[min(c0 + kRemeshTile, nn)];
    if (idx < nn && flag[idx]) {
      const int s = (int)(offs[idx] - p0);
      int ii[3];
      ii[0] = (int)(idx % g.E[0]) + g.lo[0];
      ii[1] = (int)((idx / g.E[0]) % g.E[1]) + g.lo[1];
      ii[2] = (int)(idx / ((int64_t)g.E[0] * g.E[1])) + g.lo[2];
#pragma unroll
      for (int d = 0; d < 3; ++d)
        stage[0][3 * s + d] = __dadd_rn(g.xmin[d], __dmul_rn(__dadd_rn((double)(ii[d] - 1), 0.5), g.dx[d]));
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        stage[1][3 * s + c] = (double)vort[3 * idx + c];
        stage[2][3 * s + c] = vel ? (double)vel[3 * idx + c] : 0.0;
        stage[3][3 * s + c] = dvort ? (double)dvort[3 * idx + c] : 0.0;
      }
    }
    __syncthreads();
    // words [3 p0, 3 p1) of every output, past the capacity dropped (the
    // caller reports an over-capacity count)
    const int64_t w0 = 3 * p0, w1 = 3 * (p1 < cap ? p1 : (p0 < cap ? cap : p0));
    for (int64_t wd = w0 + tid; wd < w1; wd += kRemeshTile) {
      const int k = (int)(wd - w0);
      pos[wd] = stage[0][k];
      pvort[wd] = stage[1][k];
      if (pvel) pvel[wd] = stage[2][k];
      if (pdvort) pdvort[wd] = stage[3][k];
      if (vel_rk) vel_rk[wd] = 0.0;
      if (dvort_rk) dvort_rk[wd] = 0.0;
    }
    __syncthreads();   // the stage is free for the next tile
  }
}

// ------------------------------------------------------------ advance
struct RkCoef {
  double a1, b1, c1, a2, b2, c2;
};

// the patch box [lo, hi) per axis
struct PBox {
  double lo[3], hi[3];
};

// the RK stage update of every particle word; with `outside` given, also
// flags (1) a particle that the update leaves outside the box on any axis
// dvort null: a zero right-hand side (dvort = dvort_rk = +0), dvort_rk is
// neither read nor written and vort takes the increment c2 dr that the full
// form rounds for those zeros, dr = a2 (+0) + (b2 dt) (+0)
static __global__ void k_rk_update(RkCoef k, double dt, NCount npc, double *pos, const double *vel, double *vort,
                                   const double *dvort, double *vel_rk, double *dvort_rk, PBox box, int *outside) {
  const int64_t np = npc.get();
  const bool rhs = dvort != nullptr;
  const double dinc = __dmul_rn(k.c2, __dadd_rn(__dmul_rn(k.a2, 0.0), __dmul_rn(__dmul_rn(k.b2, dt), 0.0)));
  // axis of word i, advanced with the grid stride (no 64-bit modulo)
  const int i0 = blockIdx.x * blockDim.x + threadIdx.x, s3 = (int)(((int64_t)gridDim.x * blockDim.x) % 3);
  int d = i0 % 3;
  for (int64_t i = i0; i < 3 * np; i += (int64_t)gridDim.x * blockDim.x, d = (d + s3) % 3) {
    double vr = __dadd_rn(__dmul_rn(k.a1, vel_rk[i]), __dmul_rn(__dmul_rn(k.b1, dt), vel[i]));
    vel_rk[i] = vr;
    const double x = __dadd_rn(pos[i], __dmul_rn(k.c1, vr));
    pos[i] = x;
    if (rhs) {
      double dr = __dadd_rn(__dmul_rn(k.a2, dvort_rk[i]), __dmul_rn(__dmul_rn(k.b2, dt), dvort[i]));
      dvort_rk[i] = dr;
      vort[i] = __dadd_rn(vort[i], __dmul_rn(k.c2, dr));
    } else {
      vort[i] = __dadd_rn(vort[i], dinc);
    }
    if (outside) {
      const bool o = x < box.lo[d] || x >= box.hi[d];
      const uint64_t m = __ballot(o);
      if (m && (threadIdx.x & 63) == __ffsll((unsigned long long)m) - 1) atomicOr(outside, 1);
    }
  }
}

// migration class along axis d: 0 stay, 1 west (wrap), 2 east (wrap), 3 drop;
// nothing to do when the stage's update left every particle inside
static __global__ void k_classify(NCount npc, const int *outside, const double *pos, int d, double lo, double hi,
                                  int per, int *cls, int *counts) {
  if (!*outside) return;
  const int64_t np = npc.get();
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    double x = pos[3 * q + d];
    int c = x < lo ? 1 : (x >= hi ? 2 : 0);
    if (c != 0 && !per) c = 3;
    cls[q] = c;
    if (c != 0) atomicAdd(&counts[c], 1);
  }
}

}  // namespace pmgpu
