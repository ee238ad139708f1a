// Host side of the particle kernels: workspace, scans, C ABI.
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/pmgpu.h"
#include "particles_kernels.h"


namespace pmgpu {

struct Scratch {
  void *p = nullptr;
  size_t cap = 0;
  int ensure(size_t bytes) {
    if (bytes <= cap) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    PM_HIP(hipMalloc(&p, bytes));
    cap = bytes;
    return 0;
  }
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
  Scratch() = default;
  Scratch(const Scratch &) = delete;
  Scratch &operator=(const Scratch &) = delete;
  template <typename T> T *as() { return reinterpret_cast<T *>(p); }
};

// Exclusive scan of n ints (in -> out, not aliased): out[0] = 0 and out[i +
// 1] = in[0] + ... + in[i], so out[n] is the total.  Reduce, then scan, over
// tiles of kIScanTile ints: the tile sums, one block turning them into tile
// offsets, then each tile scanned through LDS (each thread a run of
// kIScanItems) from its offset.  No memset, so a step may record it.
constexpr int kIScanThreads = 256, kIScanItems = 16, kIScanTile = kIScanThreads * kIScanItems;
constexpr int kIScanOffThreads = 1024;
__device__ __forceinline__ int block_excl_int(int v, int &total) {
  __shared__ int ws[kIScanThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[wv] = x;
  __syncthreads();
  int base = 0;
  total = 0;
#pragma unroll
  for (int k = 0; k < kIScanThreads / 64; ++k) {
    if (k < wv) base += ws[k];
    total += ws[k];
  }
  __syncthreads();   // ws is free for the next call
  return base + x - v;
}
__global__ void __launch_bounds__(kIScanThreads) k_iscan_tiles(const int *in, int64_t n, int *tsum) {
  const int64_t b = (int64_t)blockIdx.x * kIScanTile;
  int s = 0;
#pragma unroll
  for (int j = 0; j < kIScanItems; ++j) {
    const int64_t i = b + j * kIScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  int tot;
  (void)block_excl_int(s, tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}
// tile sums -> exclusive tile offsets in place: each wavefront a contiguous
// share, summed, then walked 64 tiles at a time (the next 64 prefetched)
__global__ void __launch_bounds__(kIScanOffThreads) k_iscan_offsets(int *tsum, int ntiles) {
  constexpr int NW = kIScanOffThreads / 64;
  __shared__ int wtot[NW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int per = (ntiles + NW - 1) / NW;
  const int t0 = min(wv * per, ntiles), t1 = min(t0 + per, ntiles);
  int s = 0;
  for (int t = t0 + lane; t < t1; t += 64) s += tsum[t];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) wtot[wv] = s;
  __syncthreads();
  int run = 0;
  for (int k = 0; k < wv; ++k) run += wtot[k];
  int vn = t0 + lane < t1 ? tsum[t0 + lane] : 0;
  for (int b = t0; b < t1; b += 64) {
    const int t = b + lane;
    const int v = vn;
    vn = t + 64 < t1 ? tsum[t + 64] : 0;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(x, o);
      if (lane >= o) x += y;
    }
    if (t < t1) tsum[t] = run + x - v;
    run += __shfl(x, 63);
  }
}
__global__ void __launch_bounds__(kIScanThreads) k_iscan_apply(const int *in, int64_t n, const int *toff, int *out) {
  __shared__ int sh[kIScanTile + kIScanTile / 32];
  const int64_t b = (int64_t)blockIdx.x * kIScanTile;
  if (blockIdx.x == 0 && threadIdx.x == 0) out[0] = 0;
#pragma unroll
  for (int j = 0; j < kIScanItems; ++j) {
    const int i = j * kIScanThreads + threadIdx.x;
    sh[i + i / 32] = b + i < n ? in[b + i] : 0;
  }
  __syncthreads();
  int v[kIScanItems], sum = 0;
#pragma unroll
  for (int k = 0; k < kIScanItems; ++k) {
    const int i = threadIdx.x * kIScanItems + k;
    v[k] = sh[i + i / 32];
    sum += v[k];
  }
  int tot;
  int run = block_excl_int(sum, tot) + toff[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kIScanItems; ++k) {
    const int i = threadIdx.x * kIScanItems + k;
    run += v[k];
    sh[i + i / 32] = run;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kIScanItems; ++j) {
    const int i = j * kIScanThreads + threadIdx.x;
    if (b + i < n) out[1 + b + i] = sh[i + i / 32];
  }
}
static int64_t iscan_tiles(int64_t n) { return (n + kIScanTile - 1) / kIScanTile; }
// tsum: iscan_tiles(n) ints of scratch
static int iscan_launch(const int *in, int *out, int64_t n, int *tsum, hipStream_t st) {
  const int64_t nt = iscan_tiles(n);
  PM_CHECK(nt < (int64_t)1 << 31, "pmgpu: scan too large");
  if (nt == 0) {
    PM_HIP(hipMemsetAsync(out, 0, sizeof(int), st));
    return 0;
  }
  hipLaunchKernelGGL(k_iscan_tiles, dim3((unsigned)nt), dim3(kIScanThreads), 0, st, in, n, tsum);
  hipLaunchKernelGGL(k_iscan_offsets, dim3(1), dim3(kIScanOffThreads), 0, st, tsum, (int)nt);
  hipLaunchKernelGGL(k_iscan_apply, dim3((unsigned)nt), dim3(kIScanThreads), 0, st, in, n, (const int *)tsum, out);
  PM_HIP(hipGetLastError());
  return 0;
}
static int scan_exclusive(const int *in, int *out, int64_t n, Scratch &tmp, hipStream_t st) {
  PM_CHECK(in != out, "pmgpu: internal scan aliasing");
  if (tmp.ensure(sizeof(int) * ((size_t)iscan_tiles(n) + 1))) return -1;
  return iscan_launch(in, out, n, tmp.as<int>(), st);
}

// Counts of west movers, east movers and drops (the migration classes 1, 2,
// 3).  The migrations scan them in two levels over tiles of kPartTile list
// entries: class totals per tile, one block scanning the tile totals into
// tile offsets, then each tile's own block scans in index order.
struct Mig3 {
  int w, e, d;
};
constexpr int kPartThreads = 256, kPartItems = 8, kPartTile = kPartThreads * kPartItems;
__device__ __forceinline__ Mig3 m3_of(int c) { return Mig3{c == 1 ? 1 : 0, c == 2 ? 1 : 0, c == 3 ? 1 : 0}; }
__device__ __forceinline__ Mig3 m3_add(const Mig3 &a, const Mig3 &b) { return Mig3{a.w + b.w, a.e + b.e, a.d + b.d}; }
// exclusive scan of one Mig3 per thread over a kPartThreads-thread block;
// every thread of the block calls it
__device__ Mig3 block_scan3(Mig3 v, Mig3 &total) {
  __shared__ Mig3 wsum[kPartThreads / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  Mig3 x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const Mig3 y{__shfl_up(x.w, o), __shfl_up(x.e, o), __shfl_up(x.d, o)};
    if (lane >= o) x = m3_add(x, y);
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  Mig3 base{0, 0, 0};
  total = Mig3{0, 0, 0};
#pragma unroll
  for (int k = 0; k < kPartThreads / 64; ++k) {
    if (k < wv) base = m3_add(base, wsum[k]);
    total = m3_add(total, wsum[k]);
  }
  __syncthreads();   // wsum is free for the next call
  return Mig3{base.w + x.w - v.w, base.e + x.e - v.e, base.d + x.d - v.d};
}
__device__ __forceinline__ bool part_idle(const int *counts) { return counts[1] + counts[2] + counts[3] == 0; }
// tile sums -> exclusive tile offsets, in place (one block of
// kTileScanThreads): each wavefront owns a contiguous share of the tiles,
// sums it, and after the totals are combined walks it 64 tiles at a time
// with a wavefront scan, the next 64 loaded while the current are scanned
// (one block barrier in all)
constexpr int kTileScanThreads = 1024;
__global__ void __launch_bounds__(kTileScanThreads) k_part_tile_scan(const int *counts, Mig3 *tsum, int ntiles) {
  if (part_idle(counts)) return;
  constexpr int NW = kTileScanThreads / 64;
  __shared__ Mig3 wtot[NW];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int per = (ntiles + NW - 1) / NW;
  const int t0 = min(wv * per, ntiles), t1 = min(t0 + per, ntiles);
  const Mig3 zero{0, 0, 0};
  Mig3 s = zero;
  for (int t = t0 + lane; t < t1; t += 64) s = m3_add(s, tsum[t]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s = m3_add(s, Mig3{__shfl_xor(s.w, o), __shfl_xor(s.e, o), __shfl_xor(s.d, o)});
  if (lane == 0) wtot[wv] = s;
  __syncthreads();
  Mig3 run = zero;
  for (int k = 0; k < wv; ++k) run = m3_add(run, wtot[k]);
  Mig3 vn = t0 + lane < t1 ? tsum[t0 + lane] : zero;
  for (int b = t0; b < t1; b += 64) {
    const int t = b + lane;
    const Mig3 v = vn;
    vn = t + 64 < t1 ? tsum[t + 64] : zero;
    Mig3 x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const Mig3 y{__shfl_up(x.w, o), __shfl_up(x.e, o), __shfl_up(x.d, o)};
      if (lane >= o) x = m3_add(x, y);
    }
    if (t < t1) tsum[t] = Mig3{run.w + x.w - v.w, run.e + x.e - v.e, run.d + x.d - v.d};
    run = m3_add(run, Mig3{__shfl(x.w, 63), __shfl(x.e, 63), __shfl(x.d, 63)});
  }
}

// ------------------------------------------------------------ geometry
static int make_pgeom(const pmgpu_geom *g, PGeom &pg) {
  PM_CHECK(g, "pmgpu: null geometry");
  PM_CHECK(g->interp_order == 3 || g->interp_order == 4,
           "pmgpu: Interpolation scheme/order unknown (interp_particle_mesh.f:1118)");
  pg.m6 = g->interp_order == 4;
  pg.ns = pg.m6 ? 6 : 4;
  pg.off = pg.m6 ? 2 : 1;
  for (int d = 0; d < 3; ++d) {
    PM_CHECK(g->ncell[d] > 0 && g->dx[d] > 0, "pmgpu: bad geometry");
    pg.ncell[d] = g->ncell[d];
    pg.bc[d] = g->bound_cond[d];
    pg.ngl[d] = g->bound_cond[d] == 0 ? g->nghost : (pg.m6 ? 3 : 2);
    pg.E[d] = g->ncell[d] + 2 * pg.ngl[d];
    pg.lo[d] = 1 - pg.ngl[d];
    pg.hi[d] = g->ncell[d] + pg.ngl[d];
    pg.xmin[d] = g->xmin[d];
    pg.xmax[d] = g->xmax[d];
    pg.dx[d] = g->dx[d];
    pg.c1dx[d] = 1.0 / g->dx[d];
    pg.blo[d] = pg.lo[d] + pg.off - pg.ns + 1;
    pg.nb[d] = (pg.hi[d] + pg.off) - pg.blo[d] + 1;
  }
  return 0;
}

static int64_t mesh_cells(const PGeom &g) { return (int64_t)g.E[0] * g.E[1] * g.E[2]; }
static dim3 grid_for(int64_t n, int threads = 256) {
  int64_t b = (n + threads - 1) / threads;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return dim3((unsigned)b);
}

// Workspace shared by the particle entry points of one process.
struct PartWS {
  Scratch key, ids, sort, count, start, scan, flag, offs, cls, migbuf, rec;
};
// Allocated on first use and released by pmgpu_finalize (never from a static
// destructor, which may run after the HIP runtime has shut down).
static PartWS *g_ws = nullptr;
static PartWS &ws() {
  if (!g_ws) g_ws = new PartWS;
  return *g_ws;
}

void particles_release() {
  delete g_ws;
  g_ws = nullptr;
}

// ---------------------------------------------------------------- P2M
// prec: PMGPU_FP64 (double mesh, reference arithmetic) or PMGPU_FP32 (float
// mesh, FP32 weights, per-row partial sums Kahan-added across the stencil
// rows); particles are FP64.
// Order each bin's particles by id (one thread per bin), stage their records
// (one thread per sorted particle), then gather them line by line of 64 nodes
// along x (one wavefront per line).
template <typename T, int NS>
static void p2m_stage_gather(const PGeom &g, const double *xpart, const double *part, int64_t np, int *ids,
                             PartWS &w, int nbins, T *mesh, int accumulate, hipStream_t st) {
  // the records are planar with np + 1 records per field (the allocation
  // holds kSlots fields of np + 1)
  const int64_t rs = np + 1;
  if (np > 0) {
    hipLaunchKernelGGL(k_p2m_binsort, grid_for(nbins), dim3(256), 0, st, w.start.as<int>(), nbins, ids);
    hipLaunchKernelGGL((k_p2m_stage<T, NS>), grid_for(np), dim3(256), 0, st, g, xpart, part, ids,
                       w.start.as<int>(), nbins, w.rec.as<T>(), rs);
  }
  const dim3 grid((g.E[0] + 63) / 64, g.E[1], g.E[2]);
  hipLaunchKernelGGL((k_p2m_rows<T, NS>), grid, dim3(64), 0, st, g, w.start.as<int>(), w.rec.as<T>(), rs, mesh,
                     accumulate);
}

int p2m_device(const pmgpu_geom *geom, int prec, const double *xpart, const double *part, int64_t np, void *mesh,
               int clear, hipStream_t st) {
  PGeom g;
  if (make_pgeom(geom, g)) return -1;
  PM_CHECK(prec == PMGPU_FP64 || prec == PMGPU_FP32, "pmgpu: precision must be PMGPU_FP64 or PMGPU_FP32");
  PM_CHECK(mesh && (np == 0 || (xpart && part)), "pmgpu_interp_particle_mesh: null argument");
  for (int d = 0; d < 3; ++d)
    PM_CHECK(g.bc[d] == 0 || g.ncell[d] >= 2 * g.ngl[d], "pmgpu: periodic axis shorter than two ghost layers");
  PartWS &w = ws();
  const int nbins = g.nb[0] * g.nb[1] * g.nb[2];
  PM_CHECK(np >= 0 && np < (int64_t)1 << 31, "pmgpu: particle count out of range");
  PM_CHECK((int64_t)nbins < ((int64_t)1 << 30), "pmgpu: mesh too large for deposition");
  // counting sort by bin: count and rank (keys), scan, scatter; the stage
  // kernel restores id order inside each bin
  if (w.key.ensure(sizeof(int) * ((size_t)np + 1)) || w.ids.ensure(sizeof(int) * 2 * ((size_t)np + 1)) ||
      w.count.ensure(sizeof(int) * (nbins + 1)) || w.start.ensure(sizeof(int) * (nbins + 1)))
    return -1;
  int *key = w.key.as<int>();
  int *rank = w.ids.as<int>(), *id_out = rank + np + 1;
  PM_HIP(hipMemsetAsync(w.count.p, 0, sizeof(int) * (nbins + 1), st));
  if (np > 0) {
    hipLaunchKernelGGL(k_p2m_keys, grid_for(np), dim3(256), 0, st, g, xpart, np, nbins, key, rank,
                       w.count.as<int>());
    PM_HIP(hipGetLastError());
  }
  if (scan_exclusive(w.count.as<int>(), w.start.as<int>(), nbins, w.scan, st)) return -1;
  if (np > 0) {
    hipLaunchKernelGGL(k_p2m_scatter, grid_for(np), dim3(256), 0, st, np, nbins, key, rank, w.start.as<int>(),
                       id_out);
    PM_HIP(hipGetLastError());
  }
  const int accumulate = clear ? 0 : 1;
  const size_t rec_bytes = (prec == PMGPU_FP64 ? sizeof(double) : sizeof(float)) * PRec<double, 6>::kSlots;
  if (w.rec.ensure(rec_bytes * ((size_t)np + 1))) return -1;
  if (prec == PMGPU_FP64) {
    if (g.ns == 4) p2m_stage_gather<double, 4>(g, xpart, part, np, id_out, w, nbins, (double *)mesh, accumulate, st);
    else p2m_stage_gather<double, 6>(g, xpart, part, np, id_out, w, nbins, (double *)mesh, accumulate, st);
  } else {
    if (g.ns == 4) p2m_stage_gather<float, 4>(g, xpart, part, np, id_out, w, nbins, (float *)mesh, accumulate, st);
    else p2m_stage_gather<float, 6>(g, xpart, part, np, id_out, w, nbins, (float *)mesh, accumulate, st);
  }
  PM_HIP(hipGetLastError());
  for (int d = 0; d < 3; ++d)
    if (g.bc[d] != 0) {
      const dim3 grid = grid_for(2LL * g.ngl[d] * 3 * g.E[(d + 1) % 3] * g.E[(d + 2) % 3]);
      if (prec == PMGPU_FP64)
        hipLaunchKernelGGL(k_halo_add<double>, grid, dim3(256), 0, st, (double *)mesh, g, d);
      else
        hipLaunchKernelGGL(k_halo_add<float>, grid, dim3(256), 0, st, (float *)mesh, g, d);
      PM_HIP(hipGetLastError());
    }
  return 0;
}

template <typename T> __global__ void k_zero_ghosts(PGeom g, T *mesh) {
  const int64_t nn = (int64_t)g.E[0] * g.E[1] * g.E[2];
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nn;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int i = (int)(idx % g.E[0]) + g.lo[0];
    int j = (int)((idx / g.E[0]) % g.E[1]) + g.lo[1];
    int k = (int)(idx / ((int64_t)g.E[0] * g.E[1])) + g.lo[2];
    bool interior = i >= 1 && i <= g.ncell[0] && j >= 1 && j <= g.ncell[1] && k >= 1 && k <= g.ncell[2];
    if (!interior) mesh[3 * idx] = mesh[3 * idx + 1] = mesh[3 * idx + 2] = T(0);
  }
}

int p2m_device_full(const pmgpu_geom *geom, int prec, const double *xpart, const double *part, int64_t np,
                    void *mesh, int clear, int clear_ghost, hipStream_t st) {
  if (!clear && clear_ghost) {
    PGeom g;
    if (make_pgeom(geom, g)) return -1;
    PM_CHECK(mesh, "pmgpu_interp_particle_mesh: null argument");
    if (prec == PMGPU_FP32)
      hipLaunchKernelGGL(k_zero_ghosts<float>, grid_for(mesh_cells(g)), dim3(256), 0, st, g, (float *)mesh);
    else
      hipLaunchKernelGGL(k_zero_ghosts<double>, grid_for(mesh_cells(g)), dim3(256), 0, st, g, (double *)mesh);
    PM_HIP(hipGetLastError());
  }
  return p2m_device(geom, prec, xpart, part, np, mesh, clear, st);
}

// ---------------------------------------------------------------- M2P
int m2p_device(const pmgpu_geom *geom, int prec, const void *mesh, const double *pos, int64_t np, double *out,
               hipStream_t st) {
  PGeom g;
  if (make_pgeom(geom, g)) return -1;
  PM_CHECK(prec == PMGPU_FP64 || prec == PMGPU_FP32, "pmgpu: precision must be PMGPU_FP64 or PMGPU_FP32");
  PM_CHECK(mesh && (np == 0 || (pos && out)), "pmgpu_interp_mesh_particle: null argument");
  if (np > 0) {
    if (prec == PMGPU_FP64) {
      if (g.ns == 4)
        hipLaunchKernelGGL((k_m2p<double, 4>), grid_for(np), dim3(256), 0, st, g, (const double *)mesh, pos, np, out);
      else
        hipLaunchKernelGGL((k_m2p<double, 6>), grid_for(np), dim3(256), 0, st, g, (const double *)mesh, pos, np, out);
    } else {
      if (g.ns == 4)
        hipLaunchKernelGGL((k_m2p<float, 4>), grid_for(np), dim3(256), 0, st, g, (const float *)mesh, pos, np, out);
      else
        hipLaunchKernelGGL((k_m2p<float, 6>), grid_for(np), dim3(256), 0, st, g, (const float *)mesh, pos, np, out);
    }
  }
  PM_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------- remesh
template <typename T>
static void remesh_flags_launch(const PGeom &g, int64_t nn, const void *vort_m, const void *dvort_m,
                                const uint8_t *mask, double min_vort, double min_dvort, int incl_ghost, int *flag,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_remesh_flags<T>, grid_for(nn), dim3(256), 0, st, g, (const T *)vort_m, (const T *)dvort_m,
                     mask, min_vort, min_dvort, incl_ghost, flag);
}
template <typename T>
static void remesh_create_launch(const PGeom &g, int64_t nn, const int *flag, const int *offs, const void *vort_m,
                                 const void *dvort_m, const void *vel_m, double *pos, double *vort, double *vel,
                                 double *dvort, double *vel_rk, double *dvort_rk, hipStream_t st) {
  hipLaunchKernelGGL(k_remesh_create<T>, grid_for(nn), dim3(256), 0, st, g, flag, offs, (const T *)vort_m,
                     (const T *)dvort_m, (const T *)vel_m, pos, vort, vel, dvort, vel_rk, dvort_rk, INT64_MAX);
}

int remesh_device(const pmgpu_geom *geom, int prec, const void *vort_m, const void *dvort_m, const void *vel_m,
                  const uint8_t *mask, double min_vort, double min_dvort, int copy_dvort, int copy_vel,
                  int incl_ghost, double *pos, double *vort, double *vel, double *dvort, double *vel_rk,
                  double *dvort_rk, int64_t cap, int64_t *np_out, hipStream_t st) {
  PGeom g;
  if (make_pgeom(geom, g)) return -1;
  PM_CHECK(prec == PMGPU_FP64 || prec == PMGPU_FP32, "pmgpu: precision must be PMGPU_FP64 or PMGPU_FP32");
  PartWS &w = ws();
  const int64_t nn = mesh_cells(g);
  PM_CHECK(nn < (int64_t)1 << 31, "pmgpu: mesh too large for remeshing");
  if (w.flag.ensure(sizeof(int) * (nn + 1)) || w.offs.ensure(sizeof(int) * (nn + 1))) return -1;
  PM_CHECK(vort_m, "pmgpu_remesh: null vorticity mesh");
  PM_CHECK(!copy_vel || vel_m, "pmgpu_remesh: vel requested without a velocity mesh");
  if (prec == PMGPU_FP64)
    remesh_flags_launch<double>(g, nn, vort_m, dvort_m, mask, min_vort, min_dvort, incl_ghost, w.flag.as<int>(), st);
  else
    remesh_flags_launch<float>(g, nn, vort_m, dvort_m, mask, min_vort, min_dvort, incl_ghost, w.flag.as<int>(), st);
  PM_HIP(hipGetLastError());
  if (scan_exclusive(w.flag.as<int>(), w.offs.as<int>(), nn, w.scan, st)) return -1;
  int total = 0;
  PM_HIP(hipMemcpyAsync(&total, w.offs.as<int>() + nn, sizeof(int), hipMemcpyDeviceToHost, st));
  PM_HIP(hipStreamSynchronize(st));
  *np_out = total;
  if (pos == nullptr || total > cap) return 0;  // count only
  PM_CHECK(vort, "pmgpu_remesh: null particle vorticity");
  const void *dv = copy_dvort ? dvort_m : nullptr, *vl = copy_vel ? vel_m : nullptr;
  if (prec == PMGPU_FP64)
    remesh_create_launch<double>(g, nn, w.flag.as<int>(), w.offs.as<int>(), vort_m, dv, vl, pos, vort, vel, dvort,
                                 vel_rk, dvort_rk, st);
  else
    remesh_create_launch<float>(g, nn, w.flag.as<int>(), w.offs.as<int>(), vort_m, dv, vl, pos, vort, vel, dvort,
                                vel_rk, dvort_rk, st);
  PM_HIP(hipGetLastError());
  return 0;
}

// ------------------------------------------------------------ advance
// The map of the eager path works on an index list: each pass of each axis
// partitions the list (west movers, then east movers, then stayers, each in
// list order; drops removed) and wraps the axis coordinate it carries along
// (with the columns of earlier axes that wrapped), so that the particle
// arrays are permuted once, after the last axis.  The final order and every
// wrapped coordinate are those of permuting the arrays at every pass, which
// is what the reference's map does.

// migration class of each listed particle (0 stay, 1 west, 2 east, 3 drop),
// the class totals of each tile of kPartTile entries (block b: entries
// [b kPartTile, (b + 1) kPartTile)) and the class counts; with pos given,
// the coordinate column cd is first gathered from the positions (coordinate
// d of the listed particles, the identity list when perm is null)
__global__ void __launch_bounds__(kPartThreads) k_list_classify(int64_t n, const int *perm, const double *pos, int d,
                                                                double *cd, double lo, double hi, int per, int *cls,
                                                                Mig3 *tsum, int *counts) {
  Mig3 s{0, 0, 0};
#pragma unroll
  for (int j = 0; j < kPartItems; ++j) {
    const int64_t i = (int64_t)blockIdx.x * kPartTile + j * kPartThreads + threadIdx.x;
    if (i >= n) break;
    double x;
    if (pos) {
      x = pos[3 * (int64_t)(perm ? perm[i] : (int)i) + d];
      cd[i] = x;
    } else {
      x = cd[i];
    }
    int c = x < lo ? 1 : (x >= hi ? 2 : 0);
    if (c != 0 && !per) c = 3;
    cls[i] = c;
    s = m3_add(s, m3_of(c));
  }
  Mig3 tot;
  (void)block_scan3(s, tot);
  if (threadIdx.x == 0) {
    tsum[blockIdx.x] = tot;
    if (tot.w) atomicAdd(&counts[1], tot.w);
    if (tot.e) atomicAdd(&counts[2], tot.e);
    if (tot.d) atomicAdd(&counts[3], tot.d);
  }
}
// coordinate columns moved with the list: the current axis' column (index 0)
// and those of earlier axes whose particles wrapped
struct ListCols {
  const double *src[3];
  double *dst[3];
  int n;
};
// one partition pass of the list, block b on tile b from its offset toff[b];
// movers' coordinates wrapped by one period, and those still outside
// [lo, hi) after the wrap counted in *still (the next pass's movers)
__global__ void __launch_bounds__(kPartThreads) k_list_partition(int64_t n, const int *cls, const Mig3 *toff, int n1,
                                                                 int n2, double len, double lo, double hi,
                                                                 const int *perm, int *perm2, ListCols cols,
                                                                 int *still) {
  Mig3 run = toff[blockIdx.x];
  for (int j = 0; j < kPartItems; ++j) {
    const int64_t i = (int64_t)blockIdx.x * kPartTile + j * kPartThreads + threadIdx.x;
    const int c = i < n ? cls[i] : 0;
    Mig3 tot;
    const Mig3 ex = block_scan3(m3_of(c), tot);
    const Mig3 o = m3_add(run, ex);
    run = m3_add(run, tot);
    if (i >= n || c == 3) continue;
    const int64_t to = c == 1 ? o.w : (c == 2 ? n1 + o.e : n1 + n2 + (i - o.w - o.e - o.d));
    double x = cols.src[0][i];
    if (c == 1) x = x + len;
    if (c == 2) x = x - len;
    if (c != 0 && (x < lo || x >= hi)) atomicAdd(still, 1);
    perm2[to] = perm ? perm[i] : (int)i;
    cols.dst[0][to] = x;
    for (int k = 1; k < cols.n; ++k) cols.dst[k][to] = cols.src[k][i];
  }
}
struct Arr4 {
  double *a[4];
};
// n words of each of narr arrays, src -> dst, two words per thread step
// (every pointer 16-byte aligned)
__global__ void k_copy_back4(int64_t n, const Arr4 src, Arr4 dst, int narr) {
  const int64_t n2 = n / 2;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n2; i += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < narr; ++a)
      reinterpret_cast<double2 *>(dst.a[a])[i] = reinterpret_cast<const double2 *>(src.a[a])[i];
  if (n % 2 && blockIdx.x == 0 && threadIdx.x == 0)
    for (int a = 0; a < narr; ++a) dst.a[a][n - 1] = src.a[a][n - 1];
}
// row i of each carried array from row perm[i]; the position rows take the
// wrapped coordinates from the columns of the axes that wrapped (col[d] not
// null)
__global__ void k_list_gather(int64_t n, const int *perm, Arr4 src, Arr4 dst, int narr, ListCols col) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 3 * n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t q = i / 3, c = i - 3 * q;
    const int64_t from = 3 * (int64_t)perm[q] + c;
    const double *cc = col.src[c];
    dst.a[0][i] = cc ? cc[q] : src.a[0][from];
    for (int a = 1; a < narr; ++a) dst.a[a][i] = src.a[a][from];
  }
}

static RkCoef rk_coef(int order, int stage, int euler_vort) {
  double A1[3] = {0, 0, 0}, B1[3] = {1, 0, 0}, C1[3] = {1, 0, 0};
  double A2[3] = {0, 0, 0}, B2[3] = {1, 0, 0}, C2[3] = {1, 0, 0};
  if (order == 2) {
    A1[1] = -0.5; B1[1] = 1.0; C1[0] = 0.5; C1[1] = 1.0;
    if (euler_vort) { A2[1] = 1.0; B2[1] = 0.0; C2[0] = 0.5; C2[1] = 0.5; }
    else { A2[1] = -0.5; B2[1] = 1.0; C2[0] = 0.5; C2[1] = 1.0; }
  } else if (order == 3) {
    A1[1] = -5.0 / 9.0; A1[2] = -153.0 / 128.0; B1[1] = B1[2] = 1.0;
    C1[0] = 1.0 / 3.0; C1[1] = 15.0 / 16.0; C1[2] = 8.0 / 15.0;
    if (euler_vort) { A2[1] = A2[2] = 1.0; B2[1] = B2[2] = 0.0; C2[0] = C2[1] = C2[2] = 1.0 / 3.0; }
    else {
      A2[1] = -5.0 / 9.0; A2[2] = -153.0 / 128.0; B2[1] = B2[2] = 1.0;
      C2[0] = 1.0 / 3.0; C2[1] = 15.0 / 16.0; C2[2] = 8.0 / 15.0;
    }
  }
  int s = stage - 1;
  return RkCoef{A1[s], B1[s], C1[s], A2[s], B2[s], C2[s]};
}

// arrays: pos, vel, vort, dvort, vel_rk, dvort_rk (3, np) on the device; the
// survivors are compacted in place (through a scratch copy)
// spare (may be null): a second set of the six arrays of the same capacity;
// when the map reorders particles, the carried arrays' new contents go there
// instead of back into arr, and *in_spare is set to 1 (else 0)
int advance_device(const pmgpu_geom *geom, int rk_order, int stage, double dt, int euler_vort, double **arr,
                   int64_t np, int64_t *np_out, hipStream_t st, double **spare, int *in_spare) {
  if (in_spare) *in_spare = 0;
  PM_CHECK(rk_order >= 1 && rk_order <= 3, "pmgpu: time integration order must be 1, 2 or 3");
  PM_CHECK(stage >= 1 && stage <= rk_order, "pmgpu: RK stage out of range");
  PM_CHECK(np >= 0 && np < (int64_t)1 << 31, "pmgpu: particle count out of range");
  // dvort null: a zero right-hand side, dvort_rk not used (may be null too)
  const bool rhs = arr[3] != nullptr;
  PM_CHECK(np == 0 || (arr[0] && arr[1] && arr[2] && arr[4] && (!rhs || arr[5])),
           "pmgpu_particles_advance: null array (only dvort, and with it dvort_rk, may be null)");
  RkCoef k = rk_coef(rk_order, stage, euler_vort);
  PartWS &w = ws();
  int64_t n = np;
  if (n > 0) {
    // the update flags a particle it leaves outside the patch on any axis:
    // when there is none, no axis below moves or drops one (its
    // classification sees the same positions), so one read-back decides for
    // all three
    PBox box;
    for (int d = 0; d < 3; ++d) {
      box.lo[d] = geom->xmin[d];
      box.hi[d] = geom->xmax[d];
    }
    if (w.count.ensure(sizeof(int) * 8)) return -1;
    PM_HIP(hipMemsetAsync(w.count.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_rk_update, grid_for(3 * n), dim3(256), 0, st, k, dt, n, arr[0], arr[1], arr[2], arr[3],
                       arr[4], rhs ? arr[5] : nullptr, box, w.count.as<int>());
    PM_HIP(hipGetLastError());
    int any = 0;
    PM_HIP(hipMemcpyAsync(&any, w.count.p, sizeof(int), hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    if (any == 0) {
      *np_out = n;
      return 0;
    }
  }
  // the arrays the reference's map carries (pmlib_particles_advance.f:121-166):
  // pos and vort, and above first order vel_rk and dvort_rk (not with a zero
  // right-hand side: it stays zero); vel and dvort keep what the update left
  // in them
  const int carry[4] = {0, 2, 4, 5};
  const int narr = rk_order > 1 ? (rhs ? 4 : 3) : 2;
  const int64_t n0 = n;
  // list workspace: cls, two index lists (ints), the tile totals and
  // offsets (Mig3), two buffers of each axis' coordinate column and the
  // gathered rows (doubles)
  const size_t ni = ((size_t)n0 + 2) & ~(size_t)1;   // even: 16-byte aligned rows in migbuf
  const size_t nt0 = (size_t)((n0 + kPartTile - 1) / kPartTile) + 1;
  if (w.cls.ensure(sizeof(int) * 3 * ni + sizeof(Mig3) * nt0 + 16) || w.count.ensure(sizeof(int) * 8) ||
      w.migbuf.ensure(sizeof(double) * (6 * ni + 3 * ni * narr)))
    return -1;
  int *cls = w.cls.as<int>();
  int *plist[2] = {cls + ni, cls + 2 * ni};
  Mig3 *tsum = reinterpret_cast<Mig3 *>(cls + 3 * ni);
  double *colbuf[3][2];
  for (int d = 0; d < 3; ++d)
    for (int b = 0; b < 2; ++b) colbuf[d][b] = w.migbuf.as<double>() + (size_t)(2 * d + b) * ni;
  bool carried[3] = {false, false, false};   // axis d wrapped: its column travels with the list
  int cur = 0;
  bool listed = false;   // plist[cur] holds the list (else the identity)
  for (int d = 0; d < 3 && n > 0; ++d) {
    const double lo = geom->xmin[d], hi = geom->xmax[d], len = geom->xmax[d] - geom->xmin[d];
    const int per = geom->bound_cond[d] != 0;
    bool wrapped = false;
    for (int pass = 0; pass < 64; ++pass) {
      PM_HIP(hipMemsetAsync(w.count.p, 0, sizeof(int) * 5, st));
      // the first pass of an axis gathers its coordinate column
      const int ntl = (int)((n + kPartTile - 1) / kPartTile);
      hipLaunchKernelGGL(k_list_classify, dim3(ntl), dim3(kPartThreads), 0, st, n,
                         listed ? (const int *)plist[cur] : nullptr, pass == 0 ? (const double *)arr[0] : nullptr, d,
                         colbuf[d][cur], lo, hi, per, cls, tsum, w.count.as<int>());
      int counts[4];
      PM_HIP(hipMemcpyAsync(counts, w.count.p, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
      PM_HIP(hipStreamSynchronize(st));
      if (counts[1] == 0 && counts[2] == 0 && counts[3] == 0) break;  // nobody moves
      hipLaunchKernelGGL(k_part_tile_scan, dim3(1), dim3(kTileScanThreads), 0, st, (const int *)w.count.as<int>(), tsum,
                         ntl);
      ListCols cols{};
      cols.src[0] = colbuf[d][cur];
      cols.dst[0] = colbuf[d][1 - cur];
      cols.n = 1;
      for (int e = 0; e < d; ++e)
        if (carried[e]) {
          cols.src[cols.n] = colbuf[e][cur];
          cols.dst[cols.n++] = colbuf[e][1 - cur];
        }
      hipLaunchKernelGGL(k_list_partition, dim3(ntl), dim3(kPartThreads), 0, st, n, (const int *)cls, (const Mig3 *)tsum,
                         counts[1], counts[2], len, lo, hi, listed ? (const int *)plist[cur] : nullptr, plist[1 - cur],
                         cols, w.count.as<int>() + 4);
      PM_HIP(hipGetLastError());
      cur = 1 - cur;
      listed = true;
      wrapped = wrapped || counts[1] + counts[2] > 0;
      n -= counts[3];
      if (!per || counts[1] + counts[2] == 0) break;
      // a further pass only when a wrapped particle is still outside (it
      // would find nothing to move otherwise)
      int still = 0;
      PM_HIP(hipMemcpyAsync(&still, w.count.as<int>() + 4, sizeof(int), hipMemcpyDeviceToHost, st));
      PM_HIP(hipStreamSynchronize(st));
      if (still == 0) break;
    }
    carried[d] = wrapped;
  }
  if (listed && n > 0) {
    Arr4 src{}, dst{};
    for (int a = 0; a < narr; ++a) {
      src.a[a] = arr[carry[a]];
      dst.a[a] = spare ? spare[carry[a]] : w.migbuf.as<double>() + 6 * ni + (size_t)a * 3 * ni;
    }
    ListCols col{};
    for (int d = 0; d < 3; ++d) col.src[d] = carried[d] ? colbuf[d][cur] : nullptr;
    hipLaunchKernelGGL(k_list_gather, grid_for(3 * n), dim3(256), 0, st, n, (const int *)plist[cur], src, dst, narr,
                       col);
    PM_HIP(hipGetLastError());
    if (spare) {
      *in_spare = 1;
      *np_out = n;
      return 0;
    }
    // back into the caller's arrays: one launch of 16-byte words when every
    // pointer allows it
    bool al = true;
    for (int a = 0; a < narr; ++a) al = al && ((uintptr_t)src.a[a] % 16 == 0) && ((uintptr_t)dst.a[a] % 16 == 0);
    if (al) {
      hipLaunchKernelGGL(k_copy_back4, grid_for((3 * n + 1) / 2), dim3(256), 0, st, 3 * n, dst, src,
                         narr);
    } else {
      for (int a = 0; a < narr; ++a)
        PM_HIP(hipMemcpyAsync(src.a[a], dst.a[a], sizeof(double) * 3 * n, hipMemcpyDeviceToDevice, st));
    }
  }
  PM_HIP(hipGetLastError());
  *np_out = n;
  return 0;
}

int cell_index_device(const pmgpu_geom *geom, const double *pos, int64_t np, int32_t *out, hipStream_t st) {
  PGeom g;
  if (make_pgeom(geom, g)) return -1;
  if (np > 0) hipLaunchKernelGGL(k_cell_index, grid_for(3 * np), dim3(256), 0, st, g, pos, np, out);
  PM_HIP(hipGetLastError());
  return 0;
}

}  // namespace pmgpu

using namespace pmgpu;

// ------------------------------------------------------------------ C ABI
namespace {
struct Dev {
  void *p = nullptr;
  ~Dev() {
    if (p) (void)hipFree(p);
  }
};
int to_dev(Dev &d, const void *h, size_t bytes) {
  PM_HIP(hipMalloc(&d.p, bytes > 0 ? bytes : 8));
  if (h && bytes) PM_HIP(hipMemcpy(d.p, h, bytes, hipMemcpyHostToDevice));
  return 0;
}
int mesh_bytes(const pmgpu_geom *g, size_t &b) {
  int32_t e[3];
  if (pmgpu_mesh_extent(g, e)) return -1;
  b = sizeof(double) * 3 * (size_t)e[0] * e[1] * e[2];
  return 0;
}
}  // namespace

extern "C" {

int pmgpu_interp_particle_mesh(const pmgpu_geom *g, const double *xpart, const double *part, int64_t np,
                               double *mesh, int32_t clear, int32_t clear_ghost) {
  PM_CHECK(g && mesh && (np == 0 || (xpart && part)), "pmgpu_interp_particle_mesh: null argument");
  size_t mb;
  if (mesh_bytes(g, mb)) return -1;
  Dev dp, dv, dm;
  if (to_dev(dp, xpart, sizeof(double) * 3 * np) || to_dev(dv, part, sizeof(double) * 3 * np) ||
      to_dev(dm, clear ? nullptr : mesh, mb))
    return -1;
  if (p2m_device_full(g, PMGPU_FP64, (double *)dp.p, (double *)dv.p, np, dm.p, clear, clear_ghost, nullptr))
    return -1;
  PM_HIP(hipMemcpy(mesh, dm.p, mb, hipMemcpyDeviceToHost));
  return 0;
}

int pmgpu_interp_particle_mesh_device(const pmgpu_geom *g, int32_t precision, const void *d_xpart,
                                      const void *d_part, int64_t np, void *d_mesh, int32_t clear,
                                      int32_t clear_ghost, void *stream) {
  return p2m_device_full(g, precision, (const double *)d_xpart, (const double *)d_part, np, d_mesh, clear,
                         clear_ghost, (hipStream_t)stream);
}

int pmgpu_interp_mesh_particle(const pmgpu_geom *g, const double *mesh, const double *pos, int64_t np, double *out) {
  PM_CHECK(g && mesh && (np == 0 || (pos && out)), "pmgpu_interp_mesh_particle: null argument");
  size_t mb;
  if (mesh_bytes(g, mb)) return -1;
  Dev dm, dp, dout;
  if (to_dev(dm, mesh, mb) || to_dev(dp, pos, sizeof(double) * 3 * np) || to_dev(dout, nullptr, sizeof(double) * 3 * np))
    return -1;
  if (m2p_device(g, PMGPU_FP64, dm.p, (double *)dp.p, np, (double *)dout.p, nullptr)) return -1;
  if (np) PM_HIP(hipMemcpy(out, dout.p, sizeof(double) * 3 * np, hipMemcpyDeviceToHost));
  return 0;
}

int pmgpu_interp_mesh_particle_device(const pmgpu_geom *g, int32_t precision, const void *d_mesh,
                                      const void *d_pos, int64_t np, void *d_out, void *stream) {
  return m2p_device(g, precision, d_mesh, (const double *)d_pos, np, (double *)d_out, (hipStream_t)stream);
}

int pmgpu_remesh(const pmgpu_geom *g, const double *vort_m, const double *dvort_m, const double *vel_m,
                 const uint8_t *mask, double min_vort, double min_dvort, int32_t dvort_flag, int32_t vel_flag,
                 int32_t incl_ghost, double *pos, double *vort, double *vel, double *dvort, int64_t cap,
                 int64_t *np_out) {
  PM_CHECK(g && vort_m && np_out, "pmgpu_remesh: null argument");
  size_t mb;
  if (mesh_bytes(g, mb)) return -1;
  Dev dw, ddw, dvl, dmk;
  if (to_dev(dw, vort_m, mb) || (dvort_m && to_dev(ddw, dvort_m, mb)) || (vel_m && to_dev(dvl, vel_m, mb)) ||
      (mask && to_dev(dmk, mask, mb / 24)))
    return -1;
  int64_t n = 0;
  if (remesh_device(g, PMGPU_FP64, dw.p, ddw.p, dvl.p, (uint8_t *)dmk.p, min_vort, min_dvort, dvort_flag, vel_flag,
                    incl_ghost, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, &n, nullptr))
    return -1;
  *np_out = n;
  if (!pos || n > cap) return 0;
  Dev op, ov, ol, od;
  size_t pb = sizeof(double) * 3 * (size_t)n;
  if (to_dev(op, nullptr, pb) || to_dev(ov, nullptr, pb) || to_dev(ol, nullptr, pb) || to_dev(od, nullptr, pb))
    return -1;
  if (remesh_device(g, PMGPU_FP64, dw.p, ddw.p, dvl.p, (uint8_t *)dmk.p, min_vort, min_dvort, dvort_flag, vel_flag,
                    incl_ghost, (double *)op.p, (double *)ov.p, (double *)ol.p, (double *)od.p, nullptr, nullptr, n,
                    &n, nullptr))
    return -1;
  if (n) {
    PM_HIP(hipMemcpy(pos, op.p, pb, hipMemcpyDeviceToHost));
    PM_HIP(hipMemcpy(vort, ov.p, pb, hipMemcpyDeviceToHost));
    if (vel) PM_HIP(hipMemcpy(vel, ol.p, pb, hipMemcpyDeviceToHost));
    if (dvort) PM_HIP(hipMemcpy(dvort, od.p, pb, hipMemcpyDeviceToHost));
  }
  return 0;
}

int pmgpu_remesh_device(const pmgpu_geom *g, int32_t precision, const void *d_vort_m, const void *d_dvort_m,
                        const void *d_vel_m, const void *d_mask, double min_vort, double min_dvort,
                        int32_t dvort_flag, int32_t vel_flag, int32_t incl_ghost, void *d_pos, void *d_vort,
                        void *d_vel, void *d_dvort, void *d_vel_rk, void *d_dvort_rk, int64_t cap, int64_t *np_out,
                        void *stream) {
  PM_CHECK(np_out, "pmgpu_remesh_device: null argument");
  return remesh_device(g, precision, d_vort_m, d_dvort_m, d_vel_m, (const uint8_t *)d_mask, min_vort, min_dvort,
                       dvort_flag, vel_flag, incl_ghost,
                       (double *)d_pos, (double *)d_vort, (double *)d_vel, (double *)d_dvort, (double *)d_vel_rk,
                       (double *)d_dvort_rk, cap, np_out, (hipStream_t)stream);
}

int pmgpu_particles_advance(const pmgpu_geom *g, int32_t rk_order, int32_t stage, double dt, int32_t euler_vort,
                            double *pos, double *vel, double *vort, double *dvort, double *vel_rk, double *dvort_rk,
                            int64_t np, int64_t *np_out) {
  PM_CHECK(g && np_out && (np == 0 || (pos && vel && vort && dvort && vel_rk && dvort_rk)),
           "pmgpu_particles_advance: null argument");
  double *h[6] = {pos, vel, vort, dvort, vel_rk, dvort_rk};
  Dev d[6];
  double *dp[6];
  size_t pb = sizeof(double) * 3 * (size_t)np;
  for (int a = 0; a < 6; ++a) {
    if (to_dev(d[a], h[a], pb)) return -1;
    dp[a] = (double *)d[a].p;
  }
  int64_t n = 0;
  if (advance_device(g, rk_order, stage, dt, euler_vort, dp, np, &n, nullptr, nullptr, nullptr)) return -1;
  for (int a = 0; a < 6; ++a)
    if (n) PM_HIP(hipMemcpy(h[a], dp[a], sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost));
  *np_out = n;
  return 0;
}

int pmgpu_particles_advance_device(const pmgpu_geom *g, int32_t rk_order, int32_t stage, double dt,
                                   int32_t euler_vort, void *const d_arrays[6], int64_t np, int64_t *np_out,
                                   void *stream) {
  PM_CHECK(g && d_arrays && np_out, "pmgpu_particles_advance_device: null argument");
  double *dp[6];
  for (int a = 0; a < 6; ++a) dp[a] = (double *)d_arrays[a];
  return advance_device(g, rk_order, stage, dt, euler_vort, dp, np, np_out, (hipStream_t)stream, nullptr, nullptr);
}

int pmgpu_particles_advance_device_spare(const pmgpu_geom *g, int32_t rk_order, int32_t stage, double dt,
                                         int32_t euler_vort, void *const d_arrays[6], void *const d_spare[6],
                                         int64_t np, int64_t *np_out, int32_t *in_spare, void *stream) {
  PM_CHECK(g && d_arrays && d_spare && np_out && in_spare, "pmgpu_particles_advance_device_spare: null argument");
  double *dp[6], *ds[6];
  for (int a = 0; a < 6; ++a) {
    dp[a] = (double *)d_arrays[a];
    ds[a] = (double *)d_spare[a];
  }
  const int carried[4] = {0, 2, 4, 5};
  for (int a = 0; a < (rk_order > 1 ? (dp[3] ? 4 : 3) : 2); ++a)
    PM_CHECK(ds[carried[a]] && ds[carried[a]] != dp[carried[a]],
             "pmgpu_particles_advance_device_spare: each carried spare array must be given and distinct");
  int flag = 0;
  const int rc = advance_device(g, rk_order, stage, dt, euler_vort, dp, np, np_out, (hipStream_t)stream, ds, &flag);
  *in_spare = flag;
  return rc;
}

int pmgpu_cell_index(const pmgpu_geom *g, const double *pos, int64_t np, int32_t *idx) {
  PM_CHECK(g && (np == 0 || (pos && idx)), "pmgpu_cell_index: null argument");
  Dev dp, di;
  if (to_dev(dp, pos, sizeof(double) * 3 * np) || to_dev(di, nullptr, sizeof(int32_t) * 3 * np)) return -1;
  if (cell_index_device(g, (double *)dp.p, np, (int32_t *)di.p, nullptr)) return -1;
  if (np) PM_HIP(hipMemcpy(idx, di.p, sizeof(int32_t) * 3 * np, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"

// ===================================================== device-resident step
// One vortex time step (per RK stage: P2M, solve, M2P, RK update and
// migration; then P2M and remesh) with every count kept in device memory:
// the grids cover the capacity and the kernels read the live particle count
// (NCount), so the step has no host synchronisation and is recorded once as
// a HIP graph and replayed.  The same kernels and the same per-particle
// arithmetic as the eager entry points, in the client's order
// (vortex_codes/vortex_structures/source/main.f:453-505).
namespace pmgpu {

struct Arr6 {
  double *a[6];
};


__global__ void k_fill_int(int *p, int64_t n, int v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}
// Three-way stable partition of the recorded step, gated: when the stage's
// update left no particle outside the patch (or the axis moves none), every
// kernel below returns at once.  The class offsets come from a two-level scan
// over tiles of kPartTile particles (tile sums, one block scanning the tile
// sums, then each tile's own block scans in index order), so that a step with
// nothing to migrate launches no full-size scan.
// class counts of each tile
__global__ void __launch_bounds__(kPartThreads) k_part_tiles(NCount npc, const int *counts, const int *cls,
                                                             Mig3 *tsum) {
  if (part_idle(counts)) return;
  const int64_t np = npc.get();
  Mig3 s{0, 0, 0};
#pragma unroll
  for (int j = 0; j < kPartItems; ++j) {
    const int64_t q = (int64_t)blockIdx.x * kPartTile + j * kPartThreads + threadIdx.x;
    if (q < np) s = m3_add(s, m3_of(cls[q]));
  }
  Mig3 tot;
  (void)block_scan3(s, tot);
  if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}
// west movers, east movers, then stayers, each in their original order; a
// stayer's rank is q minus the movers and drops before it
__global__ void __launch_bounds__(kPartThreads) k_part_migrate(NCount npc, const int *cls, const Mig3 *toff,
                                                               const int *counts, int d, double len, Arr6 src,
                                                               Arr6 dst) {
  if (part_idle(counts)) return;
  const int64_t np = npc.get();
  const int n1 = counts[1], n2 = counts[2];
  Mig3 run = toff[blockIdx.x];
  for (int j = 0; j < kPartItems; ++j) {
    const int64_t q = (int64_t)blockIdx.x * kPartTile + j * kPartThreads + threadIdx.x;
    const int c = q < np ? cls[q] : 0;
    Mig3 tot;
    const Mig3 ex = block_scan3(q < np ? m3_of(c) : Mig3{0, 0, 0}, tot);
    const Mig3 o = m3_add(run, ex);
    run = m3_add(run, tot);
    if (q >= np || c == 3) continue;
    const int64_t to = c == 1 ? o.w : (c == 2 ? n1 + o.e : n1 + n2 + (q - o.w - o.e - o.d));
    for (int a = 0; a < 6; ++a) {
      double v0 = src.a[a][3 * q], v1 = src.a[a][3 * q + 1], v2 = src.a[a][3 * q + 2];
      if (a == 0) {
        double &x = d == 0 ? v0 : (d == 1 ? v1 : v2);
        if (c == 1) x = x + len;
        if (c == 2) x = x - len;
      }
      dst.a[a][3 * to] = v0;
      dst.a[a][3 * to + 1] = v1;
      dst.a[a][3 * to + 2] = v2;
    }
  }
}
__global__ void k_copy_back(NCount npc, const int *counts, Arr6 src, Arr6 dst) {
  if (counts[1] + counts[2] + counts[3] == 0) return;
  const int64_t n = 3 * (npc.get() - counts[3]);
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    for (int a = 0; a < 6; ++a) dst.a[a][i] = src.a[a][i];
}
__global__ void k_np_drop(int *np, const int *counts) { *np -= counts[3]; }
// a particle still outside a periodic axis after one wrap moved more than a
// period in one stage: flagged (bit 2 of the status word)
__global__ void k_check_inside(NCount npc, const int *outside, const double *pos, int d, double lo, double hi,
                               int *err) {
  if (!*outside) return;
  const int64_t np = npc.get();
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < np; q += (int64_t)gridDim.x * blockDim.x) {
    const double x = pos[3 * q + d];
    if (x < lo || x >= hi) atomicOr(err, 2);
  }
}
// the remeshed count, capped (bit 1 of the status word: over capacity)
__global__ void k_np_from_offs(int *np, const int *offs, int64_t nn, int64_t cap, int *err) {
  const int64_t t = offs[nn];
  if (t > cap) atomicOr(err, 1);
  *np = (int)(t > cap ? cap : t);
}

struct VStep {
  pmgpu_geom geom{};
  PGeom g{};
  int prec = PMGPU_FP64, rk = 2;
  int64_t cap = 0;
  hipStream_t st = nullptr;
  pmgpu_poisson_t solver = nullptr;
  PartWS w;         // this step's own P2M / remesh workspace (the graph keeps its pointers)
  Scratch parts;    // the particle set: 6 arrays of 3 * cap doubles
  Scratch tmp;      // migration target: 6 arrays
  Scratch mesh;     // vorticity and velocity meshes
  Scratch mig;      // cls, counts, tile offsets
  Scratch state;    // np, the status word and the stage's outside flag
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double g_dt = 0.0, g_min = 0.0;

  size_t mesh_elems() const { return 3 * (size_t)g.E[0] * g.E[1] * g.E[2]; }
  Arr6 arrays(Scratch &s) {
    Arr6 r;
    for (int a = 0; a < 6; ++a) r.a[a] = s.as<double>() + (size_t)a * 3 * cap;
    return r;
  }
  int *d_np() { return state.as<int>(); }
  int *d_err() { return state.as<int>() + 1; }
  int *d_outside() { return state.as<int>() + 2; }
  int ntiles() const { return (int)((cap + kPartTile - 1) / kPartTile); }
  void *mvort() { return mesh.p; }
  void *mvel() { return (char *)mesh.p + mesh_elems() * (prec == PMGPU_FP64 ? 8 : 4); }

  ~VStep() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (solver) (void)pmgpu_poisson_destroy(solver);
    if (st) (void)hipStreamDestroy(st);
  }

  // every buffer at its final size before anything is recorded
  int prepare() {
    const int nbins = g.nb[0] * g.nb[1] * g.nb[2];
    const int64_t nn = mesh_cells(g);
    PM_CHECK(cap < ((int64_t)1 << 31) && nbins < (1 << 30) && nn < ((int64_t)1 << 31),
             "pmgpu_vstep: capacity or mesh too large");
    const size_t rec_bytes = (prec == PMGPU_FP64 ? sizeof(double) : sizeof(float)) * PRec<double, 6>::kSlots;
    if (w.key.ensure(sizeof(int) * ((size_t)cap + 1)) || w.ids.ensure(sizeof(int) * 2 * ((size_t)cap + 1)) ||
        w.count.ensure(sizeof(int) * (nbins + 1)) || w.start.ensure(sizeof(int) * (nbins + 1)) ||
        w.rec.ensure(rec_bytes * ((size_t)cap + 1)) || w.flag.ensure(sizeof(int) * (nn + 1)) ||
        w.offs.ensure(sizeof(int) * (nn + 1)))
      return -1;
    const int64_t ts = iscan_tiles(nbins > nn ? nbins : nn);
    if (w.scan.ensure(sizeof(int) * ((size_t)ts + 1))) return -1;
    const size_t pb = sizeof(double) * 3 * (size_t)cap * 6;
    if (parts.ensure(pb) || tmp.ensure(pb) || mesh.ensure(2 * mesh_elems() * (prec == PMGPU_FP64 ? 8 : 4)) ||
        mig.ensure(sizeof(int) * ((size_t)cap + 8) + sizeof(Mig3) * ((size_t)ntiles() + 1)) ||
        state.ensure(sizeof(int) * 8))
      return -1;
    return 0;
  }
  int *cls() { return mig.as<int>(); }
  int *counts() { return mig.as<int>() + cap; }
  Mig3 *toff() { return reinterpret_cast<Mig3 *>(mig.as<int>() + cap + 8); }

  // deposition of the particle vorticity onto the vorticity mesh (clear)
  int p2m() {
    const int nbins = g.nb[0] * g.nb[1] * g.nb[2];
    const NCount n(cap, d_np());
    Arr6 A = arrays(parts);
    int *key = w.key.as<int>(), *rank = w.ids.as<int>(), *id_out = rank + cap + 1;
    hipLaunchKernelGGL(k_fill_int, grid_for(nbins + 1), dim3(256), 0, st, w.count.as<int>(), (int64_t)nbins + 1, 0);
    hipLaunchKernelGGL(k_p2m_keys, grid_for(cap), dim3(256), 0, st, g, (const double *)A.a[0], n, nbins, key, rank,
                       w.count.as<int>());
    if (iscan_launch(w.count.as<int>(), w.start.as<int>(), nbins, w.scan.as<int>(), st)) return -1;
    hipLaunchKernelGGL(k_p2m_scatter, grid_for(cap), dim3(256), 0, st, n, nbins, key, rank, w.start.as<int>(),
                       id_out);
    if (prec == PMGPU_FP64) {
      if (g.ns == 4) p2m_stage_gather<double, 4>(g, A.a[0], A.a[2], cap, id_out, w, nbins, (double *)mvort(), 0, st);
      else p2m_stage_gather<double, 6>(g, A.a[0], A.a[2], cap, id_out, w, nbins, (double *)mvort(), 0, st);
    } else {
      if (g.ns == 4) p2m_stage_gather<float, 4>(g, A.a[0], A.a[2], cap, id_out, w, nbins, (float *)mvort(), 0, st);
      else p2m_stage_gather<float, 6>(g, A.a[0], A.a[2], cap, id_out, w, nbins, (float *)mvort(), 0, st);
    }
    for (int d = 0; d < 3; ++d)
      if (g.bc[d] != 0) {
        const dim3 grid = grid_for(2LL * g.ngl[d] * 3 * g.E[(d + 1) % 3] * g.E[(d + 2) % 3]);
        if (prec == PMGPU_FP64)
          hipLaunchKernelGGL(k_halo_add<double>, grid, dim3(256), 0, st, (double *)mvort(), g, d);
        else
          hipLaunchKernelGGL(k_halo_add<float>, grid, dim3(256), 0, st, (float *)mvort(), g, d);
      }
    PM_HIP(hipGetLastError());
    return 0;
  }

  int migrate(int d) {
    const NCount n(cap, d_np());
    Arr6 A = arrays(parts), T = arrays(tmp);
    const double lo = geom.xmin[d], hi = geom.xmax[d], len = geom.xmax[d] - geom.xmin[d];
    const int per = geom.bound_cond[d] != 0;
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, counts(), (int64_t)4, 0);
    hipLaunchKernelGGL(k_classify, grid_for(cap), dim3(256), 0, st, n, (const int *)d_outside(),
                       (const double *)A.a[0], d, lo, hi, per, cls(), counts());
    const dim3 tiles((unsigned)ntiles());
    hipLaunchKernelGGL(k_part_tiles, tiles, dim3(kPartThreads), 0, st, n, (const int *)counts(), (const int *)cls(),
                       toff());
    hipLaunchKernelGGL(k_part_tile_scan, dim3(1), dim3(kTileScanThreads), 0, st, (const int *)counts(), toff(), ntiles());
    hipLaunchKernelGGL(k_part_migrate, tiles, dim3(kPartThreads), 0, st, n, (const int *)cls(), (const Mig3 *)toff(),
                       (const int *)counts(), d, len, A, T);
    hipLaunchKernelGGL(k_copy_back, grid_for(3 * cap), dim3(256), 0, st, n, (const int *)counts(), T, A);
    hipLaunchKernelGGL(k_np_drop, dim3(1), dim3(1), 0, st, d_np(), (const int *)counts());
    if (per)
      hipLaunchKernelGGL(k_check_inside, grid_for(cap), dim3(256), 0, st, NCount(cap, d_np()),
                         (const int *)d_outside(), (const double *)A.a[0], d, lo, hi, d_err());
    PM_HIP(hipGetLastError());
    return 0;
  }

  // the whole step, stream-ordered (recorded into the graph)
  int enqueue(double dt, double min_vort) {
    Arr6 A = arrays(parts);
    const NCount n(cap, d_np());
    for (int s = 1; s <= rk; ++s) {
      if (p2m()) return -1;
      if (pmgpu_poisson_solve_device(solver, mvort(), mvel(), nullptr, nullptr, 0, 0, 0)) return -1;
      if (prec == PMGPU_FP64) {
        if (g.ns == 4)
          hipLaunchKernelGGL((k_m2p<double, 4>), grid_for(cap), dim3(256), 0, st, g, (const double *)mvel(),
                             (const double *)A.a[0], n, A.a[1]);
        else
          hipLaunchKernelGGL((k_m2p<double, 6>), grid_for(cap), dim3(256), 0, st, g, (const double *)mvel(),
                             (const double *)A.a[0], n, A.a[1]);
      } else {
        if (g.ns == 4)
          hipLaunchKernelGGL((k_m2p<float, 4>), grid_for(cap), dim3(256), 0, st, g, (const float *)mvel(),
                             (const double *)A.a[0], n, A.a[1]);
        else
          hipLaunchKernelGGL((k_m2p<float, 6>), grid_for(cap), dim3(256), 0, st, g, (const float *)mvel(),
                             (const double *)A.a[0], n, A.a[1]);
      }
      const RkCoef k = rk_coef(rk, s, 0);
      PBox box;
      for (int d = 0; d < 3; ++d) {
        box.lo[d] = geom.xmin[d];
        box.hi[d] = geom.xmax[d];
      }
      hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, d_outside(), (int64_t)1, 0);
      // no right-hand side in this step: dvort and dvort_rk stay zero
      hipLaunchKernelGGL(k_rk_update, grid_for(3 * cap), dim3(256), 0, st, k, dt, n, A.a[0], (const double *)A.a[1],
                         A.a[2], (const double *)nullptr, A.a[4], (double *)nullptr, box, d_outside());
      PM_HIP(hipGetLastError());
      for (int d = 0; d < 3; ++d)
        if (migrate(d)) return -1;
    }
    // remesh: deposit, flag, scan, create (velocity from the last stage's mesh)
    if (p2m()) return -1;
    const int64_t nn = mesh_cells(g);
    if (prec == PMGPU_FP64)
      remesh_flags_launch<double>(g, nn, mvort(), nullptr, nullptr, min_vort, 0.0, 0, w.flag.as<int>(), st);
    else
      remesh_flags_launch<float>(g, nn, mvort(), nullptr, nullptr, min_vort, 0.0, 0, w.flag.as<int>(), st);
    if (iscan_launch(w.flag.as<int>(), w.offs.as<int>(), nn, w.scan.as<int>(), st)) return -1;
    if (prec == PMGPU_FP64)
      hipLaunchKernelGGL(k_remesh_create<double>, grid_for(nn), dim3(256), 0, st, g, (const int *)w.flag.as<int>(),
                         (const int *)w.offs.as<int>(), (const double *)mvort(), (const double *)nullptr,
                         (const double *)mvel(), A.a[0], A.a[2], A.a[1], A.a[3], A.a[4], A.a[5], cap);
    else
      hipLaunchKernelGGL(k_remesh_create<float>, grid_for(nn), dim3(256), 0, st, g, (const int *)w.flag.as<int>(),
                         (const int *)w.offs.as<int>(), (const float *)mvort(), (const float *)nullptr,
                         (const float *)mvel(), A.a[0], A.a[2], A.a[1], A.a[3], A.a[4], A.a[5], cap);
    hipLaunchKernelGGL(k_np_from_offs, dim3(1), dim3(1), 0, st, d_np(), (const int *)w.offs.as<int>(), nn, cap,
                       d_err());
    PM_HIP(hipGetLastError());
    return 0;
  }

  int record(double dt, double min_vort) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    PM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int e = enqueue(dt, min_vort);
    hipGraph_t gr = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &gr);
    if (e) {
      if (gr) (void)hipGraphDestroy(gr);
      return -1;
    }
    PM_HIP(ce);
    graph = gr;
    PM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    g_dt = dt;
    g_min = min_vort;
    return 0;
  }
};

}  // namespace pmgpu

struct pmgpu_vstep_s {
  pmgpu::VStep s;
};

extern "C" {

int pmgpu_vstep_create(const pmgpu_geom *g, const pmgpu_params *p, int32_t precision, int32_t rk_order,
                       int64_t capacity, pmgpu_vstep_t *out) {
  PM_CHECK(g && p && out, "pmgpu_vstep_create: null argument");
  PM_CHECK(precision == PMGPU_FP64 || precision == PMGPU_FP32, "pmgpu_vstep_create: bad precision");
  PM_CHECK(rk_order >= 1 && rk_order <= 3, "pmgpu_vstep_create: time integration order must be 1, 2 or 3");
  PM_CHECK(capacity > 0, "pmgpu_vstep_create: capacity must be positive");
  auto h = new pmgpu_vstep_s;
  VStep &s = h->s;
  s.geom = *g;
  s.prec = precision;
  s.rk = rk_order;
  s.cap = capacity;
  int e = make_pgeom(g, s.g);
  if (!e) {
    for (int d = 0; d < 3; ++d)
      if (s.g.bc[d] != 0 && s.g.ncell[d] < 2 * s.g.ngl[d]) {
        set_error("pmgpu: periodic axis shorter than two ghost layers");
        e = -1;
      }
  }
  if (!e && hipStreamCreateWithFlags(&s.st, hipStreamNonBlocking) != hipSuccess) {
    set_error("pmgpu_vstep_create: stream creation failed");
    e = -1;
  }
  if (!e) e = pmgpu_poisson_create(g, p, precision, s.st, &s.solver);
  if (!e) e = pmgpu_poisson_setup(s.solver);
  if (!e) e = s.prepare();
  if (!e && (hipMemsetAsync(s.state.p, 0, sizeof(int) * 8, s.st) != hipSuccess ||
             hipMemsetAsync(s.parts.p, 0, s.parts.cap, s.st) != hipSuccess ||
             hipStreamSynchronize(s.st) != hipSuccess)) {
    set_error("pmgpu_vstep_create: initialisation failed");
    e = -1;
  }
  if (e) {
    delete h;
    return -1;
  }
  *out = h;
  return 0;
}

int pmgpu_vstep_destroy(pmgpu_vstep_t h) {
  delete h;
  return 0;
}

int pmgpu_vstep_load(pmgpu_vstep_t h, const void *d_pos, const void *d_vort, int64_t np) {
  PM_CHECK(h && (np == 0 || (d_pos && d_vort)), "pmgpu_vstep_load: null argument");
  VStep &s = h->s;
  PM_CHECK(np >= 0 && np <= s.cap, "pmgpu_vstep_load: more particles than the capacity");
  // the particle set starts from (pos, vort) with every other field zero
  PM_HIP(hipMemsetAsync(s.parts.p, 0, s.parts.cap, s.st));
  Arr6 A = s.arrays(s.parts);
  if (np) {
    PM_HIP(hipMemcpyAsync(A.a[0], d_pos, 24 * (size_t)np, hipMemcpyDeviceToDevice, s.st));
    PM_HIP(hipMemcpyAsync(A.a[2], d_vort, 24 * (size_t)np, hipMemcpyDeviceToDevice, s.st));
  }
  const int st[2] = {(int)np, 0};
  PM_HIP(hipMemcpyAsync(s.state.p, st, sizeof(st), hipMemcpyHostToDevice, s.st));
  PM_HIP(hipStreamSynchronize(s.st));
  return 0;
}

int pmgpu_vstep_step(pmgpu_vstep_t h, double dt, double min_vort, int32_t use_graph) {
  PM_CHECK(h, "pmgpu_vstep_step: null handle");
  VStep &s = h->s;
  if (!use_graph) return s.enqueue(dt, min_vort);
  if (!s.exec || s.g_dt != dt || s.g_min != min_vort)
    if (s.record(dt, min_vort)) return -1;
  PM_HIP(hipGraphLaunch(s.exec, s.st));
  return 0;
}

int pmgpu_vstep_sync(pmgpu_vstep_t h, int64_t *np, int32_t *status) {
  PM_CHECK(h, "pmgpu_vstep_sync: null handle");
  VStep &s = h->s;
  int st[2] = {0, 0};
  PM_HIP(hipMemcpyAsync(st, s.state.p, sizeof(st), hipMemcpyDeviceToHost, s.st));
  PM_HIP(hipStreamSynchronize(s.st));
  if (np) *np = st[0];
  if (status) *status = st[1];
  return 0;
}

int pmgpu_vstep_fetch(pmgpu_vstep_t h, double *pos, double *vort, double *vel, int64_t cap, int64_t *np) {
  PM_CHECK(h && np, "pmgpu_vstep_fetch: null argument");
  VStep &s = h->s;
  int st[2] = {0, 0};
  PM_HIP(hipMemcpyAsync(st, s.state.p, sizeof(st), hipMemcpyDeviceToHost, s.st));
  PM_HIP(hipStreamSynchronize(s.st));
  *np = st[0];
  PM_CHECK(st[0] <= cap, "pmgpu_vstep_fetch: host arrays too small");
  Arr6 A = s.arrays(s.parts);
  double *dst[3] = {pos, vort, vel};
  const int src[3] = {0, 2, 1};
  for (int q = 0; q < 3; ++q)
    if (dst[q] && st[0]) PM_HIP(hipMemcpy(dst[q], A.a[src[q]], 24 * (size_t)st[0], hipMemcpyDeviceToHost));
  return 0;
}

int pmgpu_vstep_arrays(pmgpu_vstep_t h, void *d_arrays[6], void **stream) {
  PM_CHECK(h && d_arrays, "pmgpu_vstep_arrays: null argument");
  VStep &s = h->s;
  Arr6 A = s.arrays(s.parts);
  for (int a = 0; a < 6; ++a) d_arrays[a] = A.a[a];
  if (stream) *stream = s.st;
  return 0;
}

}  // extern "C"
