// Global diagnostics in FP64 (HIP, gfx950): the vortex-ring particle
// diagnostics and the mesh energy integrals, reduced deterministically.
//
// Particle form (vortex_codes/vortex_structures/source/diagnostics/
// diagnostics_ring.f:72-226): impulse, angular impulse, kinetic energy,
// enstrophy, helicity and the enstrophy-weighted centroid in a first sweep,
// ring radius and circulation about that centroid in a second, core radius
// about the ring radius in a third.  Mesh form
// (.../diagnostics/diagnostics_energy_mesh.f:72-134): kinetic energy,
// enstrophy, helicity and Lamb energy over the interior cells.
//
// Each sweep runs a fixed grid (at most kBlocks blocks of kThreads threads,
// a function of the element count only); a thread sums its grid-stride
// elements in index order, the block reduces by a fixed tree into one
// partial per block, and one block sums the partials in block order, so
// every result is run-to-run identical.  Every per-element term is formed in
// the reference's operation order with explicitly rounded FP64 operations;
// only the order of the sum differs from the reference's serial loop.
// Sweeps 2 and 3 read the totals of the previous sweep from device memory,
// so the three sweeps are stream-ordered with no host round trip.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/pmgpu.h"
#include "pm_common.h"

namespace pmgpu {

constexpr int kThreads = 256;
constexpr int kBlocks = 1024;
constexpr int kParts = 16;  // partial slots per block

// totals (device): particle sweep slots 0..15 in the order of the public
// output, the mesh sweep slots 0..4
enum {
  T_IMP = 0, T_ANG = 3, T_KIN = 6, T_ENST = 7, T_HEL = 8, T_CIRC = 9, T_CEN = 10, T_RAD = 13, T_CORE = 14,
  T_MAX = 15
};

struct DiagDx {
  double x, y, z;
};

__device__ __forceinline__ double dmul(double a, double b) { return __dmul_rn(a, b); }
__device__ __forceinline__ double dadd(double a, double b) { return __dadd_rn(a, b); }
__device__ __forceinline__ double dsub(double a, double b) { return __dsub_rn(a, b); }
// (t * dx) * dy * dz, left to right as the reference writes it
__device__ __forceinline__ double vol(double t, DiagDx h) { return dmul(dmul(dmul(t, h.x), h.y), h.z); }
// a*a + b*b + c*c, left to right
__device__ __forceinline__ double sq3(double a, double b, double c) {
  return dadd(dadd(dmul(a, a), dmul(b, b)), dmul(c, c));
}
__device__ __forceinline__ double dot3(const double *a, const double *b) {
  return dadd(dadd(dmul(a[0], b[0]), dmul(a[1], b[1])), dmul(a[2], b[2]));
}

// Block reduction of NS sums (fixed tree) and NM maxima; thread 0 writes
// them to out[0..NS+NM).
template <int NS, int NM> __device__ void block_reduce(double (&v)[NS + NM], double *out) {
  __shared__ double sh[NS + NM][kThreads];
  for (int q = 0; q < NS + NM; ++q) sh[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      for (int q = 0; q < NS; ++q) sh[q][threadIdx.x] = dadd(sh[q][threadIdx.x], sh[q][threadIdx.x + o]);
      for (int q = NS; q < NS + NM; ++q) sh[q][threadIdx.x] = fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < NS + NM; ++q) out[q] = sh[q][0];
}

// ---- particle sweeps
// sweep 1: impulse(3), angular impulse(3), kinetic energy, enstrophy,
// helicity, centroid sums(3); maximum |vort|
__global__ void __launch_bounds__(kThreads)
    k_diag_p1(const double *__restrict__ pos, const double *__restrict__ vort, const double *__restrict__ vel,
              int64_t np, DiagDx h, double *__restrict__ partial) {
  const double c_1_3 = 1.0 / 3.0;
  double v[13] = {};
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < np; i += (int64_t)gridDim.x * kThreads) {
    const double x[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    const double w[3] = {vort[3 * i], vort[3 * i + 1], vort[3 * i + 2]};
    const double u[3] = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
    double m1[3], m2[3];
    m1[0] = dsub(dmul(x[1], w[2]), dmul(x[2], w[1]));
    m1[1] = dsub(dmul(x[2], w[0]), dmul(x[0], w[2]));
    m1[2] = dsub(dmul(x[0], w[1]), dmul(x[1], w[0]));
    m2[0] = dsub(dmul(x[1], m1[2]), dmul(x[2], m1[1]));
    m2[1] = dsub(dmul(x[2], m1[0]), dmul(x[0], m1[2]));
    m2[2] = dsub(dmul(x[0], m1[1]), dmul(x[1], m1[0]));
    for (int d = 0; d < 3; ++d) v[d] = dadd(v[d], vol(dmul(0.5, m1[d]), h));
    for (int d = 0; d < 3; ++d) v[3 + d] = dadd(v[3 + d], vol(dmul(c_1_3, m2[d]), h));
    v[6] = dadd(v[6], vol(dot3(u, m1), h));
    const double e = sq3(w[0], w[1], w[2]);
    v[7] = dadd(v[7], vol(e, h));
    v[8] = dadd(v[8], vol(dot3(u, w), h));
    for (int d = 0; d < 3; ++d) v[9 + d] = dadd(v[9 + d], vol(dmul(e, x[d]), h));
    v[12] = fmax(v[12], __dsqrt_rn(e));
  }
  block_reduce<12, 1>(v, partial + kParts * blockIdx.x);
}

// sweep 2: ring radius and circulation about the centroid in tot[T_CEN..]
__global__ void __launch_bounds__(kThreads)
    k_diag_p2(const double *__restrict__ pos, const double *__restrict__ vort, int64_t np, DiagDx h,
              const double *__restrict__ tot, double *__restrict__ partial) {
  const double cy = tot[T_CEN + 1], cz = tot[T_CEN + 2];
  double v[2] = {};
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < np; i += (int64_t)gridDim.x * kThreads) {
    const double w0 = vort[3 * i], w1 = vort[3 * i + 1], w2 = vort[3 * i + 2];
    const double e = sq3(w0, w1, w2);
    const double dy = dsub(pos[3 * i + 1], cy), dz = dsub(pos[3 * i + 2], cz);
    const double r2 = dadd(dmul(dy, dy), dmul(dz, dz));
    v[0] = dadd(v[0], vol(dmul(e, __dsqrt_rn(r2)), h));
    // sum - dz / r2 * w1 + dy / r2 * w2, left to right
    v[1] = dadd(dsub(v[1], dmul(__ddiv_rn(dz, r2), w1)), dmul(__ddiv_rn(dy, r2), w2));
  }
  block_reduce<2, 0>(v, partial + kParts * blockIdx.x);
}

// sweep 3: core radius about the ring of radius tot[T_RAD]
__global__ void __launch_bounds__(kThreads)
    k_diag_p3(const double *__restrict__ pos, const double *__restrict__ vort, int64_t np, DiagDx h,
              const double *__restrict__ tot, double *__restrict__ partial) {
  const double cx = tot[T_CEN], cy = tot[T_CEN + 1], cz = tot[T_CEN + 2], rad = tot[T_RAD];
  double v[1] = {};
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < np; i += (int64_t)gridDim.x * kThreads) {
    const double e = sq3(vort[3 * i], vort[3 * i + 1], vort[3 * i + 2]);
    const double dy = dsub(pos[3 * i + 1], cy), dz = dsub(pos[3 * i + 2], cz), dx = dsub(pos[3 * i], cx);
    const double s = dsub(__dsqrt_rn(dadd(dmul(dy, dy), dmul(dz, dz))), rad);
    v[0] = dadd(v[0], vol(dmul(e, __dsqrt_rn(dadd(dmul(s, s), dmul(dx, dx)))), h));
  }
  block_reduce<1, 0>(v, partial + kParts * blockIdx.x);
}

// Sum of the per-block partials in block order (thread t sums blocks t,
// t + 256, ... then a fixed tree), then the normalisations of each sweep.
template <int NS, int NM> __device__ void sum_partials(const double *partial, int nb, double (&r)[NS + NM]) {
  double v[NS + NM] = {};
  for (int b = threadIdx.x; b < nb; b += kThreads)
    for (int q = 0; q < NS + NM; ++q)
      v[q] = q < NS ? dadd(v[q], partial[kParts * b + q]) : fmax(v[q], partial[kParts * b + q]);
  __shared__ double res[NS + NM];
  block_reduce<NS, NM>(v, res);
  __syncthreads();
  for (int q = 0; q < NS + NM; ++q) r[q] = res[q];
}

__global__ void __launch_bounds__(kThreads) k_diag_p1_fin(const double *partial, int nb, double *tot) {
  double r[13];
  sum_partials<12, 1>(partial, nb, r);
  if (threadIdx.x == 0) {
    for (int q = 0; q < 9; ++q) tot[q] = r[q];
    for (int d = 0; d < 3; ++d) tot[T_CEN + d] = __ddiv_rn(r[9 + d], r[7]);
    tot[T_MAX] = r[12];
  }
}
__global__ void __launch_bounds__(kThreads) k_diag_p2_fin(const double *partial, int nb, DiagDx h, double *tot) {
  double r[2];
  sum_partials<2, 0>(partial, nb, r);
  if (threadIdx.x == 0) {
    const double PI = 3.141592653589793;  // the double nearest pi (acos(-1) in FP64)
    tot[T_RAD] = __ddiv_rn(r[0], tot[T_ENST]);
    // circulation * dx * dy * dz / (2 pi)
    tot[T_CIRC] = __ddiv_rn(vol(r[1], h), dmul(2.0, PI));
  }
}
__global__ void __launch_bounds__(kThreads) k_diag_p3_fin(const double *partial, int nb, double *tot) {
  double r[1];
  sum_partials<1, 0>(partial, nb, r);
  if (threadIdx.x == 0) tot[T_CORE] = __ddiv_rn(r[0], tot[T_ENST]);
}

// ---- mesh sweep: kinetic energy, enstrophy, helicity, Lamb energy; max
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_diag_mesh(const T *__restrict__ vel, const T *__restrict__ vort, int N0, int N1, int N2, int E0, int E1,
                int ng0, int ng1, int ng2, DiagDx h, double *__restrict__ partial) {
  double v[5] = {};
  const int64_t n = (int64_t)N0 * N1 * N2;
  for (int64_t c = blockIdx.x * (int64_t)kThreads + threadIdx.x; c < n; c += (int64_t)gridDim.x * kThreads) {
    const int i = (int)(c % N0);
    const int64_t r = c / N0;
    const int j = (int)(r % N1), k = (int)(r / N1);
    const int64_t o = 3 * ((int64_t)(i + ng0) + (int64_t)E0 * ((j + ng1) + (int64_t)E1 * (k + ng2)));
    const double w[3] = {(double)vort[o], (double)vort[o + 1], (double)vort[o + 2]};
    const double u[3] = {(double)vel[o], (double)vel[o + 1], (double)vel[o + 2]};
    const double e = sq3(w[0], w[1], w[2]);
    v[0] = dadd(v[0], vol(dmul(0.5, sq3(u[0], u[1], u[2])), h));
    v[1] = dadd(v[1], vol(e, h));
    const double hh = dot3(w, u);
    v[2] = dadd(v[2], vol(__ddiv_rn(dmul(hh, hh), e), h));
    const double l0 = dsub(dmul(w[1], u[2]), dmul(w[2], u[1]));
    const double l1 = dsub(dmul(w[2], u[0]), dmul(w[0], u[2]));
    const double l2 = dsub(dmul(w[0], u[1]), dmul(w[1], u[0]));
    v[3] = dadd(v[3], vol(__ddiv_rn(sq3(l0, l1, l2), e), h));
    v[4] = fmax(v[4], __dsqrt_rn(e));
  }
  block_reduce<4, 1>(v, partial + kParts * blockIdx.x);
}
__global__ void __launch_bounds__(kThreads) k_diag_mesh_fin(const double *partial, int nb, double *tot) {
  double r[5];
  sum_partials<4, 1>(partial, nb, r);
  if (threadIdx.x == 0)
    for (int q = 0; q < 5; ++q) tot[q] = r[q];
}

// partials of kBlocks blocks and 32 totals, allocated at first use
struct DiagWS {
  double *partial = nullptr;
  double *tot = nullptr;
};
static DiagWS g_diag;
static std::mutex g_diag_mu;

static int diag_ws() {
  std::lock_guard<std::mutex> lock(g_diag_mu);
  if (!g_diag.partial) PM_HIP(hipMalloc(&g_diag.partial, sizeof(double) * kParts * kBlocks));
  if (!g_diag.tot) PM_HIP(hipMalloc(&g_diag.tot, sizeof(double) * 32));
  return 0;
}

void diagnostics_release() {
  std::lock_guard<std::mutex> lock(g_diag_mu);
  if (g_diag.partial) (void)hipFree(g_diag.partial);
  if (g_diag.tot) (void)hipFree(g_diag.tot);
  g_diag = DiagWS{};
}

static int grid_for(int64_t n) {
  const int64_t b = (n + kThreads - 1) / kThreads;
  return (int)(b < 1 ? 1 : (b > kBlocks ? kBlocks : b));
}

// copy nv totals to d_out (device) and/or h_out (host, synchronising)
static int deliver(const double *tot, int nv, double *d_out, double *h_out, hipStream_t st) {
  if (d_out) PM_HIP(hipMemcpyAsync(d_out, tot, sizeof(double) * nv, hipMemcpyDeviceToDevice, st));
  if (h_out) {
    double h[16];
    PM_HIP(hipMemcpyAsync(h, tot, sizeof(double) * nv, hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    std::memcpy(h_out, h, sizeof(double) * nv);
  }
  return 0;
}

}  // namespace pmgpu

using namespace pmgpu;

extern "C" {

int pmgpu_diagnostics_particles(const double dx[3], const double *d_pos, const double *d_vort, const double *d_vel,
                                int64_t np, double *d_out, double *h_out, void *stream) {
  PM_CHECK(dx && d_pos && d_vort && d_vel, "pmgpu_diagnostics_particles: null argument");
  PM_CHECK(np > 0, "pmgpu_diagnostics_particles: no particles (the centroid is undefined)");
  if (diag_ws()) return -1;
  hipStream_t st = (hipStream_t)stream;
  const DiagDx h{dx[0], dx[1], dx[2]};
  const int nb = grid_for(np);
  hipLaunchKernelGGL(k_diag_p1, dim3(nb), dim3(kThreads), 0, st, d_pos, d_vort, d_vel, np, h, g_diag.partial);
  hipLaunchKernelGGL(k_diag_p1_fin, dim3(1), dim3(kThreads), 0, st, (const double *)g_diag.partial, nb, g_diag.tot);
  hipLaunchKernelGGL(k_diag_p2, dim3(nb), dim3(kThreads), 0, st, d_pos, d_vort, np, h, (const double *)g_diag.tot,
                     g_diag.partial);
  hipLaunchKernelGGL(k_diag_p2_fin, dim3(1), dim3(kThreads), 0, st, (const double *)g_diag.partial, nb, h,
                     g_diag.tot);
  hipLaunchKernelGGL(k_diag_p3, dim3(nb), dim3(kThreads), 0, st, d_pos, d_vort, np, h, (const double *)g_diag.tot,
                     g_diag.partial);
  hipLaunchKernelGGL(k_diag_p3_fin, dim3(1), dim3(kThreads), 0, st, (const double *)g_diag.partial, nb, g_diag.tot);
  PM_HIP(hipGetLastError());
  return deliver(g_diag.tot, 16, d_out, h_out, st);
}

int pmgpu_diagnostics_mesh(const pmgpu_geom *g, int32_t precision, const void *d_vel, const void *d_vort,
                           double *d_out, double *h_out, void *stream) {
  PM_CHECK(g && d_vel && d_vort, "pmgpu_diagnostics_mesh: null argument");
  PM_CHECK(precision == PMGPU_FP64 || precision == PMGPU_FP32,
           "pmgpu_diagnostics_mesh: precision must be PMGPU_FP64 or PMGPU_FP32");
  int32_t E[3];
  if (pmgpu_mesh_extent(g, E)) return -1;
  if (diag_ws()) return -1;
  hipStream_t st = (hipStream_t)stream;
  const DiagDx h{g->dx[0], g->dx[1], g->dx[2]};
  int ng[3];
  for (int d = 0; d < 3; ++d) ng[d] = (E[d] - g->ncell[d]) / 2;
  const int nb = grid_for((int64_t)g->ncell[0] * g->ncell[1] * g->ncell[2]);
  if (precision == PMGPU_FP64)
    hipLaunchKernelGGL(k_diag_mesh<double>, dim3(nb), dim3(kThreads), 0, st, (const double *)d_vel,
                       (const double *)d_vort, g->ncell[0], g->ncell[1], g->ncell[2], E[0], E[1], ng[0], ng[1],
                       ng[2], h, g_diag.partial);
  else
    hipLaunchKernelGGL(k_diag_mesh<float>, dim3(nb), dim3(kThreads), 0, st, (const float *)d_vel,
                       (const float *)d_vort, g->ncell[0], g->ncell[1], g->ncell[2], E[0], E[1], ng[0], ng[1], ng[2],
                       h, g_diag.partial);
  hipLaunchKernelGGL(k_diag_mesh_fin, dim3(1), dim3(kThreads), 0, st, (const double *)g_diag.partial, nb,
                     g_diag.tot);
  PM_HIP(hipGetLastError());
  return deliver(g_diag.tot, 5, d_out, h_out, st);
}

}  // extern "C"
