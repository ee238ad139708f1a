// Shared device/host helpers for the pmgpu HIP library (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

namespace pmgpu {

template <typename R> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };
template <typename R> using cplx_t = typename cplx_of<R>::type;

template <typename V> __host__ __device__ __forceinline__ V cmake(decltype(V::x) a, decltype(V::x) b) {
  V r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return cmake<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return cmake<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V cmul(V a, V b) {
  return cmake<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V> __device__ __forceinline__ V cconj(V a) { return cmake<V>(a.x, -a.y); }
template <typename V, typename R> __device__ __forceinline__ V cscale(V a, R s) { return cmake<V>(a.x * s, a.y * s); }
// multiply by i*s with s = +1 or -1 (compile time)
template <int S, typename V> __device__ __forceinline__ V cmul_i(V a) {
  return S > 0 ? cmake<V>(-a.y, a.x) : cmake<V>(a.y, -a.x);
}

// Error state of the last failing call on this host thread.
void set_error(const std::string &msg);
const std::string &get_error();

#define PM_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      ::pmgpu::set_error(std::string(#call " failed: ") + hipGetErrorString(e_));      \
      return -1;                                                                       \
    }                                                                                  \
  } while (0)

#define PM_CHECK(cond, msg)          \
  do {                               \
    if (!(cond)) {                   \
      ::pmgpu::set_error(msg);       \
      return -1;                     \
    }                                \
  } while (0)

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Map the linear block id to a logical tile id so that consecutive
// tiles, which share cache lines at their edges, run on the same XCD.
__device__ __forceinline__ int64_t xcd_tile(int64_t lin, int64_t n) {
  const int64_t per = n / 8, rem = n % 8;
  const int64_t x = lin % 8, k = lin / 8;
  return x * per + (x < rem ? x : rem) + k;
}
__device__ __forceinline__ int64_t block_linear() {
  return blockIdx.x + (int64_t)gridDim.x * (blockIdx.y + (int64_t)gridDim.y * blockIdx.z);
}

// explicitly rounded arithmetic (no contraction into multiply-adds): the
// stencil kernels that reproduce the reference's rounding use these
template <typename T> __device__ __forceinline__ T rmul(T a, T b);
template <typename T> __device__ __forceinline__ T radd(T a, T b);
template <typename T> __device__ __forceinline__ T rsub(T a, T b);
template <> __device__ __forceinline__ double rmul(double a, double b) { return __dmul_rn(a, b); }
template <> __device__ __forceinline__ double radd(double a, double b) { return __dadd_rn(a, b); }
template <> __device__ __forceinline__ double rsub(double a, double b) { return __dsub_rn(a, b); }
template <> __device__ __forceinline__ float rmul(float a, float b) { return __fmul_rn(a, b); }
template <> __device__ __forceinline__ float radd(float a, float b) { return __fadd_rn(a, b); }
template <> __device__ __forceinline__ float rsub(float a, float b) { return __fsub_rn(a, b); }

// Periodic forward ghost fill of an (nc, E...) mesh along axis d
// (mesh/pmlib_mesh_map_ghost.f with incl_edges = .FALSE., op = replace).
template <typename R>
__global__ void k_ghost_fill(R *m, int E0, int E1, int E2, int d, int N, int ng, int nc) {
  int E[3] = {E0, E1, E2};
  int o1 = d == 0 ? 1 : 0, o2 = d == 2 ? 1 : 2;
  int64_t st[3] = {nc, nc * (int64_t)E0, nc * (int64_t)E0 * E1};
  int64_t n = (int64_t)E[o1] * E[o2] * nc * ng;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int c = idx % nc;
    int64_t r = idx / nc;
    int t = r % ng;
    r /= ng;
    int a1 = r % E[o1];
    int a2 = r / E[o1];
    int64_t base = c + st[o1] * a1 + st[o2] * a2;
    m[base + st[d] * t] = m[base + st[d] * (N + t)];
    m[base + st[d] * (ng + N + t)] = m[base + st[d] * (ng + t)];
  }
}

// free the workspaces of particles.hip and penalise.hip
void particles_release();
void penalise_release();
void diagnostics_release();
void m2m_release();
void penalisation_solve_release();
// fetches the device address of the allocation-free penalisation scratch on
// the current device (once per device)
int penalise_noalloc_init();

}  // namespace pmgpu
typedef struct pmgpu_poisson_s *pmgpu_poisson_t;
typedef struct pmgpu_geom pmgpu_geom;
namespace pmgpu {
// geometry, precision and stream of a solver handle (poisson.hip)
int poisson_handle_info(pmgpu_poisson_t h, pmgpu_geom *g, int *precision, hipStream_t *stream);

}  // namespace pmgpu
