// Brinkman penalisation step (HIP, gfx950): one fused stencil pass
//   f = chi (us - u),  vort += relax curl(f)   on the interior cells,
// with the force-increment and enstrophy sums reduced deterministically.
//
// Restates one iteration of penalisation_solve
// (vortex_codes/vortex_penalisation/source/penalisation/penalisation_solve.f:
// 96-111 for f, :146-193 / :195-273 for the 2nd / 4th order curl).  f is never
// stored in HBM: a workgroup forms it plane by plane for its tile and halo in
// LDS while it walks along z.  Every floating-point operation is an
// explicitly rounded one, in the order the reference writes it, so the FP64
// result is bit-identical to it.
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/pmgpu.h"
#include "pm_common.h"

namespace pmgpu {

struct BGeom {
  int E[3], N[3], ng[3];
  double xmin[3], dx[3];
};

template <typename T> struct PenArgs {
  BGeom g;
  const T *vel, *chi, *us;  // us: (3, E...) field or null (then us0..2)
  T us0, us1, us2;
  T relax, cx, cy, cz;       // cx = 1 / (2 dx)
  T *vort;
  double *partial;           // 5 doubles per block, or null (no sums)
};

template <typename T>
__device__ __forceinline__ T f_at(const PenArgs<T> &a, int64_t q, int c) {
  T u = a.vel[3 * q + c];
  T s = a.us ? a.us[3 * q + c] : (c == 0 ? a.us0 : (c == 1 ? a.us1 : a.us2));
  return rmul(a.chi[q], rsub(s, u));
}

// central difference from the values at offsets -2, -1, +1, +2 (FD = 4) or
// -1, +1 (FD = 2; m2 and p2 unused)
template <typename T, int FD>
__device__ __forceinline__ T cdiff_v(T m2, T m1, T p1, T p2, T cc) {
  if (FD == 4) {
    const T c1 = T(1.0 / 6.0), c2 = T(4.0 / 3.0);
    T t = rsub(rmul(c1, m2), rmul(c2, m1));
    t = radd(t, rmul(c2, p1));
    t = rsub(t, rmul(c1, p2));
    return rmul(t, cc);
  } else {
    return rmul(radd(-m1, p1), cc);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

constexpr int kTx = 64, kTy = 4;
constexpr int kZc = 16;   // planes per workgroup

// A workgroup of kTx x kTy threads owns an (i, j) tile of interior cells and
// walks the planes of its z chunk (kZc planes).  Per plane it forms f once
// for the tile and an H = fd/2 halo in x and y into LDS; the x and y
// differences come from there, the z differences from a per-thread queue of
// the cell's f0 and f1 over the 2H + 1 planes around it, so a plane's
// vorticity is finished H planes after it was formed.  Per cell the operations
// and their order are those of the reference's loop, so the FP64 result is
// bit-identical to it; the sums are formed per (i tile, j tile, plane) and
// reduced in a fixed order.
template <typename T, int FD>
__global__ void __launch_bounds__(kTx *kTy) k_penalise(PenArgs<T> a) {
  constexpr int H = FD / 2, TW = kTx + 2 * H, TH = kTy + 2 * H, Q = 2 * H + 1;
  __shared__ T fl[3][TH][TW];
  __shared__ double red[kTy][5];
  const BGeom &g = a.g;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * kTx + tx;
  // XCD-aware tile order: the workgroups of one XCD take consecutive tiles
  // (i fastest, then j, then the z chunk), whose halo rows they share in L2
  const int64_t lt = xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y * gridDim.z);
  const int bx = (int)(lt % gridDim.x), by = (int)((lt / gridDim.x) % gridDim.y),
            bz = (int)(lt / ((int64_t)gridDim.x * gridDim.y));
  const int i0 = bx * kTx + 1, j0 = by * kTy + 1;   // interior indices are 1..N
  const int i = i0 + tx, j = j0 + ty;
  const int k0 = bz * kZc + 1, k1 = min(k0 + kZc - 1, g.N[2]);
  const bool own = i <= g.N[0] && j <= g.N[1];
  const int64_t sy = g.E[0], sz = (int64_t)g.E[0] * g.E[1];
  auto cell = [&](int ii, int jj, int kk) {
    return (int64_t)(ii - 1 + g.ng[0]) + sy * (jj - 1 + g.ng[1]) + sz * (kk - 1 + g.ng[2]);
  };
  T q0[Q], q1[Q];        // f0, f1 of the own cell, oldest plane first
  T pend[H + 1][4];      // dvdx, dwdx, dudy, dwdy of the planes not yet finished
#pragma unroll
  for (int t = 0; t < Q; ++t) q0[t] = q1[t] = T(0);
#pragma unroll
  for (int t = 0; t <= H; ++t)
#pragma unroll
    for (int c = 0; c < 4; ++c) pend[t][c] = T(0);
  for (int kk = k0 - H; kk <= k1 + H; ++kk) {
    // f of the tile and its halo at plane kk (cells no owned cell reaches are
    // skipped: they may lie past the ghost layers)
    for (int idx = tid; idx < TW * TH; idx += kTx * kTy) {
      const int r = idx / TW, c = idx - r * TW;
      const int ii = i0 - H + c, jj = j0 - H + r;
      if (ii <= g.N[0] + H && jj <= g.N[1] + H) {
        const int64_t q = cell(ii, jj, kk);
#pragma unroll
        for (int m = 0; m < 3; ++m) fl[m][r][c] = f_at(a, q, m);
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t + 1 < Q; ++t) {
      q0[t] = q0[t + 1];
      q1[t] = q1[t + 1];
    }
    q0[Q - 1] = fl[0][ty + H][tx + H];
    q1[Q - 1] = fl[1][ty + H][tx + H];
    if (kk >= k0 && kk <= k1) {
      const int r = ty + H, c = tx + H;
      const T x2 = H == 2 ? fl[1][r][c - 2] : T(0), x_2 = H == 2 ? fl[1][r][c + 2] : T(0);
      const T z2 = H == 2 ? fl[2][r][c - 2] : T(0), z_2 = H == 2 ? fl[2][r][c + 2] : T(0);
      pend[H][0] = cdiff_v<T, FD>(x2, fl[1][r][c - 1], fl[1][r][c + 1], x_2, a.cx);   // dv/dx
      pend[H][1] = cdiff_v<T, FD>(z2, fl[2][r][c - 1], fl[2][r][c + 1], z_2, a.cx);   // dw/dx
      const T y2 = H == 2 ? fl[0][r - 2][c] : T(0), y_2 = H == 2 ? fl[0][r + 2][c] : T(0);
      const T w2 = H == 2 ? fl[2][r - 2][c] : T(0), w_2 = H == 2 ? fl[2][r + 2][c] : T(0);
      pend[H][2] = cdiff_v<T, FD>(y2, fl[0][r - 1][c], fl[0][r + 1][c], y_2, a.cy);   // du/dy
      pend[H][3] = cdiff_v<T, FD>(w2, fl[2][r - 1][c], fl[2][r + 1][c], w_2, a.cy);   // dw/dy
    }
    __syncthreads();   // the tile is read; the next plane may overwrite it
    const int kf = kk - H;   // plane finished now
    if (kf >= k0) {
      double s[5] = {0, 0, 0, 0, 0};
      if (own) {
        const T dvdx = pend[0][0], dwdx = pend[0][1], dudy = pend[0][2], dwdy = pend[0][3];
        const T dudz = cdiff_v<T, FD>(q0[0], q0[H - 1], q0[H + 1], q0[Q - 1], a.cz);
        const T dvdz = cdiff_v<T, FD>(q1[0], q1[H - 1], q1[H + 1], q1[Q - 1], a.cz);
        T w0 = rmul(a.relax, rsub(dwdy, dvdz));
        T w1 = rmul(a.relax, rsub(dudz, dwdx));
        T w2 = rmul(a.relax, rsub(dvdx, dudy));
        T *v = a.vort + 3 * cell(i, j, kf);
        T v0 = radd(v[0], w0), v1 = radd(v[1], w1), v2 = radd(v[2], w2);
        v[0] = v0;
        v[1] = v1;
        v[2] = v2;
        if (a.partial) {
          const double px = __dadd_rn(g.xmin[0], __dmul_rn(__dadd_rn((double)(i - 1), 0.5), g.dx[0]));
          const double py = __dadd_rn(g.xmin[1], __dmul_rn(__dadd_rn((double)(j - 1), 0.5), g.dx[1]));
          const double pz = __dadd_rn(g.xmin[2], __dmul_rn(__dadd_rn((double)(kf - 1), 0.5), g.dx[2]));
          const double d0 = w0, d1 = w1, d2 = w2, e0 = v0, e1 = v1, e2 = v2;
          s[0] = __dsub_rn(__dmul_rn(py, d2), __dmul_rn(pz, d1));
          s[1] = __dsub_rn(__dmul_rn(pz, d0), __dmul_rn(px, d2));
          s[2] = __dsub_rn(__dmul_rn(px, d1), __dmul_rn(py, d0));
          s[3] = __dadd_rn(__dadd_rn(__dmul_rn(d0, d0), __dmul_rn(d1, d1)), __dmul_rn(d2, d2));
          s[4] = __dadd_rn(__dadd_rn(__dmul_rn(e0, e0), __dmul_rn(e1, e1)), __dmul_rn(e2, e2));
        }
      }
      if (a.partial) {
        for (int m = 0; m < 5; ++m) {
          double t = wave_sum(s[m]);
          if (tx == 0) red[ty][m] = t;
        }
        __syncthreads();
        if (ty == 0 && tx < 5) {
          double t = 0.0;
          for (int w = 0; w < kTy; ++w) t += red[w][tx];
          // the block index of the (i tile, j tile, plane) partial
          const int64_t b = bx + (int64_t)gridDim.x * (by + (int64_t)gridDim.y * (kf - 1));
          a.partial[5 * b + tx] = t;
        }
        __syncthreads();
      }
    }
#pragma unroll
    for (int t = 0; t < H; ++t)
#pragma unroll
      for (int c = 0; c < 4; ++c) pend[t][c] = pend[t + 1][c];
  }
}

// fixed-order reduction of the per-block sums (one block): thread t adds the
// blocks t, t + 256, ... in that order, then a fixed tree; the loads of 8
// blocks are issued ahead of their adds
__global__ void __launch_bounds__(256) k_penalise_sums(const double *partial, int64_t nb, double *out) {
  __shared__ double sh[5][256];
  double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int64_t b = threadIdx.x;
  for (; b + 7 * 256 < nb; b += 8 * 256) {
    double v[8][5];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int m = 0; m < 5; ++m) v[u][m] = partial[5 * (b + u * 256) + m];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int m = 0; m < 5; ++m) t[m] += v[u][m];
  }
  for (; b < nb; b += 256)
#pragma unroll
    for (int m = 0; m < 5; ++m) t[m] += partial[5 * b + m];
#pragma unroll
  for (int m = 0; m < 5; ++m) sh[m][threadIdx.x] = t[m];
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int m = 0; m < 5; ++m) sh[m][threadIdx.x] += sh[m][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < 5) out[threadIdx.x] = sh[threadIdx.x][0];
}

constexpr int kMaxDevices = 64;

// workgroups of k_penalise; the partial sums are one per (i tile, j tile,
// plane), partial_blocks of them
static dim3 penalise_grid(const pmgpu_geom *geom) {
  return dim3((geom->ncell[0] + kTx - 1) / kTx, (geom->ncell[1] + kTy - 1) / kTy, (geom->ncell[2] + kZc - 1) / kZc);
}
static int64_t partial_blocks(const pmgpu_geom *geom) {
  const dim3 g = penalise_grid(geom);
  return (int64_t)g.x * g.y * geom->ncell[2];
}

struct PenWS {
  double *partial = nullptr;
  int64_t cap = 0;
  double *sums = nullptr;
};
static PenWS *g_pen = nullptr;

void penalise_release() {
  if (g_pen) {
    if (g_pen->partial) (void)hipFree(g_pen->partial);
    if (g_pen->sums) (void)hipFree(g_pen->sums);
    delete g_pen;
    g_pen = nullptr;
  }
}

// Builds the arguments and launches the penalisation kernel and, when
// partial is non-null, the reduction of its per-block partial sums (5
// doubles per block, at most cap blocks) into sums[0..4].  Stream-ordered,
// no allocation, no host synchronisation.
template <typename T>
static int penalise_run(const pmgpu_geom *geom, const void *vel, const void *chi, const void *us_field,
                        const double *us_const, double relax, int fd_order, void *vort, double *partial,
                        int64_t cap, double *sums, hipStream_t st) {
  PM_CHECK(geom && vel && chi && vort, "pmgpu_penalise: null argument");
  PM_CHECK(fd_order == 2 || fd_order == 4, "pmgpu_penalise: fd_order must be 2 or 4");
  PenArgs<T> a;
  int32_t ext[3];
  if (pmgpu_mesh_extent(geom, ext)) return -1;
  for (int d = 0; d < 3; ++d) {
    a.g.E[d] = ext[d];
    a.g.N[d] = geom->ncell[d];
    a.g.ng[d] = (ext[d] - geom->ncell[d]) / 2;
    a.g.xmin[d] = geom->xmin[d];
    a.g.dx[d] = geom->dx[d];
    PM_CHECK(a.g.ng[d] >= fd_order / 2, "pmgpu_penalise: the stencil needs fd_order/2 ghost layers");
  }
  a.vel = (const T *)vel;
  a.chi = (const T *)chi;
  a.us = (const T *)us_field;
  a.us0 = T(us_const ? us_const[0] : 0.0);
  a.us1 = T(us_const ? us_const[1] : 0.0);
  a.us2 = T(us_const ? us_const[2] : 0.0);
  a.relax = T(relax);
  // c_1_2dx2 = 1/(2 dx) is formed in the working precision from dx
  a.cx = T(1.0) / (T(2.0) * T(geom->dx[0]));
  a.cy = T(1.0) / (T(2.0) * T(geom->dx[1]));
  a.cz = T(1.0) / (T(2.0) * T(geom->dx[2]));
  a.vort = (T *)vort;
  PM_CHECK(geom->ncell[2] < 65536 && geom->ncell[1] < (1 << 20), "pmgpu_penalise: mesh too large");
  const dim3 grid = penalise_grid(geom);
  const int64_t nb = partial_blocks(geom);
  PM_CHECK(!partial || nb <= cap, "pmgpu_penalise: more blocks than the partial-sum buffer holds");
  a.partial = partial;
  if (fd_order == 4)
    hipLaunchKernelGGL((k_penalise<T, 4>), grid, dim3(kTx, kTy), 0, st, a);
  else
    hipLaunchKernelGGL((k_penalise<T, 2>), grid, dim3(kTx, kTy), 0, st, a);
  PM_HIP(hipGetLastError());
  if (partial) {
    hipLaunchKernelGGL(k_penalise_sums, dim3(1), dim3(256), 0, st, (const double *)partial, nb, sums);
    PM_HIP(hipGetLastError());
  }
  return 0;
}

// The allocating form: the partial sums live in a library buffer grown on
// demand, and the totals are copied to the host (one stream synchronisation).
template <typename T>
static int penalise_launch(const pmgpu_geom *geom, const void *vel, const void *chi, const void *us_field,
                           const double *us_const, double relax, int fd_order, void *vort, double *diag,
                           hipStream_t st) {
  PM_CHECK(geom, "pmgpu_penalise: null argument");
  double *partial = nullptr;
  int64_t cap = 0;
  if (diag) {
    const int64_t nb = partial_blocks(geom);
    if (!g_pen) g_pen = new PenWS;
    if (g_pen->cap < nb) {
      if (g_pen->partial) (void)hipFree(g_pen->partial);
      g_pen->partial = nullptr;
      g_pen->cap = 0;
      PM_HIP(hipMalloc(&g_pen->partial, sizeof(double) * 5 * nb));
      g_pen->cap = nb;
    }
    if (!g_pen->sums) PM_HIP(hipMalloc(&g_pen->sums, sizeof(double) * 8));
    partial = g_pen->partial;
    cap = g_pen->cap;
  }
  if (penalise_run<T>(geom, vel, chi, us_field, us_const, relax, fd_order, vort, partial, cap,
                      diag ? g_pen->sums : nullptr, st))
    return -1;
  if (diag) {
    double h[5];
    PM_HIP(hipMemcpyAsync(h, g_pen->sums, sizeof(h), hipMemcpyDeviceToHost, st));
    PM_HIP(hipStreamSynchronize(st));
    std::memcpy(diag, h, sizeof(h));
  }
  return 0;
}

// Scratch of the allocation-free form: the per-block partial sums of at
// most kNoAllocBlocks blocks (10 MiB), one copy per GPU, held until the
// process ends.  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol): it uses the device address
// that hipGetSymbolAddress returns, fetched once per device.
constexpr int64_t kNoAllocBlocks = int64_t(1) << 18;
static __device__ __align__(256) double scratch[5 * kNoAllocBlocks];

static double *scratch_address() {
  static std::mutex mu;
  static double *addr[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("pmgpu_penalise_device_noalloc: no current HIP device");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("pmgpu_penalise_device_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
      return nullptr;
    }
    addr[dev] = static_cast<double *>(p);
  }
  return addr[dev];
}

int penalise_noalloc_init() { return scratch_address() ? 0 : -1; }

// ---- iterative penalisation (penalisation_solve) -------------------------
// vp = us - u on every cell, us the solid-velocity field or the constants
// (penalisation_solve.f:88-95)
template <typename T>
__global__ void k_pen_velpen(const T *__restrict__ us, T us0, T us1, T us2, const T *__restrict__ u,
                             T *__restrict__ vp, int64_t ncell) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < 3 * ncell;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % 3);
    const T s = us ? us[i] : (c == 0 ? us0 : (c == 1 ? us1 : us2));
    vp[i] = s - u[i];
  }
}
template <typename T> __global__ void k_pen_add(T *__restrict__ a, const T *__restrict__ b, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = a[i] + b[i];
}
// free-stream velocity on every cell (penalisation_solve.f:405-414); quirk:
// the reference adds vel_inf(3) to the second component, as written there
template <typename T>
__global__ void k_pen_free_stream(T *__restrict__ vel, T v0, T v1, T v2, int64_t ncell, int quirk) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < ncell; i += (int64_t)gridDim.x * blockDim.x) {
    T *u = vel + 3 * i;
    u[0] = u[0] + v0;
    u[1] = u[1] + v1;
    if (quirk)
      u[1] = u[1] + v2;
    else
      u[2] = u[2] + v2;
  }
}

struct PenSolveWS {
  void *velpen = nullptr, *upen = nullptr, *vpen = nullptr;
  size_t bytes = 0;
};
static PenSolveWS *g_psolve = nullptr;

void penalisation_solve_release() {
  if (g_psolve) {
    void *b[] = {g_psolve->velpen, g_psolve->upen, g_psolve->vpen};
    for (void *p : b)
      if (p) (void)hipFree(p);
    delete g_psolve;
    g_psolve = nullptr;
  }
}

template <typename T>
static int penalisation_solve_impl(pmgpu_poisson_t h, const pmgpu_geom &g, hipStream_t st,
                                   const pmgpu_penal_params &pp, const void *chi, const void *solid_vel,
                                   const double *us_const, void *vel, void *vort, double *force_out,
                                   int32_t *npen_out, double *residual_out) {
  int32_t E[3];
  if (pmgpu_mesh_extent(&g, E)) return -1;
  const int64_t ncell = (int64_t)E[0] * E[1] * E[2];
  const size_t bytes = sizeof(T) * 3 * (size_t)ncell;
  if (!g_psolve) g_psolve = new PenSolveWS;
  if (g_psolve->bytes < bytes) {
    penalisation_solve_release();
    g_psolve = new PenSolveWS;
    PM_HIP(hipMalloc(&g_psolve->velpen, bytes));
    PM_HIP(hipMalloc(&g_psolve->upen, bytes));
    PM_HIP(hipMalloc(&g_psolve->vpen, bytes));
    g_psolve->bytes = bytes;
  }
  T *velpen = (T *)g_psolve->velpen, *upen = (T *)g_psolve->upen, *vpen = (T *)g_psolve->vpen;
  const dim3 eg(2048), eb(256);
  // vel_pen = solid_vel - vel (the velocity mapped to the penalisation mesh)
  hipLaunchKernelGGL(k_pen_velpen<T>, eg, eb, 0, st, (const T *)solid_vel, T(us_const ? us_const[0] : 0.0),
                     T(us_const ? us_const[1] : 0.0), T(us_const ? us_const[2] : 0.0), (const T *)vel, velpen,
                     ncell);
  PM_HIP(hipGetLastError());
  // the penalisation vorticity starts at zero (cleared after every call in
  // the reference), and so does the velocity it induces
  PM_HIP(hipMemsetAsync(vpen, 0, bytes, st));
  PM_HIP(hipMemsetAsync(upen, 0, bytes, st));
  const double c_1_dt = 1.0 / pp.dtime;
  double force[3] = {0.0, 0.0, 0.0}, residual = 0.0;
  int npen = 0;
  while ((residual > pp.tolerance && npen < pp.max_iter) || npen == 0) {
    ++npen;
    // vort_pen += relax curl(chi (vel_pen - u_pen)) and the force sums
    double d[5];
    if (penalise_launch<T>(&g, upen, chi, velpen, nullptr, pp.relaxation, pp.fd_order, vpen, d, st)) return -1;
    // the velocity of the penalisation vorticity (ghost cells included)
    if (pmgpu_poisson_solve_device(h, vpen, upen, nullptr, nullptr, 0, 0, 0)) return -1;
    double dforce[3];
    for (int c = 0; c < 3; ++c) {
      dforce[c] = -d[c] * g.dx[0] * g.dx[1] * g.dx[2] * c_1_dt;
      force[c] = force[c] + dforce[c];
    }
    residual = std::sqrt(dforce[0] * dforce[0] + dforce[1] * dforce[1] + dforce[2] * dforce[2]) /
               std::sqrt(force[0] * force[0] + force[1] * force[1] + force[2] * force[2]);
  }
  // add the penalisation vorticity to the mesh vorticity (its ghost cells
  // are zero), then the periodic ghost map of the vorticity
  hipLaunchKernelGGL(k_pen_add<T>, eg, eb, 0, st, (T *)vort, (const T *)vpen, 3 * ncell);
  PM_HIP(hipGetLastError());
  for (int d = 0; d < 3; ++d)
    if (g.bound_cond[d] == 1) {
      hipLaunchKernelGGL(k_ghost_fill<T>, dim3(1024), dim3(256), 0, st, (T *)vort, E[0], E[1], E[2], d, g.ncell[d],
                         (E[d] - g.ncell[d]) / 2, 3);
      PM_HIP(hipGetLastError());
    }
  // the velocity on the whole mesh, with the reprojected vorticity if asked
  if (pmgpu_poisson_solve_device(h, vort, vel, pp.reproject ? vort : nullptr, nullptr, 0, pp.reproject ? 1 : 0, 0))
    return -1;
  hipLaunchKernelGGL(k_pen_free_stream<T>, eg, eb, 0, st, (T *)vel, T(pp.vel_inf[0]), T(pp.vel_inf[1]),
                     T(pp.vel_inf[2]), ncell, pp.fix_free_stream ? 0 : 1);
  PM_HIP(hipGetLastError());
  PM_HIP(hipStreamSynchronize(st));
  if (force_out)
    for (int c = 0; c < 3; ++c) force_out[c] = force[c];
  if (npen_out) *npen_out = npen;
  if (residual_out) *residual_out = residual;
  return 0;
}

}  // namespace pmgpu

using namespace pmgpu;

extern "C" {

int pmgpu_penalise(const pmgpu_geom *g, const double *vel, const double *chi, const double *solid_vel,
                   const double us_const[3], double relax, int32_t fd_order, double *vort, double diag[5]) {
  PM_CHECK(g && vel && chi && vort, "pmgpu_penalise: null argument");
  int32_t e[3];
  if (pmgpu_mesh_extent(g, e)) return -1;
  const size_t n = (size_t)e[0] * e[1] * e[2];
  double *dv = nullptr, *dc = nullptr, *ds = nullptr, *dw = nullptr;
  int rc = -1;
  if (hipMalloc(&dv, 24 * n) == hipSuccess && hipMalloc(&dc, 8 * n) == hipSuccess &&
      hipMalloc(&dw, 24 * n) == hipSuccess && (!solid_vel || hipMalloc(&ds, 24 * n) == hipSuccess) &&
      hipMemcpy(dv, vel, 24 * n, hipMemcpyHostToDevice) == hipSuccess &&
      hipMemcpy(dc, chi, 8 * n, hipMemcpyHostToDevice) == hipSuccess &&
      hipMemcpy(dw, vort, 24 * n, hipMemcpyHostToDevice) == hipSuccess &&
      (!solid_vel || hipMemcpy(ds, solid_vel, 24 * n, hipMemcpyHostToDevice) == hipSuccess)) {
    rc = penalise_launch<double>(g, dv, dc, ds, us_const, relax, fd_order, dw, diag, nullptr);
    if (rc == 0 && hipMemcpy(vort, dw, 24 * n, hipMemcpyDeviceToHost) != hipSuccess) rc = -1;
  } else {
    set_error("pmgpu_penalise: device allocation or copy failed");
  }
  if (dv) (void)hipFree(dv);
  if (dc) (void)hipFree(dc);
  if (ds) (void)hipFree(ds);
  if (dw) (void)hipFree(dw);
  return rc;
}

int pmgpu_penalise_device(const pmgpu_geom *g, int32_t precision, const void *d_vel, const void *d_chi,
                          const void *d_solid_vel, const double us_const[3], double relax, int32_t fd_order,
                          void *d_vort, double *diag, void *stream) {
  if (precision == PMGPU_FP32)
    return penalise_launch<float>(g, d_vel, d_chi, d_solid_vel, us_const, relax, fd_order, d_vort, diag,
                                  (hipStream_t)stream);
  PM_CHECK(precision == PMGPU_FP64, "pmgpu_penalise_device: precision must be PMGPU_FP64 or PMGPU_FP32");
  return penalise_launch<double>(g, d_vel, d_chi, d_solid_vel, us_const, relax, fd_order, d_vort, diag,
                                 (hipStream_t)stream);
}

int pmgpu_penalisation_solve_device(pmgpu_poisson_t h, const pmgpu_penal_params *pp, const void *d_chi,
                                    const void *d_solid_vel, const double us_const[3], void *d_vel, void *d_vort,
                                    double force[3], int32_t *npen, double *residual) {
  PM_CHECK(h && pp && d_chi && d_vel && d_vort, "pmgpu_penalisation_solve_device: null argument");
  PM_CHECK(pp->dtime > 0.0, "pmgpu_penalisation_solve_device: dtime must be positive");
  PM_CHECK(pp->max_iter >= 1, "pmgpu_penalisation_solve_device: penal_max must be at least 1");
  pmgpu_geom g;
  int precision = 0;
  hipStream_t st = nullptr;
  if (poisson_handle_info(h, &g, &precision, &st)) return -1;
  if (precision == PMGPU_FP32)
    return penalisation_solve_impl<float>(h, g, st, *pp, d_chi, d_solid_vel, us_const, d_vel, d_vort, force, npen,
                                          residual);
  return penalisation_solve_impl<double>(h, g, st, *pp, d_chi, d_solid_vel, us_const, d_vel, d_vort, force, npen,
                                         residual);
}

int pmgpu_penalisation_solve(pmgpu_poisson_t h, const pmgpu_penal_params *pp, const double *chi,
                             const double *solid_vel, const double us_const[3], double *vel, double *vort,
                             double force[3], int32_t *npen, double *residual) {
  PM_CHECK(h && pp && chi && vel && vort, "pmgpu_penalisation_solve: null argument");
  pmgpu_geom g;
  int precision = 0;
  hipStream_t st = nullptr;
  if (poisson_handle_info(h, &g, &precision, &st)) return -1;
  PM_CHECK(precision == PMGPU_FP64, "pmgpu_penalisation_solve: host meshes need an FP64 solver handle");
  int32_t E[3];
  if (pmgpu_mesh_extent(&g, E)) return -1;
  const size_t n = (size_t)E[0] * E[1] * E[2];
  double *dc = nullptr, *ds = nullptr, *dv = nullptr, *dw = nullptr;
  int rc = -1;
  if (hipMalloc(&dc, 8 * n) == hipSuccess && hipMalloc(&dv, 24 * n) == hipSuccess &&
      hipMalloc(&dw, 24 * n) == hipSuccess && (!solid_vel || hipMalloc(&ds, 24 * n) == hipSuccess) &&
      hipMemcpy(dc, chi, 8 * n, hipMemcpyHostToDevice) == hipSuccess &&
      hipMemcpy(dv, vel, 24 * n, hipMemcpyHostToDevice) == hipSuccess &&
      hipMemcpy(dw, vort, 24 * n, hipMemcpyHostToDevice) == hipSuccess &&
      (!solid_vel || hipMemcpy(ds, solid_vel, 24 * n, hipMemcpyHostToDevice) == hipSuccess)) {
    rc = pmgpu_penalisation_solve_device(h, pp, dc, ds, us_const, dv, dw, force, npen, residual);
    if (rc == 0 && (hipMemcpy(vel, dv, 24 * n, hipMemcpyDeviceToHost) != hipSuccess ||
                    hipMemcpy(vort, dw, 24 * n, hipMemcpyDeviceToHost) != hipSuccess)) {
      set_error("pmgpu_penalisation_solve: copy back failed");
      rc = -1;
    }
  } else {
    set_error("pmgpu_penalisation_solve: device allocation or copy failed");
  }
  void *bufs[] = {dc, ds, dv, dw};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  return rc;
}

int pmgpu_penalise_device_noalloc(const pmgpu_geom *g, int32_t precision, const void *d_vel, const void *d_chi,
                                  const void *d_solid_vel, const double us_const[3], double relax, int32_t fd_order,
                                  void *d_vort, double *d_diag, void *stream) {
  PM_CHECK(precision == PMGPU_FP64 || precision == PMGPU_FP32,
           "pmgpu_penalise_device_noalloc: precision must be PMGPU_FP64 or PMGPU_FP32");
  PM_CHECK(g, "pmgpu_penalise_device_noalloc: null argument");
  double *partial = nullptr;
  if (d_diag) {
    const int64_t nb = partial_blocks(g);
    PM_CHECK(nb <= kNoAllocBlocks, "pmgpu_penalise_device_noalloc: the mesh needs " + std::to_string(nb) +
                                       " partial-sum blocks, more than the " + std::to_string(kNoAllocBlocks) +
                                       " its fixed scratch holds (use pmgpu_penalise_device)");
    partial = scratch_address();
    if (!partial) return -1;
  }
  if (precision == PMGPU_FP32)
    return penalise_run<float>(g, d_vel, d_chi, d_solid_vel, us_const, relax, fd_order, d_vort, partial,
                               kNoAllocBlocks, d_diag, (hipStream_t)stream);
  return penalise_run<double>(g, d_vel, d_chi, d_solid_vel, us_const, relax, fd_order, d_vort, partial,
                              kNoAllocBlocks, d_diag, (hipStream_t)stream);
}

}  // extern "C"
