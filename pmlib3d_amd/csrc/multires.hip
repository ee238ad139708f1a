// Multiresolution coupling on one rank (HIP, gfx950): the parent topology of
// a refinement patch and the mesh-to-mesh interpolation between a patch and
// its parent, both directions, M'4 and M6*
// (interpolation/pmlib_interp_mesh_mesh.f:5-2078,
// topology/pmlib_topology_parent.f:5-100).
//
// Patch -> parent (the source is finer, reldx = dx_from / dx_to < 1) is a
// deposit in the reference: every source cell, ghosts included, spreads
// reldx W_x W_y W_z of its value over the target stencil; parent -> patch
// gathers W_x W_y W_z-weighted source values per target cell.  Both are
// tensor products of per-axis operators, so the default form runs each as
// three 1D gather passes over per-axis tap tables (HBM-streaming, no
// atomics).  The bit-identical forms (PMGPU_M2M_EXACT=1) keep the
// reference's term order: the deposit as a gather in which each target cell
// sums the terms of the source cells whose stencils cover it in the
// reference's source-cell order (k, j, i ascending), a block of 64 target
// cells of one (y, z) row staging the per-axis bases and weights it needs in
// LDS; the interpolation per target cell in the reference's term order.
// Compiled without multiply-add contraction.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "../../include/pmgpu.h"
// (the particle kernels of the header that this file does not launch)
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wunused-function"
#include "particles_kernels.h"
#pragma clang diagnostic pop

namespace pmgpu {
namespace {

struct MGeom {
  int ncell[3], ng[3], E[3], bc[3];
  double xmin[3], xmax[3], dx[3], c1dx[3];
};

int make_mgeom(const pmgpu_geom *g, MGeom &m, const char *who) {
  PM_CHECK(g->interp_order == 3 || g->interp_order == 4,
           std::string(who) + ": interp_order must be 3 (M'4) or 4 (M6*)");
  for (int d = 0; d < 3; ++d) {
    PM_CHECK(g->ncell[d] > 0 && g->dx[d] > 0 && (g->bound_cond[d] == 0 || g->bound_cond[d] == 1),
             std::string(who) + ": bad geometry");
    m.ncell[d] = g->ncell[d];
    m.bc[d] = g->bound_cond[d];
    m.ng[d] = g->bound_cond[d] == 0 ? g->nghost : (g->interp_order == 4 ? 3 : 2);
    PM_CHECK(m.ng[d] >= 0, std::string(who) + ": negative ghost count");
    m.E[d] = m.ncell[d] + 2 * m.ng[d];
    m.xmin[d] = g->xmin[d];
    m.xmax[d] = g->xmax[d];
    m.dx[d] = g->dx[d];
    m.c1dx[d] = 1.0 / g->dx[d];
  }
  return 0;
}

constexpr int kTX = 64;       // target cells per deposit block along x
constexpr int kMaxRatio = 8;  // finest-to-coarsest cell ratio the deposit stages
// source candidates staged per block: x covers kTX + NS target bases at up to
// kMaxRatio source cells each, y and z one row's NS bases, plus margins
template <int NS> constexpr int cap_x() { return (kTX + NS + 2) * kMaxRatio + 8; }
template <int NS> constexpr int cap_yz() { return (NS + 2) * kMaxRatio + 8; }

// Conservative range of source indices (1-ng .. N+ng numbering) whose base
// index NINT((p - xmin_t) / dx_t) may fall in [blo, bhi]; the exact test is
// made on the staged bases.
__device__ __forceinline__ void src_range(const MGeom &f, const MGeom &t, int d, int blo, int bhi, int &ilo,
                                          int &ihi) {
  const double a = (t.xmin[d] - f.xmin[d] + ((double)blo - 1.0) * t.dx[d]) / f.dx[d];
  const double b = (t.xmin[d] - f.xmin[d] + ((double)bhi + 1.0) * t.dx[d]) / f.dx[d];
  ilo = (int)floor(a) - 1;
  ihi = (int)ceil(b) + 2;
  if (ilo < 1 - f.ng[d]) ilo = 1 - f.ng[d];
  if (ihi > f.ncell[d] + f.ng[d]) ihi = f.ncell[d] + f.ng[d];
}

// Base index, live flag and the NS weights reldx W_s of source index i on
// axis d (:93-158 (M'4), :520-640 (M6*)): the cell centre x_i, its target
// base b = NINT((x_i - xmin_t) / dx_t), the normalised offset and the kernel
// weights, rounded step by step as the reference evaluates them.  On a
// periodic target axis a source cell outside [xmin_t, xmax_t] deposits
// nothing (:174-195).
template <int NS>
__device__ __forceinline__ void stage_src(const MGeom &f, const MGeom &t, int d, int i, double reldx, int &b,
                                          unsigned char &live, double *w) {
  const double p = __dadd_rn(f.xmin[d], __dmul_rn(__dadd_rn((double)(i - 1), 0.5), f.dx[d]));
  b = cell_index(p, t.xmin[d], t.c1dx[d]);
  live = !(t.bc[d] != 0 && (p < t.xmin[d] || p > t.xmax[d]));
  const double dd = stencil_offset(p, t.xmin[d], t.dx[d], b, t.c1dx[d]);
#pragma unroll
  for (int s = 0; s < NS; ++s) w[s] = __dmul_rn(reldx, weight(NS == 6, s, dd));
}

// first staged entry with base >= v (bases are non-decreasing in i)
__device__ __forceinline__ int lower_bound_b(const int *bs, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (bs[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

struct DepArgs {
  MGeom f, t;
  double reldx;
  int clear;
};

// Patch -> parent: grid (ceil(Ex_t / kTX), Ey_t, Ez_t), kTX threads.
template <typename T, int NS>
__global__ void __launch_bounds__(kTX) k_m2m_deposit(const T *__restrict__ from, T *__restrict__ to, DepArgs a) {
  constexpr int off = NS == 6 ? 2 : 1;
  constexpr int CX = cap_x<NS>(), CY = cap_yz<NS>();
  __shared__ int xb[CX], yb[CY], zb[CY];
  __shared__ unsigned char xl[CX], yl[CY], zl[CY];
  __shared__ double xw[CX][NS], yw[CY][NS], zw[CY][NS];
  const MGeom &f = a.f, &t = a.t;
  // target indices in the reference's numbering (1 - ng .. N + ng)
  const int tx0 = (int)blockIdx.x * kTX + 1 - t.ng[0];
  const int txe = min(tx0 + kTX - 1, t.ncell[0] + t.ng[0]);
  const int ty = (int)blockIdx.y + 1 - t.ng[1], tz = (int)blockIdx.z + 1 - t.ng[2];
  // a source cell of base b covers target indices b - off .. b - off + NS - 1
  int ix0, ix1, iy0, iy1, iz0, iz1;
  src_range(f, t, 0, tx0 + off - NS + 1, txe + off, ix0, ix1);
  src_range(f, t, 1, ty + off - NS + 1, ty + off, iy0, iy1);
  src_range(f, t, 2, tz + off - NS + 1, tz + off, iz0, iz1);
  // (the host admits cell ratios up to kMaxRatio, for which the ranges fit
  // the staging arrays; the bounds are enforced here all the same)
  const int nx = min(ix1 - ix0 + 1, CX), ny = min(iy1 - iy0 + 1, CY), nz = min(iz1 - iz0 + 1, CY);
  for (int q = threadIdx.x; q < nx; q += blockDim.x) stage_src<NS>(f, t, 0, ix0 + q, a.reldx, xb[q], xl[q], xw[q]);
  for (int q = threadIdx.x; q < ny; q += blockDim.x) stage_src<NS>(f, t, 1, iy0 + q, a.reldx, yb[q], yl[q], yw[q]);
  for (int q = threadIdx.x; q < nz; q += blockDim.x) stage_src<NS>(f, t, 2, iz0 + q, a.reldx, zb[q], zl[q], zw[q]);
  __syncthreads();
  const int tx = tx0 + (int)threadIdx.x;
  if (tx > txe) return;
  // the staged entries whose stencil covers this target cell, per axis
  const int qx0 = lower_bound_b(xb, nx, tx + off - NS + 1), qx1 = lower_bound_b(xb, nx, tx + off + 1);
  const int qy0 = lower_bound_b(yb, ny, ty + off - NS + 1), qy1 = lower_bound_b(yb, ny, ty + off + 1);
  const int qz0 = lower_bound_b(zb, nz, tz + off - NS + 1), qz1 = lower_bound_b(zb, nz, tz + off + 1);
  const int64_t fsx = 3, fsy = 3 * (int64_t)f.E[0], fsz = fsy * f.E[1];
  const int64_t o = 3 * ((int64_t)(tx - 1 + t.ng[0]) +
                         (int64_t)t.E[0] * ((int64_t)(ty - 1 + t.ng[1]) + (int64_t)t.E[1] * (tz - 1 + t.ng[2])));
  double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
  if (!a.clear) {
    acc0 = (double)to[o];
    acc1 = (double)to[o + 1];
    acc2 = (double)to[o + 2];
  }
  for (int qz = qz0; qz < qz1; ++qz) {
    if (!zl[qz]) continue;
    const double wz = zw[qz][tz - zb[qz] + off];
    const int64_t oz = (int64_t)(iz0 + qz - 1 + f.ng[2]) * fsz;
    for (int qy = qy0; qy < qy1; ++qy) {
      if (!yl[qy]) continue;
      const double wy = yw[qy][ty - yb[qy] + off];
      const int64_t oy = oz + (int64_t)(iy0 + qy - 1 + f.ng[1]) * fsy;
      for (int qx = qx0; qx < qx1; ++qx) {
        if (!xl[qx]) continue;
        const double wx = xw[qx][tx - xb[qx] + off];
        const double ww = __dmul_rn(__dmul_rn(wx, wy), wz);
        const T *m = from + oy + (int64_t)(ix0 + qx - 1 + f.ng[0]) * fsx;
        acc0 = __dadd_rn(acc0, __dmul_rn(ww, (double)m[0]));
        acc1 = __dadd_rn(acc1, __dmul_rn(ww, (double)m[1]));
        acc2 = __dadd_rn(acc2, __dmul_rn(ww, (double)m[2]));
      }
    }
  }
  to[o] = (T)acc0;
  to[o + 1] = (T)acc1;
  to[o + 2] = (T)acc2;
}

// Parent -> patch (:1395-2062): every target cell (interior cells only on
// periodic axes) adds sum ((W_x W_y) W_z) m over the source stencil, the
// terms in i-fastest order, stencil indices outside the source mesh clamped
// to its last layer with weight 0.  clear writes 0 to the cells it skips.
template <typename T, int NS>
__global__ void k_m2m_interp(const T *__restrict__ from, T *__restrict__ to, MGeom f, MGeom t, int clear) {
  constexpr int off = NS == 6 ? 2 : 1;
  const int64_t n = (int64_t)t.E[0] * t.E[1] * t.E[2];
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) {
    const int ii[3] = {(int)(c % t.E[0]) + 1 - t.ng[0], (int)((c / t.E[0]) % t.E[1]) + 1 - t.ng[1],
                       (int)(c / ((int64_t)t.E[0] * t.E[1])) + 1 - t.ng[2]};
    bool skip = false;
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (t.bc[d] != 0 && (ii[d] < 1 || ii[d] > t.ncell[d])) skip = true;
    if (skip) {
      if (clear) to[3 * c] = to[3 * c + 1] = to[3 * c + 2] = (T)0;
      continue;
    }
    double w[3][NS];
    int64_t idx[3][NS];
    const int64_t st[3] = {3, 3 * (int64_t)f.E[0], 3 * (int64_t)f.E[0] * f.E[1]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double p = __dadd_rn(t.xmin[d], __dmul_rn(__dadd_rn((double)(ii[d] - 1), 0.5), t.dx[d]));
      const int b = cell_index(p, f.xmin[d], f.c1dx[d]);
      const double dd = stencil_offset(p, f.xmin[d], f.dx[d], b, f.c1dx[d]);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        int q = b - off + s;
        double ws = weight(NS == 6, s, dd);
        if (q < 1 - f.ng[d]) { ws = 0.0; q = 1 - f.ng[d]; }
        if (q > f.ncell[d] + f.ng[d]) { ws = 0.0; q = f.ncell[d] + f.ng[d]; }
        w[d][s] = ws;
        idx[d][s] = (int64_t)(q - 1 + f.ng[d]) * st[d];
      }
    }
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    if (!clear) {
      acc0 = (double)to[3 * c];
      acc1 = (double)to[3 * c + 1];
      acc2 = (double)to[3 * c + 2];
    }
#pragma unroll
    for (int sz = 0; sz < NS; ++sz)
#pragma unroll
      for (int sy = 0; sy < NS; ++sy)
#pragma unroll
        for (int sx = 0; sx < NS; ++sx) {
          const double ww = __dmul_rn(__dmul_rn(w[0][sx], w[1][sy]), w[2][sz]);
          const T *m = from + idx[0][sx] + idx[1][sy] + idx[2][sz];
          acc0 = __dadd_rn(acc0, __dmul_rn(ww, (double)m[0]));
          acc1 = __dadd_rn(acc1, __dmul_rn(ww, (double)m[1]));
          acc2 = __dadd_rn(acc2, __dmul_rn(ww, (double)m[2]));
        }
    to[3 * c] = (T)acc0;
    to[3 * c + 1] = (T)acc1;
    to[3 * c + 2] = (T)acc2;
  }
}

// ---------------------------------------------------------------------------
// Separable form (the default).  Both directions are tensor products of
// per-axis operators: the deposit weight of a source cell onto a target cell
// is (reldx W_x)(reldx W_y)(reldx W_z), a cell dropped on a periodic axis or
// clamped on a free one has a zero factor on that axis, and the
// interpolation weight is W_x W_y W_z with the same zero factors.  So each
// direction runs as three 1D passes, each a gather along one axis over a
// per-axis tap table, reducing (deposit) or expanding (interpolation) that
// axis: deposit x, y, z (intermediates 1/2 and 1/4 of the patch mesh),
// interpolation z, y, x (1/4 and 1/2).  Each pass streams its input once.
// The sums are regrouped against the reference's sequential loop (which
// adds ((W_x W_y) W_z) m term by term), so the result agrees to rounding
// (~1e-16 relative) rather than bit for bit; PMGPU_M2M_EXACT=1 selects the
// bit-identical gathers above.

constexpr int kMaxTaps = 64;  // taps of one output index on one axis (deposit: NS * ratio + 2)

struct AxisTab {
  int *first;   // first input index (0-based, ghosts included) of each output index
  int *count;   // number of consecutive input indices
  double *w;    // weights, tap t of output o at w[t * stride + o] (coalesced over o)
  int stride;
};

// Tap table of axis d.  Deposit (src finer): output index T (reference
// numbering) takes the source indices whose base b satisfies b - off <= T <=
// b - off + NS - 1, weight reldx W_{T - b + off}, zero for a cell outside
// the box on a periodic axis.  Interpolation (src coarser): output index T
// takes the stencil slots inside the source mesh, weight W_s.
template <int NS>
__global__ void k_m2m_table(MGeom src, MGeom dst, int d, int deposit, double reldx, AxisTab tab) {
  constexpr int off = NS == 6 ? 2 : 1;
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= dst.E[d]) return;
  const int T = o + 1 - dst.ng[d];
  double *w = tab.w + o;
  const int ws = tab.stride;
  int first = 0, n = 0;
  if (deposit) {
    int ilo, ihi;
    src_range(src, dst, d, T + off - NS + 1, T + off, ilo, ihi);
    for (int i = ilo; i <= ihi; ++i) {
      const double p = __dadd_rn(src.xmin[d], __dmul_rn(__dadd_rn((double)(i - 1), 0.5), src.dx[d]));
      const int b = cell_index(p, dst.xmin[d], dst.c1dx[d]);
      const int s = T - b + off;
      if (s < 0 || s >= NS) continue;
      if (n == 0) first = i - 1 + src.ng[d];
      if (n >= kMaxTaps) break;
      const bool live = !(dst.bc[d] != 0 && (p < dst.xmin[d] || p > dst.xmax[d]));
      const double dd = stencil_offset(p, dst.xmin[d], dst.dx[d], b, dst.c1dx[d]);
      w[ws * n++] = live ? __dmul_rn(reldx, weight(NS == 6, s, dd)) : 0.0;
    }
  } else {
    const double p = __dadd_rn(dst.xmin[d], __dmul_rn(__dadd_rn((double)(T - 1), 0.5), dst.dx[d]));
    const int b = cell_index(p, src.xmin[d], src.c1dx[d]);
    const double dd = stencil_offset(p, src.xmin[d], src.dx[d], b, src.c1dx[d]);
    for (int s = 0; s < NS; ++s) {
      const int q = b - off + s;
      if (q < 1 - src.ng[d] || q > src.ncell[d] + src.ng[d]) continue;
      if (n == 0) first = q - 1 + src.ng[d];
      w[ws * n++] = weight(NS == 6, s, dd);
    }
  }
  tab.first[o] = first;
  tab.count[o] = n;
}

struct PassArgs {
  int d;             // axis of the pass
  int Ein[3], Eout[3];
  AxisTab tab;
  int final_pass;    // writes the caller's mesh: clear / accumulate / skip rules apply
  int clear;
  int skip_bc[3];    // final interpolation pass: periodic axes whose ghost cells are not written
  int nc[3], ng[3];  // (of the output mesh, for the skip rule)
};

// One 1D pass: grid (ceil(Eout_x Eout_y / 256), Eout_z), a thread per
// output cell (all three components).
template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256) k_m2m_pass(const Tin *__restrict__ in, Tout *__restrict__ out, PassArgs a) {
  // XCD-aware order: the workgroups of one XCD take consecutive (xy, z)
  // tiles, so a z pass's taps on neighbouring planes are shared in its L2
  const int64_t lt = xcd_tile(blockIdx.x + (int64_t)gridDim.x * blockIdx.y, (int64_t)gridDim.x * gridDim.y);
  const int bx = (int)(lt % gridDim.x), bz = (int)(lt / gridDim.x);
  const int xy = (int)(bx * blockDim.x + threadIdx.x);
  if (xy >= a.Eout[0] * a.Eout[1]) return;
  int x[3] = {xy % a.Eout[0], xy / a.Eout[0], bz};
  const int64_t c = (int64_t)x[0] + (int64_t)a.Eout[0] * (x[1] + (int64_t)a.Eout[1] * x[2]);
  const int64_t sin[3] = {3, 3 * (int64_t)a.Ein[0], 3 * (int64_t)a.Ein[0] * a.Ein[1]};
  {
    if (a.final_pass) {
      bool skip = false;
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if (a.skip_bc[q] && (x[q] < a.ng[q] || x[q] >= a.ng[q] + a.nc[q])) skip = true;
      if (skip) {
        if (a.clear) out[3 * c] = out[3 * c + 1] = out[3 * c + 2] = (Tout)0;
        return;
      }
    }
    const int o = x[a.d];
    const int first = a.tab.first[o], cnt = a.tab.count[o];
    const double *w = a.tab.w + o;
    const int ws = a.tab.stride;
    x[a.d] = first;
    const Tin *p = in + sin[0] * x[0] + sin[1] * x[1] + sin[2] * x[2];
    const int64_t st = sin[a.d];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int t = 0; t < cnt; ++t) {
      const double wt = w[t * ws];
      s0 += wt * (double)p[0];
      s1 += wt * (double)p[1];
      s2 += wt * (double)p[2];
      p += st;
    }
    if (a.final_pass && !a.clear) {
      s0 += (double)out[3 * c];
      s1 += (double)out[3 * c + 1];
      s2 += (double)out[3 * c + 2];
    }
    out[3 * c] = (Tout)s0;
    out[3 * c + 1] = (Tout)s1;
    out[3 * c + 2] = (Tout)s2;
  }
}

// Intermediate meshes and tap tables, grown on demand, kept until
// pmgpu_finalize (one set per process: calls that overlap race).
struct M2MWS {
  void *a = nullptr, *b = nullptr;
  size_t na = 0, nb = 0;
  int *first = nullptr, *count = nullptr;
  double *w = nullptr;
  int ntab = 0;
};
M2MWS g_m2m;
std::mutex g_m2m_mu;

int grow(void *&p, size_t &have, size_t want) {
  if (have >= want) return 0;
  if (p) (void)hipFree(p);
  p = nullptr;
  have = 0;
  PM_HIP(hipMalloc(&p, want));
  have = want;
  return 0;
}

template <typename T>
int m2m_separable(const MGeom &f, const MGeom &t, int ns, const void *from, void *to, int clear, hipStream_t st) {
  const bool deposit = f.dx[0] / t.dx[0] <= 0.9999;
  const double reldx = f.dx[0] / t.dx[0];
  for (int d = 0; d < 3; ++d)
    PM_CHECK(!deposit || t.dx[d] / f.dx[d] <= (double)kMaxRatio + 1e-9,
             "pmgpu_interp_mesh_mesh: cell size ratios above 8 are not supported");
  // the three shapes: deposit reduces x, y, z in turn; interpolation expands z, y, x
  int E1[3], E2[3];
  const int order[3] = {deposit ? 0 : 2, 1, deposit ? 2 : 0};
  for (int q = 0; q < 3; ++q) E1[q] = f.E[q];
  E1[order[0]] = t.E[order[0]];
  for (int q = 0; q < 3; ++q) E2[q] = E1[q];
  E2[order[1]] = t.E[order[1]];
  const size_t bytes1 = sizeof(T) * 3 * (size_t)E1[0] * E1[1] * E1[2];
  const size_t bytes2 = sizeof(T) * 3 * (size_t)E2[0] * E2[1] * E2[2];
  const int emax = std::max(t.E[0], std::max(t.E[1], t.E[2]));
  std::lock_guard<std::mutex> lock(g_m2m_mu);
  if (grow(g_m2m.a, g_m2m.na, bytes1) || grow(g_m2m.b, g_m2m.nb, bytes2)) return -1;
  if (g_m2m.ntab < emax) {
    if (g_m2m.first) (void)hipFree(g_m2m.first);
    if (g_m2m.count) (void)hipFree(g_m2m.count);
    if (g_m2m.w) (void)hipFree(g_m2m.w);
    g_m2m.first = g_m2m.count = nullptr;
    g_m2m.w = nullptr;
    g_m2m.ntab = 0;
    PM_HIP(hipMalloc(&g_m2m.first, sizeof(int) * 3 * emax));
    PM_HIP(hipMalloc(&g_m2m.count, sizeof(int) * 3 * emax));
    PM_HIP(hipMalloc(&g_m2m.w, sizeof(double) * 3 * (size_t)emax * kMaxTaps));
    g_m2m.ntab = emax;
  }
  AxisTab tabs[3];
  for (int d = 0; d < 3; ++d) {
    tabs[d] = AxisTab{g_m2m.first + d * emax, g_m2m.count + d * emax, g_m2m.w + (size_t)d * emax * kMaxTaps,
                      emax};
    const int blocks = (t.E[d] + 127) / 128;
    if (ns == 6)
      hipLaunchKernelGGL(k_m2m_table<6>, dim3(blocks), dim3(128), 0, st, f, t, d, (int)deposit, reldx, tabs[d]);
    else
      hipLaunchKernelGGL(k_m2m_table<4>, dim3(blocks), dim3(128), 0, st, f, t, d, (int)deposit, reldx, tabs[d]);
    PM_HIP(hipGetLastError());
  }
  const void *ins[3] = {from, g_m2m.a, g_m2m.b};
  void *outs[3] = {g_m2m.a, g_m2m.b, to};
  const int *Eins[3] = {f.E, E1, E2};
  const int *Eouts[3] = {E1, E2, t.E};
  for (int q = 0; q < 3; ++q) {
    PassArgs a{};
    a.d = order[q];
    for (int k = 0; k < 3; ++k) {
      a.Ein[k] = Eins[q][k];
      a.Eout[k] = Eouts[q][k];
      a.skip_bc[k] = (!deposit && t.bc[k] != 0) ? 1 : 0;
      a.nc[k] = t.ncell[k];
      a.ng[k] = t.ng[k];
    }
    a.tab = tabs[a.d];
    a.final_pass = q == 2;
    a.clear = clear;
    const dim3 grid((a.Eout[0] * a.Eout[1] + 255) / 256, a.Eout[2]);
    hipLaunchKernelGGL((k_m2m_pass<T, T>), grid, dim3(256), 0, st, static_cast<const T *>(ins[q]),
                       static_cast<T *>(outs[q]), a);
    PM_HIP(hipGetLastError());
  }
  T *dst = static_cast<T *>(to);
  if (deposit) {
    PGeom pg{};
    for (int d = 0; d < 3; ++d) {
      pg.ncell[d] = t.ncell[d];
      pg.ngl[d] = t.ng[d];
      pg.E[d] = t.E[d];
    }
    for (int d = 0; d < 3; ++d)
      if (t.bc[d] != 0) {
        hipLaunchKernelGGL(k_halo_add<T>, dim3(256), dim3(256), 0, st, dst, pg, d);
        PM_HIP(hipGetLastError());
      }
  } else {
    for (int d = 0; d < 3; ++d)
      if (t.bc[d] != 0) {
        hipLaunchKernelGGL(k_ghost_fill<T>, dim3(1024), dim3(256), 0, st, dst, t.E[0], t.E[1], t.E[2], d,
                           t.ncell[d], t.ng[d], 3);
        PM_HIP(hipGetLastError());
      }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Bit-identical gathers (PMGPU_M2M_EXACT=1)
template <typename T>
int m2m_exact(const MGeom &f, const MGeom &t, int ns, const void *from, void *to, int clear, hipStream_t st) {
  const double reldx = f.dx[0] / t.dx[0];
  const T *src = static_cast<const T *>(from);
  T *dst = static_cast<T *>(to);
  if (reldx <= 0.9999) {
    for (int d = 0; d < 3; ++d)
      PM_CHECK(t.dx[d] / f.dx[d] <= (double)kMaxRatio + 1e-9,
               "pmgpu_interp_mesh_mesh: cell size ratios above 8 are not supported");
    DepArgs a{f, t, reldx, clear};
    dim3 grid((t.E[0] + kTX - 1) / kTX, t.E[1], t.E[2]);
    if (ns == 6) hipLaunchKernelGGL((k_m2m_deposit<T, 6>), grid, dim3(kTX), 0, st, src, dst, a);
    else hipLaunchKernelGGL((k_m2m_deposit<T, 4>), grid, dim3(kTX), 0, st, src, dst, a);
    PM_HIP(hipGetLastError());
    // the ghost layers of periodic axes are added onto their images
    // (:1380-1386, pmlib_mesh_map_ghost with incl_edges, op add)
    PGeom pg{};
    for (int d = 0; d < 3; ++d) {
      pg.ncell[d] = t.ncell[d];
      pg.ngl[d] = t.ng[d];
      pg.E[d] = t.E[d];
    }
    for (int d = 0; d < 3; ++d)
      if (t.bc[d] != 0) {
        hipLaunchKernelGGL(k_halo_add<T>, dim3(256), dim3(256), 0, st, dst, pg, d);
        PM_HIP(hipGetLastError());
      }
  } else {
    const int64_t n = (int64_t)t.E[0] * t.E[1] * t.E[2];
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    if (ns == 6) hipLaunchKernelGGL((k_m2m_interp<T, 6>), dim3(blocks), dim3(256), 0, st, src, dst, f, t, clear);
    else hipLaunchKernelGGL((k_m2m_interp<T, 4>), dim3(blocks), dim3(256), 0, st, src, dst, f, t, clear);
    PM_HIP(hipGetLastError());
    // periodic ghosts copied from their images (:2057-2063, op replace)
    for (int d = 0; d < 3; ++d)
      if (t.bc[d] != 0) {
        hipLaunchKernelGGL(k_ghost_fill<T>, dim3(1024), dim3(256), 0, st, dst, t.E[0], t.E[1], t.E[2], d,
                           t.ncell[d], t.ng[d], 3);
        PM_HIP(hipGetLastError());
      }
  }
  return 0;
}

int m2m_check(const pmgpu_geom *gf, const pmgpu_geom *gt, MGeom &f, MGeom &t) {
  PM_CHECK(gf && gt, "pmgpu_interp_mesh_mesh: null geometry");
  if (make_mgeom(gf, f, "pmgpu_interp_mesh_mesh") || make_mgeom(gt, t, "pmgpu_interp_mesh_mesh")) return -1;
  PM_CHECK(gf->interp_order == gt->interp_order,
           "pmgpu_interp_mesh_mesh: both meshes must use the same interpolation order");
  for (int d = 0; d < 3; ++d)
    PM_CHECK(t.bc[d] == 0 || t.ncell[d] >= 2 * t.ng[d],
             "pmgpu_interp_mesh_mesh: a periodic target axis needs at least twice its ghost count in cells");
  return 0;
}

}  // namespace

void m2m_release() {
  std::lock_guard<std::mutex> lock(g_m2m_mu);
  void *bufs[] = {g_m2m.a, g_m2m.b, g_m2m.first, g_m2m.count, g_m2m.w};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  g_m2m = M2MWS{};
}
}  // namespace pmgpu

using namespace pmgpu;

extern "C" {

int pmgpu_topology_parent(const pmgpu_geom *g, int32_t level, int32_t parent_level, pmgpu_geom *out) {
  PM_CHECK(g && out, "pmgpu_topology_parent: null argument");
  PM_CHECK(parent_level != 0, "pmlib_topology_parent: Cannot create parent topology on level 0");
  PM_CHECK(g->interp_order == 3 || g->interp_order == 4,
           "pmlib_topology_parent: Can not define number of ghost cells.");
  const double reldx = std::pow(2.0, (double)(level - parent_level));
  const long r = std::lround(reldx);
  PM_CHECK(r >= 1, "pmgpu_topology_parent: the parent level must not be finer than the patch");
  pmgpu_geom p = *g;
  for (int d = 0; d < 3; ++d) {
    PM_CHECK(g->ncell[d] % r == 0, "pmlib_topology_parent: Number mesh cells does not match");
    p.ncell[d] = (int32_t)std::lround((double)g->ncell[d] / reldx);
    p.dx[d] = reldx * g->dx[d];
  }
  p.nghost = g->interp_order == 4 ? 3 : 2;
  *out = p;
  return 0;
}

int pmgpu_interp_mesh_mesh_device(const pmgpu_geom *g_from, const pmgpu_geom *g_to, int32_t precision,
                                  const void *d_from, void *d_to, int32_t clear, void *stream) {
  MGeom f, t;
  if (m2m_check(g_from, g_to, f, t)) return -1;
  PM_CHECK(d_from && d_to, "pmgpu_interp_mesh_mesh: null mesh");
  PM_CHECK(precision == PMGPU_FP64 || precision == PMGPU_FP32, "pmgpu_interp_mesh_mesh: bad precision");
  const int ns = g_to->interp_order == 4 ? 6 : 4;
  hipStream_t st = (hipStream_t)stream;
  const char *ex = getenv("PMGPU_M2M_EXACT");
  if (ex && atoi(ex) != 0)
    return precision == PMGPU_FP64 ? m2m_exact<double>(f, t, ns, d_from, d_to, clear, st)
                                   : m2m_exact<float>(f, t, ns, d_from, d_to, clear, st);
  return precision == PMGPU_FP64 ? m2m_separable<double>(f, t, ns, d_from, d_to, clear, st)
                                 : m2m_separable<float>(f, t, ns, d_from, d_to, clear, st);
}

int pmgpu_interp_mesh_mesh(const pmgpu_geom *g_from, const pmgpu_geom *g_to, const double *from, double *to,
                           int32_t clear) {
  MGeom f, t;
  if (m2m_check(g_from, g_to, f, t)) return -1;
  PM_CHECK(from && to, "pmgpu_interp_mesh_mesh: null mesh");
  const size_t bf = sizeof(double) * 3 * (size_t)f.E[0] * f.E[1] * f.E[2];
  const size_t bt = sizeof(double) * 3 * (size_t)t.E[0] * t.E[1] * t.E[2];
  void *df = nullptr, *dt = nullptr;
  int e = 0;
  if (hipMalloc(&df, bf) != hipSuccess || hipMalloc(&dt, bt) != hipSuccess) {
    set_error("pmgpu_interp_mesh_mesh: device allocation failed");
    e = -1;
  }
  if (!e && (hipMemcpy(df, from, bf, hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(dt, to, bt, hipMemcpyHostToDevice) != hipSuccess)) {
    set_error("pmgpu_interp_mesh_mesh: host to device copy failed");
    e = -1;
  }
  if (!e) e = pmgpu_interp_mesh_mesh_device(g_from, g_to, PMGPU_FP64, df, dt, clear, nullptr);
  if (!e && hipMemcpy(to, dt, bt, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("pmgpu_interp_mesh_mesh: device to host copy failed");
    e = -1;
  }
  if (df) (void)hipFree(df);
  if (dt) (void)hipFree(dt);
  return e;
}

}  // extern "C"
