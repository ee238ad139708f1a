// Right-hand side of the vorticity equation on the mesh (HIP, gfx950):
//   dvort_l = visc lap(w_l) + w . grad(u_l)   on the interior cells,
// then the periodic ghost map of dvort.
//
// Restates part_deriv_vorticity (vortex_codes/vortex_penalisation/source/
// particle_derivatives/part_deriv_vorticity.f:57-115 for the 2nd-order and
// :117-196 for the 4th-order stencils, :201-208 for the ghost map).  One
// thread per component of an interior cell; the stencil neighbours along i
// belong to nearby lanes, so the re-reads hit L1/L2.
// Every floating-point operation is an explicitly rounded one, in the order
// the reference writes it, so the FP64 result is bit-identical to a
// left-to-right evaluation of the reference's expressions.  Ghost cells of
// dvort on free sides are not written.
#include "../../include/pmgpu.h"
#include "pm_common.h"

namespace pmgpu {

template <typename T> struct RhsArgs {
  const T *w, *u;
  T *dw;
  int E[3], N[3], ng[3];
  T visc, cdx2[3], c2dx[3];  // 1 / dx^2 and 1 / (2 dx) per axis
};

// One thread per component word of kRows rows (consecutive threads own
// consecutive words of a row, components fastest); a block covers kRowWords
// words of kRows consecutive rows and marches through kPlanes planes.  Each
// thread keeps its words' values on the planes within H = FD / 2 of the
// current one in registers (one new plane loaded per step); the current
// plane's tile with an H-cell halo in x and y is staged in LDS (the centre
// from those registers, the halo from memory), and the x and y stencils read
// it there.
constexpr int kRowWords = 256, kRows = 4, kPlanes = 16;

template <typename T, int FD>
__global__ void __launch_bounds__(kRowWords) k_vort_rhs(RhsArgs<T> a) {
  constexpr int H = FD / 2, HW = 3 * H, TW = kRowWords + 2 * HW, TR = kRows + 2 * H;
  __shared__ T lw[TR][TW], lu[TR][TW];
  // XCD-aware tile order: consecutive row groups (which read each other's
  // rows as y neighbours) run on one XCD and share its L2
  const int64_t lt = xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y * gridDim.z);
  const int bx = (int)(lt % gridDim.x), j0 = (int)((lt / gridDim.x) % gridDim.y) * kRows;
  const int kz = (int)(lt / ((int64_t)gridDim.x * gridDim.y));
  const int tid = threadIdx.x;
  const int x0 = bx * kRowWords;                 // first interior word of the tile
  const int x3 = x0 + tid, nw = 3 * a.N[0];      // this thread's word, words per row
  const bool mine = x3 < nw;
  const int l = x3 % 3;
  const int k0 = kz * kPlanes, k1 = k0 + kPlanes < a.N[2] ? k0 + kPlanes : a.N[2];
  const int64_t sy = 3 * (int64_t)a.E[0], sz = 3 * (int64_t)a.E[0] * a.E[1];
  // word wd (interior-relative, may be a halo word), row jj, plane kk
  auto at = [&](int wd, int jj, int kk) {
    return (int64_t)(3 * a.ng[0] + wd) + sy * (jj + a.ng[1]) + sz * (kk + a.ng[2]);
  };
  // the queues also hold the ghost words and rows next to the interior (the
  // tile centre positions the neighbours of interior words read)
  T zw[kRows][2 * H + 1], zu[kRows][2 * H + 1];
#pragma unroll
  for (int r = 0; r < kRows; ++r)
#pragma unroll
    for (int m = 0; m < 2 * H; ++m) {
      const bool ok = x3 < nw + HW && j0 + r < a.N[1] + H;
      zw[r][m + 1] = ok ? a.w[at(x3, j0 + r, k0 + m - H)] : T(0);
      zu[r][m + 1] = ok ? a.u[at(x3, j0 + r, k0 + m - H)] : T(0);
    }
  for (int k = k0; k < k1; ++k) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
#pragma unroll
      for (int m = 0; m < 2 * H; ++m) {
        zw[r][m] = zw[r][m + 1];
        zu[r][m] = zu[r][m + 1];
      }
      const bool ok = x3 < nw + HW && j0 + r < a.N[1] + H;
      zw[r][2 * H] = ok ? a.w[at(x3, j0 + r, k + H)] : T(0);
      zu[r][2 * H] = ok ? a.u[at(x3, j0 + r, k + H)] : T(0);
      lw[H + r][HW + tid] = zw[r][H];
      lu[H + r][HW + tid] = zu[r][H];
    }
    // halo: H rows above and below (whole tile width), HW words left and
    // right of the centre rows; positions outside the mesh are not needed
    for (int idx = tid; idx < TR * TW; idx += kRowWords) {
      const int rr = idx / TW, c = idx - rr * TW;
      if (rr >= H && rr < H + kRows && c >= HW && c < HW + kRowWords) continue;  // centre
      const int wd = x0 - HW + c, jj = j0 - H + rr;
      if (wd >= -HW && wd < nw + HW && jj >= -H && jj < a.N[1] + H) {
        lw[rr][c] = a.w[at(wd, jj, k)];
        lu[rr][c] = a.u[at(wd, jj, k)];
      }
    }
    __syncthreads();
    if (mine) {
#pragma unroll
      for (int r = 0; r < kRows; ++r) {
        if (j0 + r >= a.N[1]) break;
        const int rr = H + r, c = HW + tid;
        const T w1 = lw[rr][c - l], w2 = lw[rr][c - l + 1], w3 = lw[rr][c - l + 2];
        T dd[3], du[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          // neighbours at -2, -1, 0, +1, +2 along d: x and y from the tile,
          // z from the register queue
          auto W = [&](int o) {
            return d == 0 ? lw[rr][c + 3 * o] : (d == 1 ? lw[rr + o][c] : zw[r][H + o]);
          };
          auto U = [&](int o) {
            return d == 0 ? lu[rr][c + 3 * o] : (d == 1 ? lu[rr + o][c] : zu[r][H + o]);
          };
          if (FD == 2) {
            // (w(-1) - 2 w(0) + w(+1)) / dx^2 and (-u(-1) + u(+1)) / (2 dx)
            dd[d] = rmul(radd(rsub(W(-1), rmul(T(2), W(0))), W(1)), a.cdx2[d]);
            du[d] = rmul(radd(-U(-1), U(1)), a.c2dx[d]);
          } else {
            // (c1, c2, -2.5, c2, c1) / dx^2 and (c3, -c4, c4, -c3) / (2 dx)
            const T c1 = T(-1.0 / 12.0), c2 = T(4.0 / 3.0), c3 = T(1.0 / 6.0), c4 = T(4.0 / 3.0);
            T v = radd(rmul(c1, W(-2)), rmul(c2, W(-1)));
            v = rsub(v, rmul(T(2.5), W(0)));
            v = radd(v, rmul(c2, W(1)));
            v = radd(v, rmul(c1, W(2)));
            dd[d] = rmul(v, a.cdx2[d]);
            T x = rsub(rmul(c3, U(-2)), rmul(c4, U(-1)));
            x = radd(x, rmul(c4, U(1)));
            x = rsub(x, rmul(c3, U(2)));
            du[d] = rmul(x, a.c2dx[d]);
          }
        }
        T res = radd(rmul(a.visc, radd(radd(dd[0], dd[1]), dd[2])), rmul(w1, du[0]));
        res = radd(res, rmul(w2, du[1]));
        res = radd(res, rmul(w3, du[2]));
        a.dw[at(x3, j0 + r, k)] = res;
      }
    }
    __syncthreads();  // the tile is rewritten for the next plane
  }
}

template <typename T>
static int vort_rhs_launch(const pmgpu_geom *geom, const void *w, const void *u, double visc, int fd_order,
                           void *dw, hipStream_t st) {
  PM_CHECK(geom && w && u && dw, "pmgpu_vorticity_rhs: null argument");
  PM_CHECK(fd_order == 2 || fd_order == 4, "pmgpu_vorticity_rhs: fd_order must be 2 or 4");
  PM_CHECK(w != dw && u != dw, "pmgpu_vorticity_rhs: dvort must not alias the vorticity or velocity mesh");
  RhsArgs<T> a;
  int32_t ext[3];
  if (pmgpu_mesh_extent(geom, ext)) return -1;
  for (int d = 0; d < 3; ++d) {
    a.E[d] = ext[d];
    a.N[d] = geom->ncell[d];
    a.ng[d] = (ext[d] - geom->ncell[d]) / 2;
    PM_CHECK(a.ng[d] >= fd_order / 2, "pmgpu_vorticity_rhs: the stencil needs fd_order/2 ghost layers");
    // part_deriv_vorticity.f:67-73: 1/dx**2 and 1/(2 dx), in the working precision
    a.cdx2[d] = T(1.0) / (T(geom->dx[d]) * T(geom->dx[d]));
    a.c2dx[d] = T(1.0) / (T(2.0) * T(geom->dx[d]));
  }
  a.w = (const T *)w;
  a.u = (const T *)u;
  a.dw = (T *)dw;
  a.visc = T(visc);
  if ((int64_t)a.N[0] * a.N[1] * a.N[2] > 0) {
    PM_CHECK(a.N[1] / kRows < 65535 && a.N[2] / kPlanes < 65535, "pmgpu_vorticity_rhs: mesh too large");
    const dim3 grid((3 * a.N[0] + kRowWords - 1) / kRowWords, (a.N[1] + kRows - 1) / kRows,
                    (a.N[2] + kPlanes - 1) / kPlanes);
    if (fd_order == 4)
      hipLaunchKernelGGL((k_vort_rhs<T, 4>), grid, dim3(kRowWords), 0, st, a);
    else
      hipLaunchKernelGGL((k_vort_rhs<T, 2>), grid, dim3(kRowWords), 0, st, a);
    PM_HIP(hipGetLastError());
  }
  // ghost map of dvort, x then y then z (periodic sides)
  for (int d = 0; d < 3; ++d)
    if (geom->bound_cond[d] == 1) {
      PM_CHECK(a.N[d] >= a.ng[d], "pmgpu_vorticity_rhs: periodic axis shorter than its ghost layers");
      hipLaunchKernelGGL(k_ghost_fill<T>, dim3(1024), dim3(256), 0, st, (T *)dw, a.E[0], a.E[1], a.E[2], d,
                         a.N[d], a.ng[d], 3);
      PM_HIP(hipGetLastError());
    }
  return 0;
}

}  // namespace pmgpu

using namespace pmgpu;

extern "C" {

int pmgpu_vorticity_rhs(const pmgpu_geom *g, const double *vort, const double *vel, double visc, int32_t fd_order,
                        double *dvort) {
  PM_CHECK(g && vort && vel && dvort, "pmgpu_vorticity_rhs: null argument");
  int32_t e[3];
  if (pmgpu_mesh_extent(g, e)) return -1;
  const size_t bytes = sizeof(double) * 3 * (size_t)e[0] * e[1] * e[2];
  double *dw = nullptr, *du = nullptr, *dd = nullptr;
  int rc = -1;
  if (hipMalloc(&dw, bytes) == hipSuccess && hipMalloc(&du, bytes) == hipSuccess &&
      hipMalloc(&dd, bytes) == hipSuccess && hipMemcpy(dw, vort, bytes, hipMemcpyHostToDevice) == hipSuccess &&
      hipMemcpy(du, vel, bytes, hipMemcpyHostToDevice) == hipSuccess &&
      hipMemcpy(dd, dvort, bytes, hipMemcpyHostToDevice) == hipSuccess) {
    rc = vort_rhs_launch<double>(g, dw, du, visc, fd_order, dd, nullptr);
    if (rc == 0 && hipMemcpy(dvort, dd, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
      set_error("pmgpu_vorticity_rhs: device copy failed");
      rc = -1;
    }
  } else {
    set_error("pmgpu_vorticity_rhs: device allocation or copy failed");
  }
  if (dw) (void)hipFree(dw);
  if (du) (void)hipFree(du);
  if (dd) (void)hipFree(dd);
  return rc;
}

int pmgpu_vorticity_rhs_device(const pmgpu_geom *g, int32_t precision, const void *d_vort, const void *d_vel,
                               double visc, int32_t fd_order, void *d_dvort, void *stream) {
  if (precision == PMGPU_FP32)
    return vort_rhs_launch<float>(g, d_vort, d_vel, visc, fd_order, d_dvort, (hipStream_t)stream);
  PM_CHECK(precision == PMGPU_FP64, "pmgpu_vorticity_rhs_device: precision must be PMGPU_FP64 or PMGPU_FP32");
  return vort_rhs_launch<double>(g, d_vort, d_vel, visc, fd_order, d_dvort, (hipStream_t)stream);
}

}  // extern "C"
