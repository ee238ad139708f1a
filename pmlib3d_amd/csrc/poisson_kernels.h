// Kernels of the free-space / periodic Poisson solve (HIP, gfx950).
//
// Data path of one solve (kernel 1, fft_seq (1,2,3)), all on device:
//   X  : real mesh (3, Ex, Ey, Ez) -> A[c][z][y][kx]   R2C, implicit zero pad
//   Y  : A -> B[c][z][ky][kx]                         C2C, implicit zero pad
//   Z  : B -> B (+ psi / vort buffers)                 fused pad, FFT, G-hat
//        product, spectral curl, inverse FFT, truncate (one kernel)
//   Y' : B -> A                                        C2C inverse, truncate
//   X' : A -> real mesh                                C2R, truncate
// Semantics follow poisson/pmlib_poisson_solve.f:119-1013 and
// fourier/pmlib_fourier_{fft,ifft}.f of the reference.
#pragma once
#include "pm_fft.h"

namespace pmgpu {

// Real-input line pass (R2C) along x with implicit zero padding.
// Line (y, z) of component c: element x at in[c*cs + x*xs + y*ys + z*zs],
// x < Min; the transform length is 2*Nh.  Output element (c, kx, y, z) at
// out[c*o_cs + kx*o_ks + y*o_ys + z*o_zs]; the solve uses the
// component-interleaved spectral layout A[((z*My + y)*Hx + kx)*3 + c].
struct XFwdArgs {
  const void *in;
  void *out;
  int64_t cs, xs, ys, zs;
  int64_t o_cs, o_ks, o_ys, o_zs;
  int ncomp, Min, My, Mz, Hx, lines_per_block;
  FftPlan plan;     // length Nh
  int twL_off;      // offset of the length-2Nh table (post-processing twiddles)
};

template <typename R>
__global__ void __launch_bounds__(1024) k_x_fwd(XFwdArgs a, const cplx_t<R> *tw) {
  using V = cplx_t<R>;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);
  const int Nh = a.plan.L, P = a.plan.P, st = lds_stride(Nh);
  const int T = a.lines_per_block, NL = T * a.ncomp;
  const int y0 = blockIdx.x * T, z = blockIdx.y;
  const R *in = reinterpret_cast<const R *>(a.in);
  // load + pack x[2m] + i x[2m+1] into line (t, c); zero the rest
  for (int idx = threadIdx.x; idx < NL * st; idx += blockDim.x) lds[idx] = cmake<V>(0, 0);
  __syncthreads();
  const int per_line = a.ncomp * a.Min;
  for (int idx = threadIdx.x; idx < T * per_line; idx += blockDim.x) {
    int t = idx / per_line, e = idx - t * per_line;
    int x = e / a.ncomp, c = e - x * a.ncomp;
    int y = y0 + t;
    if (y < a.My) {
      R v = in[c * a.cs + (int64_t)x * a.xs + (int64_t)y * a.ys + (int64_t)z * a.zs];
      R *slot = reinterpret_cast<R *>(&lds[(t * a.ncomp + c) * st + (x >> 1)]);
      slot[x & 1] = v;
    }
  }
  __syncthreads();
  const int line = threadIdx.x / P, p = threadIdx.x - line * P;
  V *myline = lds + (line < NL ? line : 0) * st;
  fft_lds<-1>(myline, a.plan, p, tw);
  // post-process: X[k] = (Z[k] + conj Z[Nh-k])/2 - i/2 W^k (Z[k] - conj Z[Nh-k])
  V *out = reinterpret_cast<V *>(a.out);
  const V *twL = tw + a.twL_off;
  const int Hx = a.Hx;
  for (int idx = threadIdx.x; idx < NL * Hx; idx += blockDim.x) {
    int l = idx / Hx, k = idx - l * Hx;
    int t = l / a.ncomp, c = l - t * a.ncomp;
    int y = y0 + t;
    if (y >= a.My) continue;
    const V *zl = lds + l * st;
    V zk = zl[k == Nh ? 0 : k];
    V zm = cconj(zl[k == 0 ? 0 : Nh - k]);
    V e = cscale(cadd(zk, zm), (R)0.5);
    V o = cscale(csub(zk, zm), (R)0.5);
    // -i * W^k * o
    V w = twL[k];
    V wo = cmul(w, o);
    V X = cadd(e, cmake<V>(wo.y, -wo.x));
    out[c * a.o_cs + k * a.o_ks + (int64_t)y * a.o_ys + (int64_t)z * a.o_zs] = X;
  }
}

// Strided complex pass along a non-contiguous axis.  Element (col, row, b) at
// base + b*bstride + row*rstride + col; a block transforms T consecutive
// columns of one batch entry.  Rows >= nin read as zero, rows < nout written,
// multiplied by scale.  in and out may alias (each block owns its columns).
struct StridedArgs {
  const void *in;
  void *out;
  int64_t rs_in, rs_out, bs_in, bs_out;
  int ncols, nin, nout, T;
  double scale;
  FftPlan plan;
  int kernel;  // 0: register-resident where the length allows, 1: LDS-resident
};

template <typename R, int S>
__global__ void __launch_bounds__(1024) k_strided(StridedArgs a, const cplx_t<R> *tw) {
  using V = cplx_t<R>;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);
  const int L = a.plan.L, P = a.plan.P, st = lds_stride(L), T = a.T;
  const int col0 = blockIdx.x * T;
  const int64_t b = blockIdx.y + (int64_t)blockIdx.z * gridDim.y;
  const V *in = reinterpret_cast<const V *>(a.in) + b * a.bs_in + col0;
  V *out = reinterpret_cast<V *>(a.out) + b * a.bs_out + col0;
  for (int idx = threadIdx.x; idx < T * L; idx += blockDim.x) {
    int row = idx / T, c = idx - row * T;
    V v = cmake<V>(0, 0);
    if (row < a.nin && col0 + c < a.ncols) v = in[(int64_t)row * a.rs_in + c];
    lds[c * st + row] = v;
  }
  __syncthreads();
  const int line = threadIdx.x / P, p = threadIdx.x - line * P;
  fft_lds<S>(lds + (line < T ? line : 0) * st, a.plan, p, tw);
  const R sc = (R)a.scale;
  for (int idx = threadIdx.x; idx < T * a.nout; idx += blockDim.x) {
    int row = idx / T, c = idx - row * T;
    if (col0 + c < a.ncols) out[(int64_t)row * a.rs_out + c] = cscale(lds[c * st + row], sc);
  }
}

// Complex-to-real line pass along x (inverse of k_x_fwd), keeping x < Mout.
// Input element (c, kx, y, z) at in[c*i_cs + kx*i_ks + y*i_ys + z*i_zs].
struct XInvArgs {
  const void *in;
  void *out;
  int64_t cs, xs, ys, zs;
  int64_t i_cs, i_ks, i_ys, i_zs;
  int ncomp, Mout, My, Mz, Hx, lines_per_block;
  double scale;
  FftPlan plan;  // length Nh
  int twL_off;
};

template <typename R>
__global__ void __launch_bounds__(1024) k_x_inv(XInvArgs a, const cplx_t<R> *tw) {
  using V = cplx_t<R>;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);
  const int Nh = a.plan.L, P = a.plan.P, st = lds_stride(Nh);
  const int T = a.lines_per_block, NL = T * a.ncomp;
  const int y0 = blockIdx.x * T, z = blockIdx.y;
  const V *in = reinterpret_cast<const V *>(a.in);
  const V *twL = tw + a.twL_off;
  // Z[k] = (X[k] + conj X[Nh-k]) + i W^-k (X[k] - conj X[Nh-k]), k < Nh;
  // imaginary parts of X[0] and X[Nh] are dropped (Hermitian projection).
  for (int idx = threadIdx.x; idx < NL * Nh; idx += blockDim.x) {
    int l = idx / Nh, k = idx - l * Nh;
    int t = l / a.ncomp, c = l - t * a.ncomp;
    int y = y0 + t;
    V Zk = cmake<V>(0, 0);
    if (y < a.My) {
      const V *xl = in + c * a.i_cs + (int64_t)y * a.i_ys + (int64_t)z * a.i_zs;
      V xk = xl[k * a.i_ks], xm = xl[(Nh - k) * a.i_ks];
      if (k == 0) { xk.y = 0; xm.y = 0; }
      V s = cadd(xk, cconj(xm));
      V d = csub(xk, cconj(xm));
      V w = cconj(twL[k]);
      V wd = cmul(w, d);
      Zk = cadd(s, cmake<V>(-wd.y, wd.x));
    }
    lds[l * st + k] = Zk;
  }
  __syncthreads();
  const int line = threadIdx.x / P, p = threadIdx.x - line * P;
  fft_lds<1>(lds + (line < NL ? line : 0) * st, a.plan, p, tw);
  R *out = reinterpret_cast<R *>(a.out);
  const R sc = (R)a.scale;
  const int per_line = a.ncomp * a.Mout;
  for (int idx = threadIdx.x; idx < T * per_line; idx += blockDim.x) {
    int t = idx / per_line, e = idx - t * per_line;
    int x = e / a.ncomp, c = e - x * a.ncomp;
    int y = y0 + t;
    if (y < a.My) {
      const R *slot = reinterpret_cast<const R *>(&lds[(t * a.ncomp + c) * st + (x >> 1)]);
      out[c * a.cs + (int64_t)x * a.xs + (int64_t)y * a.ys + (int64_t)z * a.zs] = slot[x & 1] * sc;
    }
  }
}

// Fused third-direction pass: zero pad, forward FFT of the 3 vorticity
// components, G-hat product + spectral curl (+ psi, + regularised /
// reprojected vorticity), inverse FFTs, truncation to Mz rows, 1/Lz.
// One patch's share of a group launch (one kernel instance run over several
// patches): its arguments, its twiddle table, its first tile in the launch's
// global tile order and its tiles per grid row.
template <typename A> struct GrpEntry {
  A a;
  const void *tw;
  int64_t off;
  int64_t gx;
};

// scalar z-pass choice: T < kZ1Reg the LDS-resident kernel with T columns
// per block, kZ1Reg + TC the register-resident one with TC columns per block
constexpr int kZ1Reg = 100;

struct ZArgs {
  void *B;            // in: w-hat (3 comps), out: u-hat (in place)
  void *Bpsi;         // optional psi-hat output
  void *Bvort;        // optional regularised vorticity output
  const void *Ghat;   // octant table [(kz' * Hy + ky') * Hx + kx] (kernel 1)
  int64_t rs, cs;     // row (z) stride and component stride of B
  int64_t kxs, kys;   // kx and ky strides of B (interleaved layout: 3 and 3*Hx)
  int Hx, Ly, Lz, Lx, Mz, T, Hy, Hz;
  int order;          // regularisation order for zeta
  int allper;         // all axes periodic: G = zeta / (4 pi^2 |k|^2)
  int want_psi, want_vort, reproj;
  int tune_tc;        // z-pass columns per block (0 = default)
  int zkernel;        // velocity-only z pass: 0 register-resident if available, 1 LDS-resident
  int ker2;           // kernel 2 (grad G): Ghat is the complex half-spectrum
                      // [((kz * Ly + ky) * Hx + kx) * 3 + c] of the three components
  double dk[3], ksmp[3], h3, eps2pi;
  FftPlan plan;       // length Lz
};

template <typename R>
__device__ __forceinline__ R wavenumber(int n, int L, double dk, double ksmp) {
  // poisson/pmlib_poisson_solve.f:724-766 with kmin = -ksmp on one rank,
  // evaluated in the working precision
  const R kmin = -(R)ksmp;
  const R base = kmin + (R)n * (R)dk;
  return (n < L / 2) ? base + (R)ksmp : base - (R)ksmp;
}

template <typename R>
__device__ __forceinline__ R zeta_filter(int order, R s) {
  R s2 = s * s, e = exp((R)-0.5 * s2);
  switch (order) {
    case 0: return (R)1;
    case 2: return e;
    case 4: return ((R)1 + (R)0.5 * s2) * e;
    case 6: return ((R)1 + (R)0.5 * s2 + (R)0.125 * (s2 * s2)) * e;
    case 8: return ((R)1 + (R)0.5 * s2 + (R)0.125 * (s2 * s2) + (R)(1.0 / 48.0) * (s2 * s2 * s2)) * e;
    default:
      return ((R)1 + (R)0.5 * s2 + (R)0.125 * (s2 * s2) + (R)(1.0 / 48.0) * (s2 * s2 * s2) +
              (R)(1.0 / 384.0) * (s2 * s2 * s2 * s2)) * e;
  }
}

template <typename R, int Q>
__global__ void __launch_bounds__(1024) k_z_convolve(ZArgs a, const cplx_t<R> *tw) {
  using V = cplx_t<R>;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);
  const int L = a.plan.L, P = a.plan.P, st = lds_stride(L), T = a.T;
  const int kx0 = blockIdx.x * T, ky = blockIdx.y;
  const int line = threadIdx.x / P, p = threadIdx.x - line * P;
  const int myl = line < T ? line : 0;
  V *B = reinterpret_cast<V *>(a.B);
  V w[3][Q];
  for (int q = 0; q < 3; ++q) {
    const V *src = B + q * a.cs + (int64_t)ky * a.kys + (int64_t)kx0 * a.kxs;
    for (int idx = threadIdx.x; idx < T * L; idx += blockDim.x) {
      int row = idx / T, c = idx - row * T;
      V v = cmake<V>(0, 0);
      if (row < a.Mz && kx0 + c < a.Hx) v = src[(int64_t)row * a.rs + c * a.kxs];
      lds[c * st + row] = v;
    }
    __syncthreads();
    fft_lds<-1>(lds + myl * st, a.plan, p, tw);
#pragma unroll
    for (int e = 0; e < Q; ++e) {
      int n = p + e * P;
      w[q][e] = n < L ? lds[myl * st + n] : cmake<V>(0, 0);
    }
    __syncthreads();
  }
  // wavenumbers of this column
  const int kx = kx0 + myl;
  const bool col_ok = line < T && kx < a.Hx;
  const R ki = wavenumber<R>(kx, a.Lx, a.dk[0], a.ksmp[0]);
  const R kj = wavenumber<R>(ky, a.Ly, a.dk[1], a.ksmp[1]);
  const bool nx = (2 * kx == a.Lx), ny = (2 * ky == a.Ly);
  const int kyf = ky <= a.Ly - ky ? ky : a.Ly - ky;
  const R pi = (R)3.14159265358979323846;
  const R twopi = (R)6.28318530717958647692;
  const R *G = reinterpret_cast<const R *>(a.Ghat);
  const int nout = 3 + (a.want_psi ? 3 : 0) + (a.want_vort ? 3 : 0);
  const R scale = (R)1 / (R)L;
  for (int o = 0; o < nout; ++o) {
    int kind = o / 3, comp = o - kind * 3;  // kind 0 vel, 1 psi or vort, 2 vort
    if (kind == 1 && !a.want_psi) kind = 2;
#pragma unroll
    for (int e = 0; e < Q; ++e) {
      int n = p + e * P;
      if (n >= L) continue;
      const R kk = wavenumber<R>(n, a.Lz, a.dk[2], a.ksmp[2]);
      const bool nz = (2 * n == a.Lz);
      const int kzf = n <= a.Lz - n ? n : a.Lz - n;
      R zeta = (R)1;
      if (a.want_vort || a.allper) {
        if (a.want_vort) {
          R s = (R)a.eps2pi * sqrt(ki * ki + kj * kj + kk * kk);
          zeta = zeta_filter<R>(a.order, s);
        }
      }
      R g = (R)0;
      V K[3] = {cmake<V>(0, 0), cmake<V>(0, 0), cmake<V>(0, 0)};
      if (a.ker2) {
        // kernel 2: K = FFT(grad G) h^3, a complex half-spectrum
        if (col_ok) {
          const V *Kp = reinterpret_cast<const V *>(a.Ghat) + (((int64_t)n * a.Ly + ky) * a.Hx + kx) * 3;
          for (int c = 0; c < 3; ++c) K[c] = cscale(Kp[c], (R)a.h3);
        }
      } else if (a.allper) {
        R kmag = ki * ki + kj * kj + kk * kk;
        g = kmag > (R)0 ? zeta / ((R)4 * (pi * pi) * kmag) : (R)0;
      } else {
        g = col_ok ? G[((int64_t)kzf * a.Hy + kyf) * a.Hx + kx] * (R)a.h3 : (R)0;
      }
      V r;
      if (a.ker2 && kind == 0) {
        // u-hat = K x w-hat (poisson/pmlib_poisson_solve.f:898-900)
        const int c1 = comp == 2 ? 0 : comp + 1, c2 = comp == 0 ? 2 : comp - 1;
        r = csub(cmul(K[c1], w[c2][e]), cmul(K[c2], w[c1][e]));
      } else if (a.ker2) {
        // filtered vorticity + K_c 2 pi i (k . w-hat) (:903-910); a Nyquist
        // wavenumber drops out of the Hermitian part
        r = cscale(w[comp][e], zeta);
        if (a.reproj) {
          V acc = cmake<V>(0, 0);
          if (!nx) acc = cadd(acc, cscale(w[0][e], ki));
          if (!ny) acc = cadd(acc, cscale(w[1][e], kj));
          if (!nz) acc = cadd(acc, cscale(w[2][e], kk));
          r = cadd(r, cmul(K[comp], cmake<V>(-twopi * acc.y, twopi * acc.x)));
        }
      } else if (kind == 0) {
        // u-hat = G 2 pi i (k x w-hat); Nyquist wavenumbers drop out of the
        // real part of the inverse (Hermitian projection)
        R kx_ = nx ? (R)0 : ki, ky_ = ny ? (R)0 : kj, kz_ = nz ? (R)0 : kk;
        V cr;
        if (comp == 0) cr = csub(cscale(w[2][e], ky_), cscale(w[1][e], kz_));
        else if (comp == 1) cr = csub(cscale(w[0][e], kz_), cscale(w[2][e], kx_));
        else cr = csub(cscale(w[1][e], kx_), cscale(w[0][e], ky_));
        R f = g * twopi;
        r = cmake<V>(-cr.y * f, cr.x * f);
      } else if (kind == 1) {
        r = cscale(w[comp][e], g);
      } else {
        r = cscale(w[comp][e], zeta);
        if (a.reproj) {
          // + (2 pi i k_c) G 2 pi i (k . w-hat) = -4 pi^2 G k_c sum_d k_d w_d,
          // keeping only pairs (c, d) of equal Nyquist parity
          R kv[3] = {ki, kj, kk};
          bool nq[3] = {nx, ny, nz};
          V acc = cmake<V>(0, 0);
#pragma unroll
          for (int d = 0; d < 3; ++d)
            if (nq[d] == nq[comp]) acc = cadd(acc, cscale(w[d][e], kv[d]));
          R f = -(R)4 * (pi * pi) * g * kv[comp];
          r = cadd(r, cscale(acc, f));
        }
      }
      lds[myl * st + n] = r;
    }
    __syncthreads();
    fft_lds<1>(lds + myl * st, a.plan, p, tw);
    V *dst = kind == 0 ? B : (kind == 1 ? reinterpret_cast<V *>(a.Bpsi) : reinterpret_cast<V *>(a.Bvort));
    dst += comp * a.cs + (int64_t)ky * a.kys + (int64_t)kx0 * a.kxs;
    for (int idx = threadIdx.x; idx < T * a.Mz; idx += blockDim.x) {
      int row = idx / T, c = idx - row * T;
      if (kx0 + c < a.Hx) dst[(int64_t)row * a.rs + c * a.kxs] = cscale(lds[c * st + row], scale);
    }
    __syncthreads();
  }
}

// Regularised Green's function (kernel 1) on the kernel grid, written in
// FFT order (index n <-> grid point m = n or n - L, the fft_shift of
// fourier/pmlib_fourier_fft.f:139-150).  Cell centres follow
// topology/pmlib_topology_kernel.f:109-110: x = xk + ((i-1) + 1/2) dx with
// xk = -(L/2 + 1/2) dx, i = m + L/2 + 1; periodic axes sum images +-1.
struct GreenArgs {
  double *out;       // real, [(z * Ly + y) * Lx + x]
  int L[3];
  int bc[3];
  double dx[3], xk[3], len[3];
  double eps;
  int order;
  double ca;         // refinement patch: 1 / a, a = 2^(level - parent level); 0 on a base patch
  double hmax;       // max(dx): order 100 takes rho = pi r / hmax
};

// The sine integral Si(x) by the published cisi method (Numerical Recipes,
// sec. 6.9), with the truncation of poisson/pmlib_poisson_setup.f:1739-1841:
// for |x| > 2 the continued fraction of E1(ix) by the modified Lentz method
// to a relative 1e-10, Si = pi/2 + Im(e^{-it} h); otherwise the power series
// to 1e-10, odd and even terms accumulated alternately.
__device__ __forceinline__ double2 z_mul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 z_inv(double2 a) {
  const double m = a.x * a.x + a.y * a.y;
  return make_double2(a.x / m, -a.y / m);
}
__device__ inline double si_cisi(double x) {
  const double pi = 3.14159265358979323846, eps = 1.0e-10, tmin = 2.0;
  const double t = fabs(x);
  if (t == 0.0) return 0.0;
  double si;
  if (t > tmin) {
    double2 b = make_double2(1.0, t), c = make_double2(1.0e30, 0.0), d = z_inv(b), h = d;
    for (int i = 2; i <= 1000; ++i) {
      const double a = -(double)((i - 1) * (i - 1));
      b.x += 2.0;
      d = z_inv(make_double2(a * d.x + b.x, a * d.y + b.y));
      const double2 ac = z_inv(c);
      c = make_double2(b.x + a * ac.x, b.y + a * ac.y);
      const double2 del = z_mul(c, d);
      h = z_mul(h, del);
      if (fabs(del.x - 1.0) + fabs(del.y) <= eps) break;
    }
    h = z_mul(make_double2(cos(t), -sin(t)), h);
    si = 0.5 * pi + h.y;
  } else if (t < 1.0e-15) {
    si = t;
  } else {
    double sum = 0.0, sums = 0.0, sumc = 0.0, sign = 1.0, fact = 1.0;
    bool odd = true;
    for (int k = 1; k <= 1000; ++k) {
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

// poly(rho) E(rho) + F(rho) of the regularised kernels (order > 0):
// kernel 1 G = (poly E + F) / (4 pi r), kernel 2 grad G = -x (poly E + F) / (4 pi r^3)
__device__ __forceinline__ double green_part(int kernel, int order, double rho) {
  const double c_1_sqrt2 = 0.70710678118654752440;
  const double c_1_sqrt2pi = 0.39894228040143267794;
  const double r2 = rho * rho, r3 = r2 * rho, r5 = r3 * r2, r7 = r5 * r2, r9 = r7 * r2;
  double poly;
  if (kernel == 1) {
    switch (order) {
      case 2: poly = 0.0; break;
      case 4: poly = rho; break;
      case 6: poly = 1.75 * rho - 0.25 * r3; break;
      case 8: poly = 2.375 * rho + (-2.0 / 3.0) * r3 + (1.0 / 24.0) * r5; break;
      default: poly = 2.921875 * rho + (-233.0 / 192.0) * r3 + (29.0 / 192.0) * r5 + (-1.0 / 192.0) * r7; break;
    }
  } else {
    switch (order) {
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
  return poly * (exp(-0.5 * (rho * rho)) * c_1_sqrt2pi) + erf(c_1_sqrt2 * rho);
}

// G(0) / (sqrt(2) / (4 pi^1.5 eps)) of the kernel-1 orders 2..10
__device__ __forceinline__ double green_origin(int order) {
  switch (order) {
    case 2: return 1.0;
    case 4: return 1.5;
    case 6: return 1.875;
    case 8: return 2.1875;
    default: return 2.4609375;
  }
}

// Refinement-patch kernels (level > 1, poisson/pmlib_poisson_setup.f:446-528
// (order 2), :640-731 (4), :840-950 (6), :1060-1180 (8), :1300-1420 (10)):
// the kernel regularised at eps minus the one regularised at a eps, i.e.
// part(rho) - part(rho / a), at the grid point alone.  The reference adds the
// image offsets ii, jj, kk there outside the loops that set them; they are
// taken as 0 (no images).  Kernel 1 at r = 0: c (1 - 1/a) sqrt(2) / (4 pi^1.5
// eps); kernel 2 at r = 0: 0.
__device__ __forceinline__ void green_refine(const GreenArgs &a, int kernel, const double base[3], double *out) {
  const double pi = 3.14159265358979323846;
  const double c_1_4pi = 0.25 / pi;
  const double r = sqrt(base[0] * base[0] + base[1] * base[1] + base[2] * base[2]);
  if (r == 0.0) {
    if (kernel == 1) out[0] = green_origin(a.order) * (1.41421356237309504880 / (4.0 * pow(pi, 1.5) * a.eps)) *
                              (1.0 - a.ca);
    else out[0] = out[1] = out[2] = 0.0;
    return;
  }
  const double rho = r / a.eps;
  const double f = green_part(kernel, a.order, rho) - green_part(kernel, a.order, rho * a.ca);
  if (kernel == 1) {
    out[0] = c_1_4pi / r * f;
  } else {
    const double s = c_1_4pi / (r * r * r) * f;
    for (int c = 0; c < 3; ++c) out[c] = -base[c] * s;
  }
}

static __global__ void k_green(GreenArgs a) {
  const int64_t n = (int64_t)a.L[0] * a.L[1] * a.L[2];
  const double pi = 3.14159265358979323846;
  const double c_1_4pi = 0.25 / pi;
  const double c_1_sqrt2 = 0.70710678118654752440;
  const double c_1_sqrt2pi = 0.39894228040143267794;
  const double c_sq2_4pieps = 1.41421356237309504880 / (4.0 * pow(pi, 1.5) * a.eps);
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int ix = idx % a.L[0];
    int iy = (idx / a.L[0]) % a.L[1];
    int iz = idx / ((int64_t)a.L[0] * a.L[1]);
    int nn[3] = {ix, iy, iz};
    double base[3];
    for (int d = 0; d < 3; ++d) {
      int m = nn[d] < a.L[d] / 2 ? nn[d] : nn[d] - a.L[d];
      int i1 = m + a.L[d] / 2 + 1;
      base[d] = a.xk[d] + ((double)(i1 - 1) + 0.5) * a.dx[d];
    }
    if (a.ca > 0.0) {
      green_refine(a, 1, base, a.out + idx);
      continue;
    }
    double acc = 0.0;
    for (int kk = -a.bc[2]; kk <= a.bc[2]; ++kk)
      for (int jj = -a.bc[1]; jj <= a.bc[1]; ++jj)
        for (int ii = -a.bc[0]; ii <= a.bc[0]; ++ii) {
          double px = base[0] + a.len[0] * (double)ii;
          double py = base[1] + a.len[1] * (double)jj;
          double pz = base[2] + a.len[2] * (double)kk;
          double r = sqrt(px * px + py * py + pz * pz);
          double rho = r / a.eps;
          double v;
          if (a.order == 100) {
            // band-limited kernel (poisson/pmlib_poisson_setup.f:1419-1470)
            v = r == 0.0 ? 1.0 / (2.0 * pi * a.hmax) : si_cisi(pi * r / a.hmax) / (2.0 * (pi * pi) * r);
          } else if (r == 0.0) {
            switch (a.order) {
              case 0: v = 1.0; break;
              case 2: v = c_sq2_4pieps; break;
              case 4: v = 1.5 * c_sq2_4pieps; break;
              case 6: v = 1.875 * c_sq2_4pieps; break;
              case 8: v = 2.1875 * c_sq2_4pieps; break;
              default: v = 2.4609375 * c_sq2_4pieps; break;
            }
          } else if (a.order == 0) {
            v = c_1_4pi / r;
          } else {
            double E = exp(-0.5 * (rho * rho)) * c_1_sqrt2pi;
            double F = erf(c_1_sqrt2 * rho);
            double r2 = rho * rho, r3 = r2 * rho, r5 = r3 * r2, r7 = r5 * r2;
            double poly;
            switch (a.order) {
              case 2: poly = 0.0; break;
              case 4: poly = rho; break;
              case 6: poly = 1.75 * rho - 0.25 * r3; break;
              case 8: poly = 2.375 * rho + (-2.0 / 3.0) * r3 + (1.0 / 24.0) * r5; break;
              default:
                poly = 2.921875 * rho + (-233.0 / 192.0) * r3 + (29.0 / 192.0) * r5 + (-1.0 / 192.0) * r7;
                break;
            }
            v = c_1_4pi / r * (poly * E + F);
          }
          acc += v;
        }
    a.out[idx] = acc;
  }
}

// Regularised kernel 2 (grad G, the three components interleaved
// [idx * 3 + c]) on the same grid and in the same FFT order:
// -x_c / (4 pi r^3) (poly(rho) E + F), 0 at r = 0
// (poisson/pmlib_poisson_setup.f:580-635 (order 4), :287-320 (0), :389-430
// (2), :783-839 (6), :1001-1060 (8), :1231-1293 (10)).
static __global__ void k_green2(GreenArgs a) {
  const int64_t n = (int64_t)a.L[0] * a.L[1] * a.L[2];
  const double pi = 3.14159265358979323846;
  const double c_1_4pi = 0.25 / pi;
  const double c_1_sqrt2 = 0.70710678118654752440;
  const double c_1_sqrt2pi = 0.39894228040143267794;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int ix = idx % a.L[0];
    int iy = (idx / a.L[0]) % a.L[1];
    int iz = idx / ((int64_t)a.L[0] * a.L[1]);
    int nn[3] = {ix, iy, iz};
    double base[3];
    for (int d = 0; d < 3; ++d) {
      int m = nn[d] < a.L[d] / 2 ? nn[d] : nn[d] - a.L[d];
      int i1 = m + a.L[d] / 2 + 1;
      base[d] = a.xk[d] + ((double)(i1 - 1) + 0.5) * a.dx[d];
    }
    if (a.ca > 0.0) {
      green_refine(a, 2, base, a.out + idx * 3);
      continue;
    }
    double acc[3] = {0.0, 0.0, 0.0};
    for (int kk = -a.bc[2]; kk <= a.bc[2]; ++kk)
      for (int jj = -a.bc[1]; jj <= a.bc[1]; ++jj)
        for (int ii = -a.bc[0]; ii <= a.bc[0]; ++ii) {
          const double p[3] = {base[0] + a.len[0] * (double)ii, base[1] + a.len[1] * (double)jj,
                               base[2] + a.len[2] * (double)kk};
          const double r = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
          if (r == 0.0) continue;
          if (a.order == 100) {
            // x (sin rho - Si(rho)) / (2 pi^2 r^3), rho = pi r / h (:1472-1505)
            const double rb = pi * r / a.hmax;
            const double f = (sin(rb) - si_cisi(rb)) / (2.0 * (pi * pi) * (r * r * r));
            for (int c = 0; c < 3; ++c) acc[c] += p[c] * f;
            continue;
          }
          const double rho = r / a.eps;
          double f;
          if (a.order == 0) {
            f = 1.0;
          } else {
            const double r2 = rho * rho, r3 = r2 * rho, r5 = r3 * r2, r7 = r5 * r2, r9 = r7 * r2;
            double poly;
            switch (a.order) {
              case 2: poly = -2.0 * rho; break;
              case 4: poly = -2.0 * rho + r3; break;
              case 6: poly = -2.0 * rho + 2.25 * r3 - 0.25 * r5; break;
              case 8: poly = -2.0 * rho + (89.0 / 24.0) * r3 + (-20.0 / 24.0) * r5 + (1.0 / 24.0) * r7; break;
              default:
                poly = -2.0 * rho + (1027.0 / 192.0) * r3 + (-349.0 / 192.0) * r5 + (35.0 / 192.0) * r7 +
                       (-1.0 / 192.0) * r9;
                break;
            }
            f = poly * exp(-0.5 * (rho * rho)) * c_1_sqrt2pi + erf(c_1_sqrt2 * rho);
          }
          const double s = c_1_4pi / (r * r * r) * f;
          for (int c = 0; c < 3; ++c) acc[c] -= p[c] * s;
        }
    for (int c = 0; c < 3; ++c) a.out[idx * 3 + c] = acc[c];
  }
}

// Working-precision copy of the kernel-2 half-spectrum
template <typename R> __global__ void k_spec_convert(const double2 *in, cplx_t<R> *out, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = cmake<cplx_t<R>>((R)in[i].x, (R)in[i].y);
}

// Octant extraction of the real, even kernel spectrum:
// Ghat[(kz' * Hy + ky') * Hx + kx] = Re B[(kz' * Ly + ky') * Hx + kx]: kx
// fastest, as the z passes read it (consecutive lanes, consecutive kx).
template <typename R>
__global__ void k_g_octant(const double2 *B, R *G, int Hx, int Ly, int Hy, int Hz) {
  int64_t n = (int64_t)Hy * Hx * Hz;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int kx = idx % Hx;
    int64_t r = idx / Hx;
    int ky = r % Hy;
    int kz = r / Hy;
    G[idx] = (R)B[((int64_t)kz * Ly + ky) * Hx + kx].x;
  }
}

}  // namespace pmgpu
