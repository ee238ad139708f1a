// Poisson solve kernels specialised on the transform length (see
// poisson_kernels.h for the data path; argument structs are shared).
#pragma once
#include "pm_fft_fixed.h"
#include "poisson_kernels.h"

namespace pmgpu {
namespace fx {

// LDS layout of the x passes (padding mode, line stride in complex words),
// chosen with a bank model of their staging and transform accesses: for the
// hot half-lengths i + i / 16 padding with the modelled best stride; other
// lengths keep the general padding.
template <int NH, int VB> struct XLayout {
  static constexpr int PM = (NH == 256 || NH == 128) ? 2 : 0;
  static constexpr int ST = NH == 256 ? (VB == 8 ? 272 : 283) : (NH == 128 ? (VB == 8 ? 136 : 140) : Geo<NH>::ST);
};

// Line stride (complex words) of the register-resident kernels' LDS lines
// for T columns per block (thread tid: column tid % T, position tid / T):
// room for the padded line (i + i / 8), and g = 32 / T mod 32 words for
// 8-byte FP32 complex (16 / T mod 16 for 16-byte FP64), g being the number
// of positions a column has in one pass of the LDS (32 or 16 lanes), so the
// T columns' runs of g consecutive positions tile the banks.
template <int L, int VB, int T> constexpr int reg_line_stride() {
  const int mod = VB == 8 ? 32 : 16;
  const int g = (mod / T) % mod;
  int st = L + L / 8;
  while (st % mod != g) ++st;
  return st;
}

// ---------------------------------------------------------------- strided
// Kernel bodies take the block's logical tile lt and the tiles per grid row
// gx: the single-patch kernels derive them from the block index, the group
// kernels (end of file) from a table of patches.
template <typename R, int L, int T, int S>
__device__ __forceinline__ void strided_f_body(const StridedArgs &a, const cplx_t<R> *tw, int64_t lt, int gx) {
  using V = cplx_t<R>;
  constexpr int P = Geo<L>::P, ST = Geo<L>::ST, NT = P * T;
  __shared__ V lds[T * ST];
  const int col0 = (int)(lt % gx) * T;
  const int64_t b = lt / gx;
  const V *in = reinterpret_cast<const V *>(a.in) + b * a.bs_in + col0;
  V *out = reinterpret_cast<V *>(a.out) + b * a.bs_out + col0;
  const int tid = threadIdx.x;
  Twid<L, V> twd;
  twid_load<L>(twd, tid % P, tw + a.plan.tw_off);
  const int c_ld = tid % T, r0 = tid / T;  // NT / T == P rows per sweep
  const bool col_ok = col0 + c_ld < a.ncols;
  constexpr int E = (L + P - 1) / P;
  // issue every load of this thread before the first LDS write
  V v[E];
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int row = r0 + i * P;
    v[i] = cmake<V>(0, 0);
    if (row < a.nin && col_ok) v[i] = in[(int64_t)row * a.rs_in + c_ld];
  }
#pragma unroll
  for (int i = 0; i < E; ++i) {
    const int row = r0 + i * P;
    if (row < L) lds[c_ld * ST + pad(row)] = v[i];
  }
  __syncthreads();
  const int line = tid / P, p = tid % P;
  fft<L, S>(lds + line * ST, p, twd);
  __syncthreads();
  const R sc = (R)a.scale;
  if (col_ok)
    for (int row = tid / T; row < a.nout; row += NT / T)
      out[(int64_t)row * a.rs_out + c_ld] = cscale(lds[c_ld * ST + pad(row)], sc);
}
template <typename R, int L, int T, int S>
__global__ void __launch_bounds__(Geo<L>::P *T) k_strided_f(StridedArgs a, const cplx_t<R> *tw) {
  strided_f_body<R, L, T, S>(a, tw, xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y * gridDim.z), gridDim.x);
}

// Register-resident form: thread (col, p) owns rows p + P m of column
// col0 + col; loads and stores are T consecutive complex words per row across
// the lanes; first and last FFT stages on registers, LDS for the exchanges.
template <typename R, int L, int T, int S, bool HALF>
__device__ __forceinline__ void strided_r_body(const StridedArgs &a, const cplx_t<R> *tw, int64_t lt, int gx) {
  using V = cplx_t<R>;
  constexpr int P = Geo<L>::P, ST = reg_line_stride<L, sizeof(V), T>(), E = L / P;
  __shared__ V lds[T * ST];
  const int col0 = (int)(lt % gx) * T;
  const int64_t b = lt / gx;
  const int tid = threadIdx.x;
  const int col = tid % T, p = tid / T;
  Twid<L, V> twd;
  twid_load<L>(twd, p, tw + a.plan.tw_off);
  const bool col_ok = col0 + col < a.ncols;
  const V *in = reinterpret_cast<const V *>(a.in) + b * a.bs_in + col0 + col;
  V *out = reinterpret_cast<V *>(a.out) + b * a.bs_out + col0 + col;
  V x[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int row = p + P * m;
    x[m] = (col_ok && row < a.nin) ? in[(int64_t)row * a.rs_in] : cmake<V>(0, 0);
  }
  // HALF: forward with the upper half of the rows zero, or inverse with only
  // the lower half wanted
  fft_regs<L, S, V, HALF && (S < 0), HALF && (S > 0), 1>(x, lds + col * ST, p, twd);
  const R sc = (R)a.scale;
  if (col_ok) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int row = p + P * m;
      if (row < a.nout) out[(int64_t)row * a.rs_out] = cscale(x[m], sc);
    }
  }
}
template <typename R, int L, int T, int S, bool HALF>
__global__ void __launch_bounds__(Geo<L>::P *T) k_strided_r(StridedArgs a, const cplx_t<R> *tw) {
  strided_r_body<R, L, T, S, HALF>(a, tw, xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y * gridDim.z),
                                   gridDim.x);
}

// -------------------------------------------------------------- x forward
// Lines of ncomp (= 3 or 1) components; NH = half transform length.
template <typename R, int NH, int T, int NC>
__device__ __forceinline__ void x_fwd_body(const XFwdArgs &a, const cplx_t<R> *tw, int64_t lt, int gx) {
  using V = cplx_t<R>;
  constexpr int P = Geo<NH>::P, ST = XLayout<NH, sizeof(V)>::ST, PM = XLayout<NH, sizeof(V)>::PM, NL = T * NC,
                NT = P * NL;
  __shared__ V lds[NL * ST];
  const int y0 = (int)(lt % gx) * T, z = (int)(lt / gx);
  const R *in = reinterpret_cast<const R *>(a.in);
  const int tid = threadIdx.x;
  Twid<NH, V> twd;
  twid_load<NH>(twd, tid % P, tw + a.plan.tw_off);
  // packed pairs x[2m] + i x[2m+1]; m >= ceil(Min/2) are zero
  const int half = (a.Min + 1) / 2;
  if ((NC == 1 && a.xs == 1) || (a.cs == 1 && a.xs == NC)) {
    // the NC components of a line are one contiguous row of NC*Min words:
    // read it word by word across the lanes and scatter into the LDS lines
    // When a row is W = J NT words, word j NT + tid of row t is component
    // tid % NC of element x = j NT / NC + tid / NC: no division in the loop
    // (NT is a multiple of NC).  Otherwise the general indexing.
    constexpr int W = NC * 2 * NH;
    R *ldsr = reinterpret_cast<R *>(lds);
    if constexpr (W % NT == 0) {
      constexpr int J = W / NT, K = T * J;
      const int c = NC == 1 ? 0 : tid % NC, xb = NC == 1 ? tid : tid / NC;
      R v[K];
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int t = i / J, j = i % J;
        const int w = j * NT + tid;
        const int y = y0 + t;
        v[i] = (R)0;
        if (w < NC * a.Min && y < a.My) v[i] = in[(int64_t)y * a.ys + (int64_t)z * a.zs + w];
      }
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int t = i / J, j = i % J;
        const int x = j * (NT / NC) + xb;
        ldsr[2 * ((t * NC + c) * ST + padm<PM>(x >> 1)) + (x & 1)] = v[i];
      }
    } else {
      constexpr int K = (T * W + NT - 1) / NT;
      R v[K];
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int idx = tid + i * NT;
        const int t = idx / W, w = idx - t * W;
        const int y = y0 + t;
        v[i] = (R)0;
        if (t < T && w < NC * a.Min && y < a.My) v[i] = in[(int64_t)y * a.ys + (int64_t)z * a.zs + w];
      }
#pragma unroll
      for (int i = 0; i < K; ++i) {
        const int idx = tid + i * NT;
        const int t = idx / W, w = idx - t * W;
        if (t < T) {
          const int cc = w % NC, x = w / NC;
          ldsr[2 * ((t * NC + cc) * ST + padm<PM>(x >> 1)) + (x & 1)] = v[i];
        }
      }
    }
  } else
  for (int idx = tid; idx < NL * NH; idx += NT) {
    int l = idx / NH, m = idx - l * NH;
    int t = l / NC, c = l - t * NC;
    int y = y0 + t;
    V v = cmake<V>(0, 0);
    if (m < half && y < a.My) {
      const R *src = in + (int64_t)c * a.cs + (int64_t)y * a.ys + (int64_t)z * a.zs;
      int x0 = 2 * m;
      v.x = src[(int64_t)x0 * a.xs];
      v.y = x0 + 1 < a.Min ? src[(int64_t)(x0 + 1) * a.xs] : (R)0;
    }
    lds[l * ST + padm<PM>(m)] = v;
  }
  __syncthreads();
  const int line = tid / P, p = tid % P;
  fft<NH, -1, V, PM>(lds + line * ST, p, twd);
  __syncthreads();
  V *out = reinterpret_cast<V *>(a.out);
  const V *twL = tw + a.twL_off;
  constexpr int HX = NH + 1;
  // c fastest: with the interleaved spectral layout consecutive threads
  // write consecutive complex words (thread: component tid % NC, wavenumbers
  // tid / NC + n NT / NC)
  const int pc = NC == 1 ? 0 : tid % NC, pk = NC == 1 ? tid : tid / NC;
  auto post = [&](int t, int k) {
    const V *zl = lds + (t * NC + pc) * ST;
    V zk = zl[padm<PM>(k == NH ? 0 : k)];
    V zm = cconj(zl[padm<PM>(k == 0 ? 0 : NH - k)]);
    V e = cscale(cadd(zk, zm), (R)0.5);
    V o = cscale(csub(zk, zm), (R)0.5);
    V wo = cmul(twL[k], o);
    out[pc * a.o_cs + k * a.o_ks + (int64_t)(y0 + t) * a.o_ys + (int64_t)z * a.o_zs] = cadd(e, cmake<V>(wo.y, -wo.x));
  };
  if constexpr (NT / NC > HX) {
    // more threads per component than wavenumbers: walk (line, k) flattened
    for (int idx = pk; idx < T * HX; idx += NT / NC) {
      const int t = idx / HX, k = idx - t * HX;
      if (y0 + t >= a.My) break;
      post(t, k);
    }
  } else {
    for (int t = 0; t < T; ++t) {
      if (y0 + t >= a.My) break;
      for (int k = pk; k < HX; k += NT / NC) post(t, k);
    }
  }
}
template <typename R, int NH, int T, int NC>
__global__ void __launch_bounds__(Geo<NH>::P *T *NC) k_x_fwd_f(XFwdArgs a, const cplx_t<R> *tw) {
  x_fwd_body<R, NH, T, NC>(a, tw, blockIdx.x + (int64_t)gridDim.x * blockIdx.y, gridDim.x);
}

// -------------------------------------------------------------- x inverse
template <typename R, int NH, int T, int NC>
__device__ __forceinline__ void x_inv_body(const XInvArgs &a, const cplx_t<R> *tw, int64_t lt, int gx) {
  using V = cplx_t<R>;
  constexpr int P = Geo<NH>::P, ST = XLayout<NH, sizeof(V)>::ST, PM = XLayout<NH, sizeof(V)>::PM, NL = T * NC,
                NT = P * NL;
  __shared__ V lds[NL * ST];
  const int y0 = (int)(lt % gx) * T, z = (int)(lt / gx);
  const V *in = reinterpret_cast<const V *>(a.in);
  const V *twL = tw + a.twL_off;
  const int tid = threadIdx.x;
  Twid<NH, V> twd;
  twid_load<NH>(twd, tid % P, tw + a.plan.tw_off);
  // thread: component tid % NC, wavenumbers tid / NC + n NT / NC
  const int pc = NC == 1 ? 0 : tid % NC, pk = NC == 1 ? tid : tid / NC;
  auto pre = [&](int t, int k) {
    const int y = y0 + t;
    V Zk = cmake<V>(0, 0);
    if (y < a.My) {
      const V *xl = in + pc * a.i_cs + (int64_t)y * a.i_ys + (int64_t)z * a.i_zs;
      V xk = xl[k * a.i_ks], xm = xl[(NH - k) * a.i_ks];
      if (k == 0) {
        xk.y = 0;
        xm.y = 0;
      }
      V s = cadd(xk, cconj(xm));
      V d = csub(xk, cconj(xm));
      V wd = cmul(cconj(twL[k]), d);
      Zk = cadd(s, cmake<V>(-wd.y, wd.x));
    }
    lds[(t * NC + pc) * ST + padm<PM>(k)] = Zk;
  };
  if constexpr (NT / NC > NH) {
    // more threads per component than wavenumbers: walk (line, k) flattened
    for (int idx = pk; idx < T * NH; idx += NT / NC) {
      const int t = idx / NH;
      pre(t, idx - t * NH);
    }
  } else {
    for (int t = 0; t < T; ++t)
      for (int k = pk; k < NH; k += NT / NC) pre(t, k);
  }
  __syncthreads();
  const int line = tid / P, p = tid % P;
  fft<NH, 1, V, PM>(lds + line * ST, p, twd);
  __syncthreads();
  R *out = reinterpret_cast<R *>(a.out);
  const R sc = (R)a.scale;
  const int halfo = (a.Mout + 1) / 2;
  if ((NC == 1 && a.xs == 1) || (a.cs == 1 && a.xs == NC)) {
    // when row t is W = J NT words, word j NT + tid is component tid % NC
    // of element j NT / NC + tid / NC (no division in the loop)
    constexpr int W = NC * 2 * NH;
    const R *ldsr = reinterpret_cast<const R *>(lds);
    if constexpr (W % NT == 0) {
      constexpr int J = W / NT;
      const int c = NC == 1 ? 0 : tid % NC, xb = NC == 1 ? tid : tid / NC;
#pragma unroll
      for (int i = 0; i < T * J; ++i) {
        const int t = i / J, j = i % J;
        const int w = j * NT + tid;
        const int y = y0 + t;
        if (w >= NC * a.Mout || y >= a.My) continue;
        const int x = j * (NT / NC) + xb;
        out[(int64_t)y * a.ys + (int64_t)z * a.zs + w] = ldsr[2 * ((t * NC + c) * ST + padm<PM>(x >> 1)) + (x & 1)] * sc;
      }
    } else {
      for (int idx = tid; idx < T * W; idx += NT) {
        const int t = idx / W, w = idx - t * W;
        const int y = y0 + t;
        if (w >= NC * a.Mout || y >= a.My) continue;
        const int c = w % NC, x = w / NC;
        out[(int64_t)y * a.ys + (int64_t)z * a.zs + w] = ldsr[2 * ((t * NC + c) * ST + padm<PM>(x >> 1)) + (x & 1)] * sc;
      }
    }
    return;
  }
  for (int idx = tid; idx < NL * halfo; idx += NT) {
    int l = idx / halfo, m = idx - l * halfo;
    int t = l / NC, c = l - t * NC;
    int y = y0 + t;
    if (y >= a.My) continue;
    V v = lds[l * ST + padm<PM>(m)];
    R *dst = out + (int64_t)c * a.cs + (int64_t)y * a.ys + (int64_t)z * a.zs;
    int x0 = 2 * m;
    dst[(int64_t)x0 * a.xs] = v.x * sc;
    if (x0 + 1 < a.Mout) dst[(int64_t)(x0 + 1) * a.xs] = v.y * sc;
  }
}
template <typename R, int NH, int T, int NC>
__global__ void __launch_bounds__(Geo<NH>::P *T *NC) k_x_inv_f(XInvArgs a, const cplx_t<R> *tw) {
  x_inv_body<R, NH, T, NC>(a, tw, blockIdx.x + (int64_t)gridDim.x * blockIdx.y, gridDim.x);
}

// ------------------------------------------------------------ fused z pass
// Component-interleaved spectra B[((z*Ly + ky)*Hx + kx)*3 + c]: a block owns
// TC consecutive kx columns of one ky row, i.e. 3*TC contiguous complex words
// per z row, and transforms the 3*TC component lines together.  The curl at
// (kx, ky, kz) reads the other two components from LDS, so no spectrum has to
// stay in registers across an FFT (only the thread's own forward values when
// psi or the filtered vorticity are also wanted).
template <typename R, int L, int TC, bool EXTRA>
__global__ void __launch_bounds__(Geo<L>::P * 3 * TC) k_z3_f(ZArgs a, const cplx_t<R> *tw_all) {
  using V = cplx_t<R>;
  constexpr int P = Geo<L>::P, ST = Geo<L>::ST;
  constexpr int NLN = 3 * TC, NT = P * NLN;
  constexpr int E = (L + P - 1) / P;
  constexpr bool FULL = E * P == L;
  __shared__ V lds[NLN * ST];
  const V *tw = tw_all + a.plan.tw_off;
  const int64_t lt = xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y);
  const int kx0 = (int)(lt % gridDim.x) * TC, ky = (int)(lt / gridDim.x);
  const int tid = threadIdx.x;
  const int line = tid / P, p = tid % P;
  Twid<L, V> twd;
  twid_load<L>(twd, p, tw);
  const int col = line / 3, comp = line - 3 * col;
  const int kx = kx0 + col;
  const bool col_ok = kx < a.Hx;
  const int ncl = (a.Hx - kx0 < TC ? a.Hx - kx0 : TC) * 3;  // valid lines in this block
  V *B = reinterpret_cast<V *>(a.B);
  const int64_t tile = (int64_t)ky * a.kys + (int64_t)kx0 * 3;  // kys = padded row
  V *myl = lds + line * ST;
  // load the 3*TC interleaved lines, zero rows >= Mz; thread tid owns line
  // word tid % NLN of rows tid / NLN + i P (NT = P * NLN), and issues all of
  // its loads before the first LDS write
  {
    const int cl = tid % NLN, r0 = tid / NLN;
    V v[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int row = r0 + i * P;
      v[i] = cmake<V>(0, 0);
      if (row < a.Mz && cl < ncl) v[i] = B[tile + (int64_t)row * a.rs + cl];
    }
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int row = r0 + i * P;
      if (row < L) lds[cl * ST + pad(row)] = v[i];
    }
  }
  __syncthreads();
  fft<L, -1>(myl, p, twd);
  __syncthreads();
  const R ki = wavenumber<R>(kx, a.Lx, a.dk[0], a.ksmp[0]);
  const R kj = wavenumber<R>(ky, a.Ly, a.dk[1], a.ksmp[1]);
  const bool nx = (2 * kx == a.Lx), ny = (2 * ky == a.Ly);
  const int kyf = ky <= a.Ly - ky ? ky : a.Ly - ky;
  const R pi = (R)3.14159265358979323846;
  const R twopi = (R)6.28318530717958647692;
  const R *G = reinterpret_cast<const R *>(a.Ghat);
  const R scale = (R)1 / (R)L;
  // EXTRA = psi and/or filtered vorticity requested (else velocity only)
  const V *cl3 = lds + (3 * col) * ST;  // the three component lines of my column
  V own[EXTRA ? E : 1];                // my component's forward spectrum (extra outputs)
  const int nkind = EXTRA ? 1 + (a.want_psi ? 1 : 0) + (a.want_vort ? 1 : 0) : 1;
  for (int kd = 0; kd < nkind; ++kd) {
    const int kind = kd == 0 ? 0 : ((kd == 1 && a.want_psi) ? 1 : 2);
    if (EXTRA && kd > 0 && kind == 2) {
      // the filtered vorticity needs all three forward spectra again
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int n = p + e * P;
        if (FULL || n < L) myl[pad(n)] = own[EXTRA ? e : 0];
      }
      __syncthreads();
    }
    V r[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int n = p + e * P;
      r[e] = cmake<V>(0, 0);
      if (!(FULL || n < L)) continue;
      const R kk = wavenumber<R>(n, L, a.dk[2], a.ksmp[2]);
      const bool nz = (2 * n == L);
      R g = (R)0;
      R zeta = (R)1;
      V K0 = cmake<V>(0, 0), K1 = K0, K2 = K0;
      if (kind == 2 || a.allper)
        if (a.want_vort) zeta = zeta_filter<R>(a.order, (R)a.eps2pi * sqrt(ki * ki + kj * kj + kk * kk));
      if (a.ker2) {
        // kernel 2: K = FFT(grad G) h^3, a complex half-spectrum
        if (col_ok) {
          const V *Kp = reinterpret_cast<const V *>(a.Ghat) + (((int64_t)n * a.Ly + ky) * a.Hx + kx) * 3;
          K0 = cscale(Kp[0], (R)a.h3);
          K1 = cscale(Kp[1], (R)a.h3);
          K2 = cscale(Kp[2], (R)a.h3);
        }
      } else if (a.allper) {
        const R kmag = ki * ki + kj * kj + kk * kk;
        g = kmag > (R)0 ? zeta / ((R)4 * (pi * pi) * kmag) : (R)0;
      } else {
        const int kzf = n <= L - n ? n : L - n;
        g = col_ok ? G[((int64_t)kzf * a.Hy + kyf) * a.Hx + kx] * (R)a.h3 : (R)0;
      }
      if (EXTRA && kind == 1) {
        r[e] = cscale(own[EXTRA ? e : 0], g);
        continue;
      }
      const V w0 = cl3[pad(n)], w1 = cl3[ST + pad(n)], w2 = cl3[2 * ST + pad(n)];
      const V wc = comp == 0 ? w0 : (comp == 1 ? w1 : w2);
      if (EXTRA && kd == 0) own[EXTRA ? e : 0] = wc;
      if (a.ker2) {
        if (!EXTRA || kind == 0) {
          // u-hat = K x w-hat (poisson/pmlib_poisson_solve.f:898-900)
          if (comp == 0) r[e] = csub(cmul(K1, w2), cmul(K2, w1));
          else if (comp == 1) r[e] = csub(cmul(K2, w0), cmul(K0, w2));
          else r[e] = csub(cmul(K0, w1), cmul(K1, w0));
        } else {
          V q = cscale(wc, zeta);
          if (a.reproj) {
            // + K_c 2 pi i (k . w-hat) (:903-910); a Nyquist wavenumber drops
            // out of the Hermitian part
            V acc = cmake<V>(0, 0);
            if (!nx) acc = cadd(acc, cscale(w0, ki));
            if (!ny) acc = cadd(acc, cscale(w1, kj));
            if (!nz) acc = cadd(acc, cscale(w2, kk));
            const V Kc = comp == 0 ? K0 : (comp == 1 ? K1 : K2);
            q = cadd(q, cmul(Kc, cmake<V>(-twopi * acc.y, twopi * acc.x)));
          }
          r[e] = q;
        }
      } else if (!EXTRA || kind == 0) {
        // u-hat = G 2 pi i (k x w-hat); Nyquist wavenumbers drop out of the
        // real part of the inverse (Hermitian projection)
        const R kx_ = nx ? (R)0 : ki, ky_ = ny ? (R)0 : kj, kz_ = nz ? (R)0 : kk;
        V cr;
        if (comp == 0) cr = csub(cscale(w2, ky_), cscale(w1, kz_));
        else if (comp == 1) cr = csub(cscale(w0, kz_), cscale(w2, kx_));
        else cr = csub(cscale(w1, kx_), cscale(w0, ky_));
        const R f = g * twopi;
        r[e] = cmake<V>(-cr.y * f, cr.x * f);
      } else {
        V q = cscale(wc, zeta);
        if (a.reproj) {
          // -4 pi^2 G k_c sum_d k_d w_d over pairs of equal Nyquist parity
          const bool nc = comp == 0 ? nx : (comp == 1 ? ny : nz);
          const R kc = comp == 0 ? ki : (comp == 1 ? kj : kk);
          V acc = cmake<V>(0, 0);
          if (nx == nc) acc = cadd(acc, cscale(w0, ki));
          if (ny == nc) acc = cadd(acc, cscale(w1, kj));
          if (nz == nc) acc = cadd(acc, cscale(w2, kk));
          q = cadd(q, cscale(acc, -(R)4 * (pi * pi) * g * kc));
        }
        r[e] = q;
      }
    }
    __syncthreads();  // every read of the forward spectra is done
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int n = p + e * P;
      if (FULL || n < L) myl[pad(n)] = r[e];
    }
    line_sync<Geo<L>::WAVE>();
    fft<L, 1>(myl, p, twd);
    __syncthreads();
    V *dst = (kind == 0 ? B : (kind == 1 ? reinterpret_cast<V *>(a.Bpsi) : reinterpret_cast<V *>(a.Bvort))) + tile;
    for (int idx = tid; idx < a.Mz * NLN; idx += NT) {
      int row = idx / NLN, cl = idx - row * NLN;
      if (cl < ncl) dst[(int64_t)row * a.rs + cl] = cscale(lds[cl * ST + pad(row)], scale);
    }
    __syncthreads();
  }
}

// ------------------------------------------- register-resident fused z pass
// Buffer-resource loads (vector memory, one 32-bit VGPR offset per thread
// plus a scalar offset per row): the resource's size bounds every access.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float2 buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff, float2) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ double2 buf_load(__amdgpu_buffer_rsrc_t r, int voff, int soff, double2) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}

// Velocity only.  Thread (col, p) of a block owns kx column kx0 + col (of TC
// consecutive ones) and the z rows p + P m of all three components: its loads
// and stores are whole cache lines across the block's lanes (consecutive
// columns, components interleaved), the first and last FFT stages run on the
// registers, and the curl is formed in registers at the thread's own
// wavenumbers; LDS only carries the exchanges between FFT stages, one
// component at a time.  The host guarantees Mz * rs complex words fit a
// 31-bit byte offset.
// HALF (free-space z, Mz = L / 2): rows >= L / 2 are zero on the way in and
// not wanted on the way out, so the first forward and the last inverse stage
// skip them.
template <typename R, int L, int TC, bool ALLPER, bool HALF>
__global__ void __launch_bounds__(Geo<L>::P *TC) k_z3_r(ZArgs a, const cplx_t<R> *tw_all) {
  using V = cplx_t<R>;
  constexpr int P = Geo<L>::P, E = L / P;
  // line stride between the block's columns for the i + i / 8 padding
  constexpr int ST = reg_line_stride<L, sizeof(V), TC>();
  __shared__ V lds[TC * ST];
  const V *tw = tw_all + a.plan.tw_off;
  const int64_t lt = xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y);
  const int kx0 = (int)(lt % gridDim.x) * TC, ky = (int)(lt / gridDim.x);
  const int tid = threadIdx.x;
  const int col = tid % TC, p = tid / TC;
  Twid<L, V> twd;
  twid_load<L>(twd, p, tw);
  const int kx = kx0 + col;
  const bool col_ok = kx < a.Hx;
  // this block's tile: rows [0, Mz) of columns kx0 .. kx0 + TC - 1 (3 words each)
  V *base = reinterpret_cast<V *>(a.B) + (int64_t)ky * a.kys + (int64_t)kx0 * 3;
  const int64_t rsb = a.rs * (int64_t)sizeof(V);
  const int ncols_here = a.Hx - kx0 < TC ? a.Hx - kx0 : TC;
  const uint32_t extent = (uint32_t)((a.Mz - 1) * rsb + (int64_t)ncols_here * 3 * sizeof(V));
  const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(base, extent);
  const int voff = (int)(p * rsb) + col * 3 * (int)sizeof(V);
  V w[3][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const bool ok = col_ok && (HALF ? m < E / 2 : p + P * m < a.Mz);
    const int soff = (int)((int64_t)P * m * rsb);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      w[c][m] = ok ? buf_load(rsrc, voff + c * (int)sizeof(V), soff, V{}) : cmake<V>(0, 0);
  }
  V *line = lds + col * ST;
#pragma unroll
  for (int c = 0; c < 3; ++c) fft_regs<L, -1, V, HALF, false, 1>(w[c], line, p, twd);
  // u-hat = G 2 pi i (k x w-hat); Nyquist wavenumbers drop out (Hermitian
  // projection of the real inverse); the 1 / L of the inverse transform
  // rides on the same factor
  const R ki = wavenumber<R>(kx, a.Lx, a.dk[0], a.ksmp[0]);
  const R kj = wavenumber<R>(ky, a.Ly, a.dk[1], a.ksmp[1]);
  const R kx_ = (2 * kx == a.Lx) ? (R)0 : ki, ky_ = (2 * ky == a.Ly) ? (R)0 : kj;
  const int kyf = ky <= a.Ly - ky ? ky : a.Ly - ky;
  const R twopi = (R)6.28318530717958647692 / (R)L;
  // G-hat of the thread's rows, all loads issued before any is used (a
  // column past Hx reads column 0 and is masked)
  const R *G = reinterpret_cast<const R *>(a.Ghat) + (int64_t)kyf * a.Hx + (col_ok ? kx : 0);
  const int64_t gz = (int64_t)a.Hy * a.Hx;
  R gv[E];
  if constexpr (!ALLPER) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = p + P * m;
      const int kzf = n <= L - n ? n : L - n;
      gv[m] = G[kzf * gz];
    }
  }
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const int n = p + P * m;
    const R kk = wavenumber<R>(n, L, a.dk[2], a.ksmp[2]);
    R g;
    if constexpr (ALLPER) {
      const R kmag = ki * ki + kj * kj + kk * kk;
      const R pi = (R)3.14159265358979323846;
      g = kmag > (R)0 ? (R)1 / ((R)4 * (pi * pi) * kmag) : (R)0;
    } else {
      g = col_ok ? gv[m] * (R)a.h3 : (R)0;
    }
    const R kz_ = (2 * n == L) ? (R)0 : kk;
    const V w0 = w[0][m], w1 = w[1][m], w2 = w[2][m];
    const V c0 = csub(cscale(w2, ky_), cscale(w1, kz_));
    const V c1 = csub(cscale(w0, kz_), cscale(w2, kx_));
    const V c2 = csub(cscale(w1, kx_), cscale(w0, ky_));
    const R f = g * twopi;
    w[0][m] = cmake<V>(-c0.y * f, c0.x * f);
    w[1][m] = cmake<V>(-c1.y * f, c1.x * f);
    w[2][m] = cmake<V>(-c2.y * f, c2.x * f);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) fft_regs<L, 1, V, false, HALF, 1>(w[c], line, p, twd);
  // Stores are plain global stores: buffer-resource stores of these
  // 8-byte-aligned records, which the compiler merges into 16-byte stores
  // across components, have returned corrupted first components.
  if (col_ok) {
    V *dst = base + (int64_t)p * a.rs + col * 3;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      if (HALF ? m < E / 2 : p + P * m < a.Mz) {
#pragma unroll
        for (int c = 0; c < 3; ++c) dst[(int64_t)P * m * a.rs + c] = w[c][m];
      }
    }
  }
}

// Full-output form of k_z3_r: velocity plus the stream function and/or the
// filtered (and optionally reprojected) vorticity, kernel 1.  The three
// forward spectra stay in registers; each requested field is formed from
// them in registers, transformed back (one component at a time through the
// same LDS line) and stored, velocity in place, psi-hat and zeta-hat into
// their own spectra (poisson/pmlib_poisson_solve.f:767-872: psi = G w-hat,
// vort = zeta w-hat [+ reprojection]).  With every axis periodic the
// spectral G carries the filter zeta as well when the vorticity is filtered
// (:339-349).
// SL: the forward spectra wait in LDS (each thread reads back only the words
// it wrote, so no barrier) instead of registers, for FP64's register budget.
template <typename R, int L, int TC, bool ALLPER, bool HALF, bool SL>
__global__ void __launch_bounds__(Geo<L>::P *TC) k_z3_rx(ZArgs a, const cplx_t<R> *tw_all) {
  using V = cplx_t<R>;
  constexpr int P = Geo<L>::P, E = L / P;
  constexpr int ST = reg_line_stride<L, sizeof(V), TC>();
  __shared__ V lds[TC * ST];
  __shared__ V spec[SL ? TC * 3 * L : 1];
  const V *tw = tw_all + a.plan.tw_off;
  const int64_t lt = xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y);
  const int kx0 = (int)(lt % gridDim.x) * TC, ky = (int)(lt / gridDim.x);
  const int tid = threadIdx.x;
  const int col = tid % TC, p = tid / TC;
  Twid<L, V> twd;
  twid_load<L>(twd, p, tw);
  const int kx = kx0 + col;
  const bool col_ok = kx < a.Hx;
  const int64_t toff = (int64_t)ky * a.kys + (int64_t)kx0 * 3;
  V *base = reinterpret_cast<V *>(a.B) + toff;
  const int64_t rsb = a.rs * (int64_t)sizeof(V);
  const int ncols_here = a.Hx - kx0 < TC ? a.Hx - kx0 : TC;
  const uint32_t extent = (uint32_t)((a.Mz - 1) * rsb + (int64_t)ncols_here * 3 * sizeof(V));
  const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(base, extent);
  const int voff = (int)(p * rsb) + col * 3 * (int)sizeof(V);
  V w[3][E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const bool ok = col_ok && (HALF ? m < E / 2 : p + P * m < a.Mz);
    const int soff = (int)((int64_t)P * m * rsb);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      w[c][m] = ok ? buf_load(rsrc, voff + c * (int)sizeof(V), soff, V{}) : cmake<V>(0, 0);
  }
  V *line = lds + col * ST;
#pragma unroll
  for (int c = 0; c < 3; ++c) fft_regs<L, -1, V, HALF, false, 1>(w[c], line, p, twd);
  V *sp = spec + (SL ? col * 3 * L + p : 0);   // word (c, m) at sp[(c E + m) P]
  if constexpr (SL) {
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int m = 0; m < E; ++m) sp[(c * E + m) * P] = w[c][m];
  }
  auto spw = [&](int c, int m) -> V { return SL ? sp[(c * E + m) * P] : w[c][m]; };
  const R ki = wavenumber<R>(kx, a.Lx, a.dk[0], a.ksmp[0]);
  const R kj = wavenumber<R>(ky, a.Ly, a.dk[1], a.ksmp[1]);
  const bool nx = (2 * kx == a.Lx), ny = (2 * ky == a.Ly);
  const R kx_ = nx ? (R)0 : ki, ky_ = ny ? (R)0 : kj;
  const int kyf = ky <= a.Ly - ky ? ky : a.Ly - ky;
  const R pi = (R)3.14159265358979323846;
  const R twopi = (R)6.28318530717958647692;
  // per row: G h^3 (with zeta when all axes are periodic and the vorticity
  // is filtered) and zeta
  R gz[E], zt[E];
  {
    const R *G = reinterpret_cast<const R *>(a.Ghat) + (int64_t)kyf * a.Hx + (col_ok ? kx : 0);
    const int64_t gzs = (int64_t)a.Hy * a.Hx;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = p + P * m;
      const R kk = wavenumber<R>(n, L, a.dk[2], a.ksmp[2]);
      zt[m] = a.want_vort ? zeta_filter<R>(a.order, (R)a.eps2pi * sqrt(ki * ki + kj * kj + kk * kk)) : (R)1;
      if constexpr (ALLPER) {
        const R kmag = ki * ki + kj * kj + kk * kk;
        gz[m] = kmag > (R)0 ? zt[m] / ((R)4 * (pi * pi) * kmag) : (R)0;
      } else {
        const int kzf = n <= L - n ? n : L - n;
        gz[m] = col_ok ? G[kzf * gzs] * (R)a.h3 : (R)0;
      }
    }
  }
  const R scale = (R)1 / (R)L;
  const int nkind = 1 + (a.want_psi ? 1 : 0) + (a.want_vort ? 1 : 0);
  // one output component at a time: only its E words are live beside the
  // three forward spectra
#pragma unroll 1
  for (int kd = 0; kd < nkind; ++kd) {
    const int kind = kd == 0 ? 0 : ((kd == 1 && a.want_psi) ? 1 : 2);
    V *out = (kind == 0 ? reinterpret_cast<V *>(a.B)
                        : (kind == 1 ? reinterpret_cast<V *>(a.Bpsi) : reinterpret_cast<V *>(a.Bvort))) + toff;
#pragma unroll 1
    for (int c = 0; c < 3; ++c) {
      V r[E];
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const int n = p + P * m;
        const R kk = wavenumber<R>(n, L, a.dk[2], a.ksmp[2]);
        const bool nz = (2 * n == L);
        const V w0 = spw(0, m), w1 = spw(1, m), w2 = spw(2, m);
        const V wc = c == 0 ? w0 : (c == 1 ? w1 : w2);
        if (kind == 0) {
          // u-hat = G 2 pi i (k x w-hat), Nyquist wavenumbers dropped
          const R kz_ = nz ? (R)0 : kk;
          V cr;
          if (c == 0) cr = csub(cscale(w2, ky_), cscale(w1, kz_));
          else if (c == 1) cr = csub(cscale(w0, kz_), cscale(w2, kx_));
          else cr = csub(cscale(w1, kx_), cscale(w0, ky_));
          const R f = gz[m] * twopi;
          r[m] = cmake<V>(-cr.y * f, cr.x * f);
        } else if (kind == 1) {
          r[m] = cscale(wc, gz[m]);
        } else {
          V q = cscale(wc, zt[m]);
          if (a.reproj) {
            // -4 pi^2 G k_c sum_d k_d w_d over the pairs (c, d) of equal
            // Nyquist parity (:861-872)
            const bool nc = c == 0 ? nx : (c == 1 ? ny : nz);
            const R kc = c == 0 ? ki : (c == 1 ? kj : kk);
            V acc = cmake<V>(0, 0);
            if (nx == nc) acc = cadd(acc, cscale(w0, ki));
            if (ny == nc) acc = cadd(acc, cscale(w1, kj));
            if (nz == nc) acc = cadd(acc, cscale(w2, kk));
            q = cadd(q, cscale(acc, -(R)4 * (pi * pi) * gz[m] * kc));
          }
          r[m] = q;
        }
      }
      fft_regs<L, 1, V, false, HALF, 1>(r, line, p, twd);
      if (col_ok) {
        V *dst = out + (int64_t)p * a.rs + col * 3 + c;
#pragma unroll
        for (int m = 0; m < E; ++m)
          if (HALF ? m < E / 2 : p + P * m < a.Mz) dst[(int64_t)P * m * a.rs] = cscale(r[m], scale);
      }
    }
  }
}

// ------------------------------------------------------- scalar z pass
// One component (psi-hat = G w-hat, Cin = Cout = 1): a block owns T
// consecutive kx columns of one ky row, B[((z*Ly + ky)*rowp + kx)], and
// transforms them forward, multiplies by G h^3 (or 1/(4 pi^2 |k|^2) when all
// axes are periodic) and transforms back, in place.
template <typename R, int L, int T>
__device__ __forceinline__ void z1_f_body(const ZArgs &a, const cplx_t<R> *tw_all, int64_t lt, int gx) {
  using V = cplx_t<R>;
  constexpr int P = Geo<L>::P, ST = Geo<L>::ST;
  constexpr int E = (L + P - 1) / P;
  constexpr bool FULL = E * P == L;
  __shared__ V lds[T * ST];
  const V *tw = tw_all + a.plan.tw_off;
  const int kx0 = (int)(lt % gx) * T, ky = (int)(lt / gx);
  const int tid = threadIdx.x;
  const int line = tid / P, p = tid % P;
  Twid<L, V> twd;
  twid_load<L>(twd, p, tw);
  const int kx = kx0 + line;
  const int ncl = a.Hx - kx0 < T ? a.Hx - kx0 : T;
  V *B = reinterpret_cast<V *>(a.B);
  const int64_t tile = (int64_t)ky * a.kys + kx0;
  const int cl = tid % T, r0 = tid / T;  // NT / T == P rows per sweep
  {
    V v[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int row = r0 + i * P;
      v[i] = cmake<V>(0, 0);
      if (row < a.Mz && cl < ncl) v[i] = B[tile + (int64_t)row * a.rs + cl];
    }
#pragma unroll
    for (int i = 0; i < E; ++i) {
      const int row = r0 + i * P;
      if (row < L) lds[cl * ST + pad(row)] = v[i];
    }
  }
  __syncthreads();
  V *myl = lds + line * ST;
  fft<L, -1>(myl, p, twd);
  const R ki = wavenumber<R>(kx, a.Lx, a.dk[0], a.ksmp[0]);
  const R kj = wavenumber<R>(ky, a.Ly, a.dk[1], a.ksmp[1]);
  const int kyf = ky <= a.Ly - ky ? ky : a.Ly - ky;
  const R pi = (R)3.14159265358979323846;
  const R *G = reinterpret_cast<const R *>(a.Ghat);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int n = p + e * P;
    if (!(FULL || n < L)) continue;
    R g = (R)0;
    if (a.allper) {
      const R kk = wavenumber<R>(n, L, a.dk[2], a.ksmp[2]);
      const R kmag = ki * ki + kj * kj + kk * kk;
      g = kmag > (R)0 ? (R)1 / ((R)4 * (pi * pi) * kmag) : (R)0;
    } else if (kx < a.Hx) {
      const int kzf = n <= L - n ? n : L - n;
      g = G[((int64_t)kzf * a.Hy + kyf) * a.Hx + kx] * (R)a.h3;
    }
    myl[pad(n)] = cscale(myl[pad(n)], g);
  }
  line_sync<Geo<L>::WAVE>();
  fft<L, 1>(myl, p, twd);
  __syncthreads();
  const R scale = (R)1 / (R)L;
  if (cl < ncl)
    for (int row = r0; row < a.Mz; row += P) B[tile + (int64_t)row * a.rs + cl] = cscale(lds[cl * ST + pad(row)], scale);
}
template <typename R, int L, int T>
__global__ void __launch_bounds__(Geo<L>::P *T) k_z1_f(ZArgs a, const cplx_t<R> *tw_all) {
  z1_f_body<R, L, T>(a, tw_all, xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y), gridDim.x);
}

// Register-resident scalar z pass: thread (col, p) of a block owns column
// kx0 + col (of TC consecutive ones, TC * sizeof(V) contiguous bytes per z
// row) and the rows p + P m; the first and last FFT stages run on the
// registers and G-hat multiplies in registers, LDS only carries the
// exchanges between stages.  HALF (free-space z, Mz = L / 2) skips the zero
// upper half in the first forward and the last inverse stage.  The host
// guarantees Mz * rs complex words fit a 31-bit byte offset.
template <typename R, int L, int TC, bool ALLPER, bool HALF>
__device__ __forceinline__ void z1_r_body(const ZArgs &a, const cplx_t<R> *tw_all, int64_t lt, int gx) {
  using V = cplx_t<R>;
  constexpr int P = Geo<L>::P, E = L / P;
  constexpr int ST = reg_line_stride<L, sizeof(V), TC>();
  __shared__ V lds[TC * ST];
  const V *tw = tw_all + a.plan.tw_off;
  const int kx0 = (int)(lt % gx) * TC, ky = (int)(lt / gx);
  const int tid = threadIdx.x;
  const int col = tid % TC, p = tid / TC;
  Twid<L, V> twd;
  twid_load<L>(twd, p, tw);
  const int kx = kx0 + col;
  const bool col_ok = kx < a.Hx;
  V *base = reinterpret_cast<V *>(a.B) + (int64_t)ky * a.kys + kx0;
  const int64_t rsb = a.rs * (int64_t)sizeof(V);
  const int ncols_here = a.Hx - kx0 < TC ? a.Hx - kx0 : TC;
  const uint32_t extent = (uint32_t)((a.Mz - 1) * rsb + (int64_t)ncols_here * sizeof(V));
  const __amdgpu_buffer_rsrc_t rsrc = make_rsrc(base, extent);
  const int voff = (int)(p * rsb) + col * (int)sizeof(V);
  V w[E];
#pragma unroll
  for (int m = 0; m < E; ++m) {
    const bool ok = col_ok && (HALF ? m < E / 2 : p + P * m < a.Mz);
    w[m] = ok ? buf_load(rsrc, voff, (int)((int64_t)P * m * rsb), V{}) : cmake<V>(0, 0);
  }
  V *line = lds + col * ST;
  fft_regs<L, -1, V, HALF, false, 1>(w, line, p, twd);
  const int kyf = ky <= a.Ly - ky ? ky : a.Ly - ky;
  if constexpr (ALLPER) {
    const R ki = wavenumber<R>(kx, a.Lx, a.dk[0], a.ksmp[0]);
    const R kj = wavenumber<R>(ky, a.Ly, a.dk[1], a.ksmp[1]);
    const R pi = (R)3.14159265358979323846;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const R kk = wavenumber<R>(p + P * m, L, a.dk[2], a.ksmp[2]);
      const R kmag = ki * ki + kj * kj + kk * kk;
      w[m] = cscale(w[m], kmag > (R)0 ? (R)1 / ((R)4 * (pi * pi) * kmag) : (R)0);
    }
  } else {
    // G-hat of the thread's rows, all loads issued before any is used (a
    // column past Hx reads column 0; its results are not stored)
    const R *G = reinterpret_cast<const R *>(a.Ghat) + (int64_t)kyf * a.Hx + (col_ok ? kx : 0);
    const int64_t gz = (int64_t)a.Hy * a.Hx;
    R gv[E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int n = p + P * m;
      gv[m] = G[(n <= L - n ? n : L - n) * gz];
    }
#pragma unroll
    for (int m = 0; m < E; ++m) w[m] = cscale(w[m], gv[m] * (R)a.h3);
  }
  fft_regs<L, 1, V, false, HALF, 1>(w, line, p, twd);
  const R scale = (R)1 / (R)L;
  if (col_ok) {
    V *dst = base + (int64_t)p * a.rs + col;
#pragma unroll
    for (int m = 0; m < E; ++m)
      if (HALF ? m < E / 2 : p + P * m < a.Mz) dst[(int64_t)P * m * a.rs] = cscale(w[m], scale);
  }
}
template <typename R, int L, int TC, bool ALLPER, bool HALF>
__global__ void __launch_bounds__(Geo<L>::P *TC) k_z1_r(ZArgs a, const cplx_t<R> *tw_all) {
  z1_r_body<R, L, TC, ALLPER, HALF>(a, tw_all, xcd_tile(block_linear(), (int64_t)gridDim.x * gridDim.y), gridDim.x);
}

// ------------------------------------------------------------ group kernels
// One launch runs one kernel instance over several patches: entry e of the
// table covers global tiles [off, off + tiles) of patch e's own grid (gx
// tiles per row) with its own arguments and twiddle table.  The global tile
// comes from the XCD-aware block mapping; the entry by binary search over
// the offsets (uniform per block).
// the block's entry, its arguments and its tile within the patch's grid
template <typename A> struct GrpTile {
  A a;
  const void *tw;
  int64_t lt;
  int gx;
};
template <typename A>
__device__ __forceinline__ GrpTile<A> grp_tile(const GrpEntry<A> *__restrict__ t, int n, int64_t nb) {
  const int64_t g = xcd_tile(block_linear(), nb);
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (t[mid].off <= g) lo = mid;
    else hi = mid - 1;
  }
  return GrpTile<A>{t[lo].a, t[lo].tw, g - t[lo].off, (int)t[lo].gx};
}

template <typename R, int L, int T, int S>
__global__ void __launch_bounds__(Geo<L>::P *T)
    kg_strided_f(const GrpEntry<StridedArgs> *__restrict__ t, int n, int64_t nb) {
  const GrpTile<StridedArgs> g = grp_tile(t, n, nb);
  strided_f_body<R, L, T, S>(g.a, reinterpret_cast<const cplx_t<R> *>(g.tw), g.lt, g.gx);
}
template <typename R, int L, int T, int S, bool HALF>
__global__ void __launch_bounds__(Geo<L>::P *T)
    kg_strided_r(const GrpEntry<StridedArgs> *__restrict__ t, int n, int64_t nb) {
  const GrpTile<StridedArgs> g = grp_tile(t, n, nb);
  strided_r_body<R, L, T, S, HALF>(g.a, reinterpret_cast<const cplx_t<R> *>(g.tw), g.lt, g.gx);
}
template <typename R, int NH, int T, int NC>
__global__ void __launch_bounds__(Geo<NH>::P *T *NC)
    kg_x_fwd_f(const GrpEntry<XFwdArgs> *__restrict__ t, int n, int64_t nb) {
  const GrpTile<XFwdArgs> g = grp_tile(t, n, nb);
  x_fwd_body<R, NH, T, NC>(g.a, reinterpret_cast<const cplx_t<R> *>(g.tw), g.lt, g.gx);
}
template <typename R, int NH, int T, int NC>
__global__ void __launch_bounds__(Geo<NH>::P *T *NC)
    kg_x_inv_f(const GrpEntry<XInvArgs> *__restrict__ t, int n, int64_t nb) {
  const GrpTile<XInvArgs> g = grp_tile(t, n, nb);
  x_inv_body<R, NH, T, NC>(g.a, reinterpret_cast<const cplx_t<R> *>(g.tw), g.lt, g.gx);
}
template <typename R, int L, int T>
__global__ void __launch_bounds__(Geo<L>::P *T) kg_z1_f(const GrpEntry<ZArgs> *__restrict__ t, int n, int64_t nb) {
  const GrpTile<ZArgs> g = grp_tile(t, n, nb);
  z1_f_body<R, L, T>(g.a, reinterpret_cast<const cplx_t<R> *>(g.tw), g.lt, g.gx);
}
template <typename R, int L, int TC, bool ALLPER, bool HALF>
__global__ void __launch_bounds__(Geo<L>::P *TC) kg_z1_r(const GrpEntry<ZArgs> *__restrict__ t, int n, int64_t nb) {
  const GrpTile<ZArgs> g = grp_tile(t, n, nb);
  z1_r_body<R, L, TC, ALLPER, HALF>(g.a, reinterpret_cast<const cplx_t<R> *>(g.tw), g.lt, g.gx);
}

}  // namespace fx
}  // namespace pmgpu
