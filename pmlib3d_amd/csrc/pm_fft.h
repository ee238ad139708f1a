// LDS-resident mixed-radix FFT engine for gfx950.
//
// A workgroup holds NL lines of length L in LDS (line stride L + pad) and
// transforms them together: each line is served by P = ceil(L/8) consecutive
// threads, every Stockham stage reads its radix-R butterflies into registers,
// barriers, applies twiddles + the radix-R DFT and writes them back in
// autosorted order, so the result is in natural order with no bit reversal.
// Radices 8, 4, 2, 3, 5 (closed-form butterflies) and 7, 11, 13 (direct
// summation) are supported; twiddles come from a per-length
// table tw[n] = exp(-2 pi i n / L) built on the host in extended precision.
#pragma once
#include "pm_common.h"

namespace pmgpu {

constexpr int kMaxStages = 16;

struct FftPlan {
  int L;              // transform length
  int P;              // threads per line
  int nstage;
  int tw_off;         // offset (complex elements) of this length's table
  int radix[kMaxStages];
};

// maximum butterflies one thread owns in a stage of radix R when P >= L/8
template <int R> struct MaxBf { static constexpr int v = (8 + R - 1) / R; };

template <int R, int S, typename V> struct Dft;

template <int S, typename V> struct Dft<2, S, V> {
  static __device__ __forceinline__ void run(V *v) {
    V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int S, typename V> struct Dft<4, S, V> {
  static __device__ __forceinline__ void run(V *v) {
    V s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    V s13 = cadd(v[1], v[3]), d13 = cmul_i<S>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};

template <int S, typename V> struct Dft<8, S, V> {
  static __device__ __forceinline__ void run(V *v) {
    using R = decltype(V::x);
    const R h = (R)0.70710678118654752440;
    V a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      a[k] = cadd(v[k], v[k + 4]);
      b[k] = csub(v[k], v[k + 4]);
    }
    // b[k] *= exp(S * 2 pi i k / 8)
    b[1] = cscale(cadd(b[1], cmul_i<S>(b[1])), h);
    b[2] = cmul_i<S>(b[2]);
    b[3] = cscale(csub(cmul_i<S>(b[3]), b[3]), h);
    Dft<4, S, V>::run(a);
    Dft<4, S, V>::run(b);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      v[2 * m] = a[m];
      v[2 * m + 1] = b[m];
    }
  }
};

template <int S, typename V> struct Dft<3, S, V> {
  static __device__ __forceinline__ void run(V *v) {
    using R = decltype(V::x);
    const R s3 = (R)0.86602540378443864676;  // sin(2 pi / 3)
    V t1 = cadd(v[1], v[2]);
    V t2 = csub(v[0], cscale(t1, (R)0.5));
    V t3 = cscale(cmul_i<S>(csub(v[1], v[2])), s3);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  }
};

template <int S, typename V> struct Dft<5, S, V> {
  static __device__ __forceinline__ void run(V *v) {
    using R = decltype(V::x);
    const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
    const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
    V a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
    V b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
    V x0 = v[0];
    V r1 = cadd(x0, cadd(cscale(a1, c1), cscale(a2, c2)));
    V r2 = cadd(x0, cadd(cscale(a1, c2), cscale(a2, c1)));
    // i*S*(s1 b1 + s2 b2) and i*S*(s2 b1 - s1 b2)
    V q1 = cmul_i<S>(cadd(cscale(b1, s1), cscale(b2, s2)));
    V q2 = cmul_i<S>(csub(cscale(b1, s2), cscale(b2, s1)));
    v[0] = cadd(x0, cadd(a1, a2));
    v[1] = cadd(r1, q1);
    v[4] = csub(r1, q1);
    v[2] = cadd(r2, q2);
    v[3] = csub(r2, q2);
  }
};

// Odd prime radix by direct summation; roots of unity come from the
// length-L table (exp(-2 pi i m / R) = tw[m * L / R]).
template <int R, int S, typename V> struct DftPrime {
  static __device__ __forceinline__ void run(V *v, const V *tw, int L) {
    const int step = L / R;
    V out[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      V acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        V w = tw[((r * m) % R) * step];
        if (S > 0) w.y = -w.y;
        acc = cadd(acc, cmul(v[r], w));
      }
      out[m] = acc;
    }
#pragma unroll
    for (int m = 0; m < R; ++m) v[m] = out[m];
  }
};

template <int R, int S, typename V> struct DftAny {
  static __device__ __forceinline__ void run(V *v, const V *, int) { Dft<R, S, V>::run(v); }
};
template <int S, typename V> struct DftAny<7, S, V> {
  static __device__ __forceinline__ void run(V *v, const V *tw, int L) { DftPrime<7, S, V>::run(v, tw, L); }
};
template <int S, typename V> struct DftAny<11, S, V> {
  static __device__ __forceinline__ void run(V *v, const V *tw, int L) { DftPrime<11, S, V>::run(v, tw, L); }
};
template <int S, typename V> struct DftAny<13, S, V> {
  static __device__ __forceinline__ void run(V *v, const V *tw, int L) { DftPrime<13, S, V>::run(v, tw, L); }
};

// One Stockham stage of radix R on one line (p = thread index in the line).
template <int R, int S, typename V>
__device__ __forceinline__ void fft_stage(V *line, int L, int Ns, int p, int P, const V *tw) {
  constexpr int MB = MaxBf<R>::v;
  const int nbf = L / R;
  V v[MB][R];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    int j = p + m * P;
    if (j < nbf) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[m][r] = line[j + r * nbf];
    }
  }
  __syncthreads();
  const int tstride = L / (Ns * R);
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    int j = p + m * P;
    if (j < nbf) {
      int k = j % Ns;
      int t = k * tstride;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        V w = tw[r * t];
        if (S > 0) w.y = -w.y;
        v[m][r] = cmul(v[m][r], w);
      }
      DftAny<R, S, V>::run(v[m], tw, L);
      int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) line[base + r * Ns] = v[m][r];
    }
  }
  __syncthreads();
}

// Full transform of the line(s) held in LDS; all threads of the block call it.
template <int S, typename V>
__device__ __forceinline__ void fft_lds(V *line, const FftPlan &pl, int p, const V *tw_all) {
  const V *tw = tw_all + pl.tw_off;
  int Ns = 1;
  for (int s = 0; s < pl.nstage; ++s) {
    int R = pl.radix[s];
    switch (R) {
      case 8: fft_stage<8, S>(line, pl.L, Ns, p, pl.P, tw); break;
      case 4: fft_stage<4, S>(line, pl.L, Ns, p, pl.P, tw); break;
      case 2: fft_stage<2, S>(line, pl.L, Ns, p, pl.P, tw); break;
      case 3: fft_stage<3, S>(line, pl.L, Ns, p, pl.P, tw); break;
      case 5: fft_stage<5, S>(line, pl.L, Ns, p, pl.P, tw); break;
      case 7: fft_stage<7, S>(line, pl.L, Ns, p, pl.P, tw); break;
      case 11: fft_stage<11, S>(line, pl.L, Ns, p, pl.P, tw); break;
      default: fft_stage<13, S>(line, pl.L, Ns, p, pl.P, tw); break;
    }
    Ns *= R;
  }
}

// LDS line stride: one extra element breaks the power-of-two bank pattern of
// the transposed tile loads.
__host__ __device__ inline int lds_stride(int L) { return L + 1; }

}  // namespace pmgpu
