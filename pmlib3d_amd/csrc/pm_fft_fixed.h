// Compile-time specialised FFT engine (hot lengths).
//
// Same Stockham autosort algorithm as pm_fft.h, but the length, the radix
// sequence, the threads per line P (a power of two) and the butterflies per
// thread are template constants, so every index is folded, every register
// array is statically indexed (no scratch) and a line that fits in one
// wavefront (P <= 64) synchronises with wave-level fences instead of block
// barriers.  LDS indices are padded, i -> i + i/8 + i/64, which spreads the
// stride-8 and stride-64 element writes of the first stages over the banks.
// Twiddles: each thread loads the base twiddle of each of its butterflies
// once (twid_load), higher powers by multiplication.
#pragma once
#include "pm_fft.h"

namespace pmgpu {
namespace fx {

// Radix sequence of a length.  Powers of two: radix 8 stages, the remaining
// 4 or 2 last.  q 2^k with q in {3, 5} and k >= 6: q first, then 8s, the
// remaining 4 or 2 last, so that the threads per line P (below) divide both
// the first and the last stage's butterfly counts and a thread's first- and
// last-stage elements are the ones it holds in registers; where that would
// leave more than 12 elements per thread the last 8 (or 4) is split into
// 4 (or 2) and 2.  Other lengths: 8, 4, 2, 3, 5 in that order.
struct RadixSeq {
  int r[16];
  int n;
  int P;
};
constexpr int pow2_part(int v) {
  int p = 1;
  while (v % (2 * p) == 0) p *= 2;
  return p;
}
constexpr int next_pow2(int v) {
  int p = 1;
  while (p < v) p *= 2;
  return p;
}
constexpr RadixSeq radix_seq(int L) {
  RadixSeq s{};
  const int p2 = L > 0 ? pow2_part(L) : 1, q = L > 0 ? L / p2 : 1;
  int k = 0;
  while ((1 << k) < p2) ++k;
  if ((q == 3 || q == 5) && k >= 6) {
    s.r[s.n++] = q;
    for (int i = 0; i < k / 3; ++i) s.r[s.n++] = 8;
    if (k % 3) s.r[s.n++] = 1 << (k % 3);
    auto threads = [&](const RadixSeq &t) {
      const int a = pow2_part(L / t.r[0]), b = pow2_part(L / t.r[t.n - 1]);
      const int np = next_pow2((L + 7) / 8);
      int P = a < b ? a : b;
      return P < np ? P : np;
    };
    s.P = threads(s);
    if (L / s.P > 12 && s.r[s.n - 1] >= 4) {
      s.r[s.n - 1] /= 2;
      s.r[s.n++] = 2;
      s.P = threads(s);
    }
    return s;
  }
  int r = L;
  const int rads[5] = {8, 4, 2, 3, 5};
  for (int i = 0; i < 5; ++i)
    while (r % rads[i] == 0 && r > 1 && s.n < 16) {
      s.r[s.n++] = rads[i];
      r /= rads[i];
    }
  if (r != 1) s.n = 0;  // unsupported
  s.P = next_pow2((L + 7) / 8);
  return s;
}
constexpr int radix_at(int L, int s) {
  const RadixSeq q = radix_seq(L);
  return s < q.n ? q.r[s] : 0;
}
constexpr int nstages(int L) { return radix_seq(L).n; }
constexpr bool supported(int L) {
  int p = 1;
  for (int s = 0; s < nstages(L); ++s) p *= radix_at(L, s);
  return p == L && L >= 8;
}
constexpr int ns_before(int L, int s) {
  int p = 1;
  for (int q = 0; q < s; ++q) p *= radix_at(L, q);
  return p;
}
template <int L> struct Geo {
  static constexpr int P = radix_seq(L).P;           // threads per line
  static constexpr int PADDED = L + L / 8 + L / 64 + 1;
  static constexpr int ST = PADDED | 1;               // odd line stride
  static constexpr bool WAVE = P <= 64;
};
__device__ __forceinline__ int pad(int i) { return i + (i >> 3) + (i >> 6); }
// Padding of the register-resident transforms (PM = 1): i + i / 8 keeps the
// first stage's stride-8 writes, the unit-stride reads and the middle stage's
// grouped writes on distinct banks (with a line stride of 8 mod 32 complex
// FP32 words, 4 mod 16 FP64 words between the columns of a block).
template <int PM> __device__ __forceinline__ int padm(int i) {
  return PM == 1 ? i + (i >> 3) : (PM == 2 ? i + (i >> 4) : pad(i));
}


template <bool WAVE> __device__ __forceinline__ void line_sync() {
  if (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// per-thread twiddles of every stage (the base power w1 of each butterfly),
// loaded once per kernel and reused by forward and inverse transforms
template <int L> constexpr int max_mb() {
  int m = 1;
  for (int q = 0; q < nstages(L); ++q) {
    int nbf = L / radix_at(L, q);
    int mb = (nbf + Geo<L>::P - 1) / Geo<L>::P;
    if (mb > m) m = mb;
  }
  return m;
}
template <int L, typename V> struct Twid {
  V w[nstages(L) > 0 ? nstages(L) : 1][max_mb<L>()];
};

template <int L, int I, typename V, bool DONE = (I >= nstages(L))> struct TwidLoad;
template <int L, int I, typename V> struct TwidLoad<L, I, V, true> {
  static __device__ __forceinline__ void run(Twid<L, V> &, int, const V *) {}
};
template <int L, int I, typename V> struct TwidLoad<L, I, V, false> {
  static __device__ __forceinline__ void run(Twid<L, V> &t, int p, const V *tw) {
    constexpr int R = radix_at(L, I), NS = ns_before(L, I), P = Geo<L>::P;
    constexpr int NBF = L / R, MB = (NBF + P - 1) / P, TSTRIDE = L / (NS * R);
    if (NS > 1) {
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int j = p + m * P;
        t.w[I][m] = (MB * P == NBF || j < NBF) ? tw[(j % NS) * TSTRIDE] : tw[0];
      }
    }
    TwidLoad<L, I + 1, V>::run(t, p, tw);
  }
};
template <int L, typename V> __device__ __forceinline__ void twid_load(Twid<L, V> &t, int p, const V *tw) {
  TwidLoad<L, 0, V>::run(t, p, tw);
}

// v[r] *= w^r (conjugated for S > 0), r = 1 .. R - 1, for butterfly m of
// stage I, the powers of the base twiddle by a multiplication chain.  For
// lines of 12 or more elements per thread the chain is formed anew at every
// use: the empty asm hides the base value's identity, so that the compiler
// does not keep the powers of every stage live across the transforms of a
// kernel (at length 768 that held the three-component z pass at one
// wavefront per SIMD; forming them again costs 7 % more VALU and gives two)
template <int L, int I, int S, int R, typename V>
__device__ __forceinline__ void twiddle(V *v, const Twid<L, V> &t, int m) {
  V w1 = t.w[I][m];
  if constexpr (L / Geo<L>::P >= 12) asm volatile("" : "+v"(w1.x), "+v"(w1.y));
  if (S > 0) w1.y = -w1.y;
  V w = w1;
#pragma unroll
  for (int r = 1; r < R; ++r) {
    v[r] = cmul(v[r], w);
    if (r + 1 < R) w = cmul(w, w1);
  }
}

template <int L, int S, int I, bool WAVE, typename V, int PM>
__device__ __forceinline__ void stage(V *line, int p, const Twid<L, V> &tws) {
  constexpr int R = radix_at(L, I), NS = ns_before(L, I);
  constexpr int P = Geo<L>::P;
  constexpr int NBF = L / R;
  constexpr int MB = (NBF + P - 1) / P;
  V v[MB][R];
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int j = p + m * P;
    if (MB * P == NBF || j < NBF) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[m][r] = line[padm<PM>(j + r * NBF)];
    }
  }
  line_sync<WAVE>();
#pragma unroll
  for (int m = 0; m < MB; ++m) {
    const int j = p + m * P;
    if (MB * P == NBF || j < NBF) {
      const int k = NS == 1 ? 0 : j % NS;
      if (NS > 1) twiddle<L, I, S, R>(v[m], tws, m);
      Dft<R, S, V>::run(v[m]);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) line[padm<PM>(base + r * NS)] = v[m][r];
    }
  }
  line_sync<WAVE>();
}

template <int L, int S, int I, bool WAVE, typename V, int PM = 0, bool DONE = (I >= nstages(L))> struct StageLoop;
template <int L, int S, int I, bool WAVE, typename V, int PM> struct StageLoop<L, S, I, WAVE, V, PM, true> {
  static __device__ __forceinline__ void run(V *, int, const Twid<L, V> &) {}
};
template <int L, int S, int I, bool WAVE, typename V, int PM> struct StageLoop<L, S, I, WAVE, V, PM, false> {
  static __device__ __forceinline__ void run(V *line, int p, const Twid<L, V> &t) {
    stage<L, S, I, WAVE, V, PM>(line, p, t);
    StageLoop<L, S, I + 1, WAVE, V, PM>::run(line, p, t);
  }
};

// In-place transform of one padded line held in LDS with twiddles preloaded
// by twid_load from the length-L table.  The data must be visible to the
// line's threads (caller syncs before).
template <int L, int S, typename V, int PM = 0>
__device__ __forceinline__ void fft(V *line, int p, const Twid<L, V> &t) {
  StageLoop<L, S, 0, Geo<L>::WAVE, V, PM>::run(line, p, t);
}

// ------------------------------------------------ register-resident lines
// Thread p of a line holds elements p + P m (m < L / P) in registers on entry
// and on exit; the first stage reads its inputs from those registers and the
// last stage leaves its outputs there, so the line buffer in LDS is used only
// for the exchanges between stages.  Lines may span several wavefronts: every
// thread of the block calls this together (block barriers).
template <int L> constexpr bool regs_ok() {
  constexpr int P = Geo<L>::P;
  if (!supported(L) || nstages(L) < 2 || (L / P) * P != L) return false;
  const int n0 = L / radix_at(L, 0), n1 = L / radix_at(L, nstages(L) - 1);
  return n0 % P == 0 && n1 % P == 0;
}

template <int L, int S, int I, typename V, int PM = 0, bool DONE = (I >= nstages(L) - 1)> struct MidStages;
template <int L, int S, int I, typename V, int PM> struct MidStages<L, S, I, V, PM, true> {
  static __device__ __forceinline__ void run(V *, int, const Twid<L, V> &) {}
};
template <int L, int S, int I, typename V, int PM> struct MidStages<L, S, I, V, PM, false> {
  static __device__ __forceinline__ void run(V *line, int p, const Twid<L, V> &t) {
    stage<L, S, I, false, V, PM>(line, p, t);
    MidStages<L, S, I + 1, V, PM>::run(line, p, t);
  }
};

// ZIN: the upper half of the line's elements (indices >= L / 2) are zero on
// entry, so the first stage skips them; HALFOUT: only the lower half is
// wanted on exit, so the last stage leaves the upper half unset.
template <int L, int S, typename V, bool ZIN = false, bool HALFOUT = false, int PM = 0>
__device__ __forceinline__ void fft_regs(V (&x)[L / Geo<L>::P], V *line, int p, const Twid<L, V> &t) {
  constexpr int P = Geo<L>::P, E = L / P;
  {  // first stage: registers -> line (no twiddles, NS = 1)
    constexpr int R = radix_at(L, 0), NBF = L / R, MB = NBF / P;
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      V v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = (ZIN && b + MB * r >= E / 2) ? cmake<V>(0, 0) : x[b + MB * r];
      Dft<R, S, V>::run(v);
      const int jj = p + b * P;
#pragma unroll
      for (int r = 0; r < R; ++r) line[padm<PM>(jj * R + r)] = v[r];
    }
  }
  __syncthreads();
  MidStages<L, S, 1, V, PM>::run(line, p, t);
  {  // last stage: line -> registers (NS = NBF, outputs at p + P m)
    constexpr int I = nstages(L) - 1, R = radix_at(L, I), NBF = L / R, MB = NBF / P;
#pragma unroll
    for (int b = 0; b < MB; ++b) {
      const int jj = p + b * P;
      V v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = line[padm<PM>(jj + r * NBF)];
      twiddle<L, I, S, R>(v, t, b);
      Dft<R, S, V>::run(v);
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (!(HALFOUT && b + MB * r >= E / 2)) x[b + MB * r] = v[r];
    }
  }
  __syncthreads();  // the line buffer is free again
}

}  // namespace fx
}  // namespace pmgpu
