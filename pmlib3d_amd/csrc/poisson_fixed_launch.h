// Launchers of the length-specialised Poisson kernels.  Each returns 1 when
// it launched, 0 when the length has no specialisation (the caller then uses
// the runtime-planned kernels of poisson_kernels.h), -1 on a launch error.
// Included by the fx_len_*.hip units, each of which instantiates the
// launcher table (fx_ops) for a few transform lengths, so that the lengths
// compile in parallel.
#pragma once
#include <cstdlib>
#include "poisson_fixed.h"
#include "poisson_fixed_ops.h"

namespace pmgpu {
namespace fx {

constexpr int kLdsBudget = 80 * 1024;  // two workgroups per CU (160 KB LDS)

template <typename R, int L> constexpr int tile_cols() {
  constexpr int P = Geo<L>::P;
  constexpr int want = sizeof(R) == 4 ? 16 : 8;
  int t = want;
  while (t > 1 && (P * t > 1024 || t * Geo<L>::ST * (int)sizeof(cplx_t<R>) > kLdsBudget)) t /= 2;
  return t;
}
// lines per x-pass block: 2 for three-component lines (small blocks overlap
// one block's loads with another's transforms: at 256^3 the solve measured
// FP32 1.35 -> 1.30 ms and FP64 2.59 -> 2.51 ms against 8 / 4 lines; 1 line
// was slower), 8 (FP32) or 4 (FP64) for one component
template <typename R, int NH, int NC> constexpr int tile_lines() {
  constexpr int P = Geo<NH>::P;
  int t = NC == 3 ? 2 : (sizeof(R) == 4 ? 8 : 4);
  while (t > 1 && (P * t * NC > 1024 || t * NC * Geo<NH>::ST * (int)sizeof(cplx_t<R>) > kLdsBudget)) t /= 2;
  return t;
}

template <typename R, int L>
int strided_L(int S, StridedArgs a, int nb_y, int nb_z, const cplx_t<R> *tw, hipStream_t st) {
  if constexpr (regs_ok<L>()) {
    if (a.kernel == 0) {
      constexpr int T = tile_cols<R, L>();
      a.T = T;
      dim3 grid((a.ncols + T - 1) / T, nb_y, nb_z);
      const bool half = (L / Geo<L>::P) % 2 == 0 && (S < 0 ? 2 * a.nin == L : 2 * a.nout == L);
      if (S < 0 && half)
        hipLaunchKernelGGL((k_strided_r<R, L, T, -1, true>), grid, dim3(Geo<L>::P * T), 0, st, a, tw);
      else if (S < 0)
        hipLaunchKernelGGL((k_strided_r<R, L, T, -1, false>), grid, dim3(Geo<L>::P * T), 0, st, a, tw);
      else if (half)
        hipLaunchKernelGGL((k_strided_r<R, L, T, 1, true>), grid, dim3(Geo<L>::P * T), 0, st, a, tw);
      else
        hipLaunchKernelGGL((k_strided_r<R, L, T, 1, false>), grid, dim3(Geo<L>::P * T), 0, st, a, tw);
      return hipGetLastError() == hipSuccess ? 1 : -1;
    }
  }
  constexpr int T = tile_cols<R, L>();
  a.T = T;
  dim3 grid((a.ncols + T - 1) / T, nb_y, nb_z);
  if (S < 0)
    hipLaunchKernelGGL((k_strided_f<R, L, T, -1>), grid, dim3(Geo<L>::P * T), 0, st, a, tw);
  else
    hipLaunchKernelGGL((k_strided_f<R, L, T, 1>), grid, dim3(Geo<L>::P * T), 0, st, a, tw);
  return hipGetLastError() == hipSuccess ? 1 : -1;
}

// columns per block of the interleaved z pass (3 component lines each)
template <typename R, int L, bool EXTRA> constexpr int z_cols() {
  constexpr int P = Geo<L>::P;
  int t = sizeof(R) == 4 ? (EXTRA ? 2 : 5) : (EXTRA ? 1 : 2);
  while (t > 1 && (P * 3 * t > 1024 || 3 * t * Geo<L>::ST * (int)sizeof(cplx_t<R>) > kLdsBudget)) --t;
  return t;
}

template <typename R, int L, int TC> int z_tc(ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  a.T = TC;
  dim3 grid((a.Hx + TC - 1) / TC, Ly);
  hipLaunchKernelGGL((k_z3_f<R, L, TC, false>), grid, dim3(Geo<L>::P * 3 * TC), 0, st, a, tw);
  return hipGetLastError() == hipSuccess ? 1 : -1;
}

// register-resident velocity z pass with TC columns per block
template <typename R, int L, int TC> int zr_T(ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  if constexpr (!regs_ok<L>() || TC * reg_line_stride<L, sizeof(cplx_t<R>), TC>() * (int)sizeof(cplx_t<R>) > kLdsBudget ||
                Geo<L>::P * TC > 1024) {
    return 0;
  } else {
    // buffer offsets of the tile are 31-bit
    if ((double)a.Mz * (double)a.rs * sizeof(cplx_t<R>) >= 2147483647.0) return 0;
    dim3 grid((a.Hx + TC - 1) / TC, Ly);
    if (a.allper)
      hipLaunchKernelGGL((k_z3_r<R, L, TC, true, false>), grid, dim3(Geo<L>::P * TC), 0, st, a, tw);
    else if (2 * a.Mz == L && (L / Geo<L>::P) % 2 == 0)
      hipLaunchKernelGGL((k_z3_r<R, L, TC, false, true>), grid, dim3(Geo<L>::P * TC), 0, st, a, tw);
    else
      hipLaunchKernelGGL((k_z3_r<R, L, TC, false, false>), grid, dim3(Geo<L>::P * TC), 0, st, a, tw);
    return hipGetLastError() == hipSuccess ? 1 : -1;
  }
}
// register-resident full-output z pass (psi and/or filtered vorticity); SL
// keeps the forward spectra in LDS
template <typename R, int L, int TC, bool SL>
int zrx_T(ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  constexpr int lds_bytes = (TC * reg_line_stride<L, sizeof(cplx_t<R>), TC>() + (SL ? TC * 3 * L : 0)) *
                            (int)sizeof(cplx_t<R>);
  if constexpr (!regs_ok<L>() || lds_bytes > kLdsBudget || Geo<L>::P * TC > 1024) {
    return 0;
  } else {
    if ((double)a.Mz * (double)a.rs * sizeof(cplx_t<R>) >= 2147483647.0) return 0;
    dim3 grid((a.Hx + TC - 1) / TC, Ly);
    if (a.allper)
      hipLaunchKernelGGL((k_z3_rx<R, L, TC, true, false, SL>), grid, dim3(Geo<L>::P * TC), 0, st, a, tw);
    else if (2 * a.Mz == L && (L / Geo<L>::P) % 2 == 0)
      hipLaunchKernelGGL((k_z3_rx<R, L, TC, false, true, SL>), grid, dim3(Geo<L>::P * TC), 0, st, a, tw);
    else
      hipLaunchKernelGGL((k_z3_rx<R, L, TC, false, false, SL>), grid, dim3(Geo<L>::P * TC), 0, st, a, tw);
    return hipGetLastError() == hipSuccess ? 1 : -1;
  }
}
template <typename R, int L> int zrx_L(ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  // FP32: spectra in registers, 4 columns per block; FP64: spectra in LDS
  // (the register form runs at one wave per SIMD), 2 or 1 columns as LDS allows
  static const int form = [] {
    const char *e = getenv("PMGPU_ZRX");
    return e ? atoi(e) : -1;
  }();
  const bool sl = form >= 0 ? form == 1 : sizeof(R) == 8;
  if (!sl) return zrx_T<R, L, 4, false>(a, Ly, tw, st);
  const int r = zrx_T<R, L, 2, true>(a, Ly, tw, st);
  return r != 0 ? r : zrx_T<R, L, 1, true>(a, Ly, tw, st);
}

template <typename R, int L> int zr_L(ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  const int want = a.tune_tc > 0 ? a.tune_tc : 4;
  switch (want) {
    case 2: return zr_T<R, L, 2>(a, Ly, tw, st);
    case 4: return zr_T<R, L, 4>(a, Ly, tw, st);
    case 8: return zr_T<R, L, 8>(a, Ly, tw, st);
    default: return 0;
  }
}

template <typename R, int L> int z_L(ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  if (!(a.want_psi || a.want_vort) && a.zkernel != 1 && !a.ker2) {
    const int r = zr_L<R, L>(a, Ly, tw, st);
    if (r != 0) return r;
  }
  if ((a.want_psi || a.want_vort) && a.zkernel != 1 && !a.ker2) {
    const int r = zrx_L<R, L>(a, Ly, tw, st);
    if (r != 0) return r;
  }
  // tuned tile width (columns per block) for the hot lengths, velocity only
  if constexpr (L == 256 || L == 512) {
    if (!(a.want_psi || a.want_vort) && a.tune_tc > 0 && a.zkernel == 1) {
      switch (a.tune_tc) {
        case 1: return z_tc<R, L, 1>(a, Ly, tw, st);
        case 2: return z_tc<R, L, 2>(a, Ly, tw, st);
        case 3: return z_tc<R, L, 3>(a, Ly, tw, st);
        case 4: return z_tc<R, L, 4>(a, Ly, tw, st);
        default: break;
      }
    }
  }
  if (a.want_psi || a.want_vort) {
    constexpr int TC = z_cols<R, L, true>();
    a.T = TC;
    dim3 grid((a.Hx + TC - 1) / TC, Ly);
    hipLaunchKernelGGL((k_z3_f<R, L, TC, true>), grid, dim3(Geo<L>::P * 3 * TC), 0, st, a, tw);
  } else {
    constexpr int TC = z_cols<R, L, false>();
    a.T = TC;
    dim3 grid((a.Hx + TC - 1) / TC, Ly);
    hipLaunchKernelGGL((k_z3_f<R, L, TC, false>), grid, dim3(Geo<L>::P * 3 * TC), 0, st, a, tw);
  }
  return hipGetLastError() == hipSuccess ? 1 : -1;
}

template <typename R, int NH> int xfwd_L(int NC, XFwdArgs a, const cplx_t<R> *tw, hipStream_t st) {
  if (NC == 3) {
    constexpr int T = tile_lines<R, NH, 3>();
    a.lines_per_block = T;
    dim3 grid((a.My + T - 1) / T, a.Mz);
    hipLaunchKernelGGL((k_x_fwd_f<R, NH, T, 3>), grid, dim3(Geo<NH>::P * T * 3), 0, st, a, tw);
  } else {
    constexpr int T = tile_lines<R, NH, 1>();
    a.lines_per_block = T;
    dim3 grid((a.My + T - 1) / T, a.Mz);
    hipLaunchKernelGGL((k_x_fwd_f<R, NH, T, 1>), grid, dim3(Geo<NH>::P * T), 0, st, a, tw);
  }
  return hipGetLastError() == hipSuccess ? 1 : -1;
}

template <typename R, int NH> int xinv_L(XInvArgs a, const cplx_t<R> *tw, hipStream_t st) {
  if (a.ncomp == 1) {
    constexpr int T = tile_lines<R, NH, 1>();
    a.lines_per_block = T;
    dim3 grid((a.My + T - 1) / T, a.Mz);
    hipLaunchKernelGGL((k_x_inv_f<R, NH, T, 1>), grid, dim3(Geo<NH>::P * T), 0, st, a, tw);
  } else {
    constexpr int T = tile_lines<R, NH, 3>();
    a.lines_per_block = T;
    dim3 grid((a.My + T - 1) / T, a.Mz);
    hipLaunchKernelGGL((k_x_inv_f<R, NH, T, 3>), grid, dim3(Geo<NH>::P * T * 3), 0, st, a, tw);
  }
  return hipGetLastError() == hipSuccess ? 1 : -1;
}

// scalar z pass: T columns per block, T in {1, 2, 4, 8, 16} where the LDS fits
template <typename R, int L, int T> int z1_T(ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  if constexpr (T * Geo<L>::ST * (int)sizeof(cplx_t<R>) > kLdsBudget || Geo<L>::P * T > 1024) {
    return 0;
  } else {
    dim3 grid((a.Hx + T - 1) / T, Ly);
    hipLaunchKernelGGL((k_z1_f<R, L, T>), grid, dim3(Geo<L>::P * T), 0, st, a, tw);
    return hipGetLastError() == hipSuccess ? 1 : -1;
  }
}
// register-resident scalar z pass, TC columns per block
template <typename R, int L, int TC> int z1r_T(ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  if constexpr (!regs_ok<L>() || TC * reg_line_stride<L, sizeof(cplx_t<R>), TC>() * (int)sizeof(cplx_t<R>) > kLdsBudget ||
                Geo<L>::P * TC > 1024) {
    return 0;
  } else {
    if ((double)a.Mz * (double)a.rs * sizeof(cplx_t<R>) >= 2147483647.0) return 0;  // 31-bit buffer offsets
    dim3 grid((a.Hx + TC - 1) / TC, Ly);
    if (a.allper)
      hipLaunchKernelGGL((k_z1_r<R, L, TC, true, false>), grid, dim3(Geo<L>::P * TC), 0, st, a, tw);
    else if (2 * a.Mz == L && (L / Geo<L>::P) % 2 == 0)
      hipLaunchKernelGGL((k_z1_r<R, L, TC, false, true>), grid, dim3(Geo<L>::P * TC), 0, st, a, tw);
    else
      hipLaunchKernelGGL((k_z1_r<R, L, TC, false, false>), grid, dim3(Geo<L>::P * TC), 0, st, a, tw);
    return hipGetLastError() == hipSuccess ? 1 : -1;
  }
}
// whether z1_T / z1r_T have that variant for length L (without launching)
template <typename R, int L, int T> constexpr bool z1_fits() {
  return T * Geo<L>::ST * (int)sizeof(cplx_t<R>) <= kLdsBudget && Geo<L>::P * T <= 1024;
}
template <typename R, int L, int TC> constexpr bool z1r_fits() {
  return regs_ok<L>() && TC * reg_line_stride<L, sizeof(cplx_t<R>), TC>() * (int)sizeof(cplx_t<R>) <= kLdsBudget &&
         Geo<L>::P * TC <= 1024;
}
template <typename R, int L> int z1_avail(int T) {
  switch (T) {
    case 1: return z1_fits<R, L, 1>();
    case 2: return z1_fits<R, L, 2>();
    case 4: return z1_fits<R, L, 4>();
    case 8: return z1_fits<R, L, 8>();
    case 16: return z1_fits<R, L, 16>();
    case kZ1Reg + 4: return z1r_fits<R, L, 4>();
    case kZ1Reg + 8: return z1r_fits<R, L, 8>();
    case kZ1Reg + 16: return z1r_fits<R, L, 16>();
    default: return 0;
  }
}
// T selects the kernel: 1..16 the LDS-resident one with T columns per
// block, kZ1Reg + TC the register-resident one with TC columns per block
template <typename R, int L> int z1_L(int T, ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  switch (T) {
    case 1: return z1_T<R, L, 1>(a, Ly, tw, st);
    case 2: return z1_T<R, L, 2>(a, Ly, tw, st);
    case 4: return z1_T<R, L, 4>(a, Ly, tw, st);
    case 8: return z1_T<R, L, 8>(a, Ly, tw, st);
    case 16: return z1_T<R, L, 16>(a, Ly, tw, st);
    case kZ1Reg + 4: return z1r_T<R, L, 4>(a, Ly, tw, st);
    case kZ1Reg + 8: return z1r_T<R, L, 8>(a, Ly, tw, st);
    case kZ1Reg + 16: return z1r_T<R, L, 16>(a, Ly, tw, st);
    default: return 0;
  }
}

// ---- group launches (one kernel instance over several patches; nb tiles)
template <typename R, int NH> int grp_xfwd_L(const GrpEntry<XFwdArgs> *t, int n, int64_t nb, hipStream_t st) {
  constexpr int T = tile_lines<R, NH, 1>();
  hipLaunchKernelGGL((kg_x_fwd_f<R, NH, T, 1>), dim3((unsigned)nb), dim3(Geo<NH>::P * T), 0, st, t, n, nb);
  return hipGetLastError() == hipSuccess ? 1 : -1;
}
template <typename R, int NH> int grp_xinv_L(const GrpEntry<XInvArgs> *t, int n, int64_t nb, hipStream_t st) {
  constexpr int T = tile_lines<R, NH, 1>();
  hipLaunchKernelGGL((kg_x_inv_f<R, NH, T, 1>), dim3((unsigned)nb), dim3(Geo<NH>::P * T), 0, st, t, n, nb);
  return hipGetLastError() == hipSuccess ? 1 : -1;
}
// the same kernel choice as strided_L
template <typename R, int L>
int grp_strided_L(int S, int kernel, bool half, const GrpEntry<StridedArgs> *t, int n, int64_t nb, hipStream_t st) {
  constexpr int T = tile_cols<R, L>();
  const dim3 grid((unsigned)nb), block(Geo<L>::P * T);
  if constexpr (regs_ok<L>()) {
    if (kernel == 0) {
      if (S < 0 && half)
        hipLaunchKernelGGL((kg_strided_r<R, L, T, -1, true>), grid, block, 0, st, t, n, nb);
      else if (S < 0)
        hipLaunchKernelGGL((kg_strided_r<R, L, T, -1, false>), grid, block, 0, st, t, n, nb);
      else if (half)
        hipLaunchKernelGGL((kg_strided_r<R, L, T, 1, true>), grid, block, 0, st, t, n, nb);
      else
        hipLaunchKernelGGL((kg_strided_r<R, L, T, 1, false>), grid, block, 0, st, t, n, nb);
      return hipGetLastError() == hipSuccess ? 1 : -1;
    }
  }
  if (S < 0)
    hipLaunchKernelGGL((kg_strided_f<R, L, T, -1>), grid, block, 0, st, t, n, nb);
  else
    hipLaunchKernelGGL((kg_strided_f<R, L, T, 1>), grid, block, 0, st, t, n, nb);
  return hipGetLastError() == hipSuccess ? 1 : -1;
}
// the same kernel choice and availability as z1_T / z1r_T
template <typename R, int L, int T>
int grp_z1_T(const GrpEntry<ZArgs> *t, int n, int64_t nb, hipStream_t st) {
  if constexpr (T * Geo<L>::ST * (int)sizeof(cplx_t<R>) > kLdsBudget || Geo<L>::P * T > 1024) {
    return 0;
  } else {
    hipLaunchKernelGGL((kg_z1_f<R, L, T>), dim3((unsigned)nb), dim3(Geo<L>::P * T), 0, st, t, n, nb);
    return hipGetLastError() == hipSuccess ? 1 : -1;
  }
}
template <typename R, int L, int TC>
int grp_z1r_T(bool allper, bool half, const GrpEntry<ZArgs> *t, int n, int64_t nb, hipStream_t st) {
  if constexpr (!regs_ok<L>() || TC * reg_line_stride<L, sizeof(cplx_t<R>), TC>() * (int)sizeof(cplx_t<R>) > kLdsBudget ||
                Geo<L>::P * TC > 1024) {
    return 0;
  } else {
    const dim3 grid((unsigned)nb), block(Geo<L>::P * TC);
    if (allper)
      hipLaunchKernelGGL((kg_z1_r<R, L, TC, true, false>), grid, block, 0, st, t, n, nb);
    else if (half)
      hipLaunchKernelGGL((kg_z1_r<R, L, TC, false, true>), grid, block, 0, st, t, n, nb);
    else
      hipLaunchKernelGGL((kg_z1_r<R, L, TC, false, false>), grid, block, 0, st, t, n, nb);
    return hipGetLastError() == hipSuccess ? 1 : -1;
  }
}
template <typename R, int L>
int grp_z1_L(int T, bool allper, bool half, const GrpEntry<ZArgs> *t, int n, int64_t nb, hipStream_t st) {
  switch (T) {
    case 1: return grp_z1_T<R, L, 1>(t, n, nb, st);
    case 2: return grp_z1_T<R, L, 2>(t, n, nb, st);
    case 4: return grp_z1_T<R, L, 4>(t, n, nb, st);
    case 8: return grp_z1_T<R, L, 8>(t, n, nb, st);
    case 16: return grp_z1_T<R, L, 16>(t, n, nb, st);
    case kZ1Reg + 4: return grp_z1r_T<R, L, 4>(allper, half, t, n, nb, st);
    case kZ1Reg + 8: return grp_z1r_T<R, L, 8>(allper, half, t, n, nb, st);
    case kZ1Reg + 16: return grp_z1r_T<R, L, 16>(allper, half, t, n, nb, st);
    default: return 0;
  }
}

// the launcher table of one transform length
template <typename R, int L> const FxOps<R> *fx_ops() {
  static const FxOps<R> ops{&strided_L<R, L>,      &z_L<R, L>,          &xfwd_L<R, L>,       &xinv_L<R, L>,
                            &z1_L<R, L>,           &z1_avail<R, L>,     tile_lines<R, L, 1>(), tile_cols<R, L>(),
                            (L / Geo<L>::P) % 2 == 0 ? 1 : 0, &grp_xfwd_L<R, L>, &grp_xinv_L<R, L>,
                            &grp_strided_L<R, L>, &grp_z1_L<R, L>};
  return &ops;
}

}  // namespace fx
}  // namespace pmgpu

#define PM_FX_INSTANTIATE(n)                                                  \
  namespace pmgpu {                                                           \
  namespace fx {                                                              \
  template const FxOps<float> *fx_ops<float, n>();                            \
  template const FxOps<double> *fx_ops<double, n>();                          \
  }                                                                           \
  }
