// Dispatch of the length-specialised Poisson launchers: each entry point
// returns 1 when it launched, 0 when the length has no specialisation (the
// caller then uses the runtime-planned kernels of poisson_kernels.h), -1 on
// a launch error.  The launchers themselves are compiled per length
// (fx_len_*.hip).
#include "poisson_fixed_ops.h"

namespace pmgpu {
namespace {

template <typename R> const fx::FxOps<R> *ops(int L) {
  switch (L) {
#define X(n) \
  case n: return fx::fx_ops<R, n>();
    PM_FX_LENGTHS(X)
#undef X
    default: return nullptr;
  }
}

}  // namespace

template <typename R>
int fx_strided(int L, int S, StridedArgs a, int nb_y, int nb_z, const cplx_t<R> *tw, hipStream_t st) {
  const auto *o = ops<R>(L);
  return o ? o->strided(S, a, nb_y, nb_z, tw, st) : 0;
}
template <typename R> int fx_z(int L, ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  const auto *o = ops<R>(L);
  return o ? o->z(a, Ly, tw, st) : 0;
}
template <typename R> int fx_xfwd(int NH, int NC, XFwdArgs a, const cplx_t<R> *tw, hipStream_t st) {
  const auto *o = ops<R>(NH);
  return o ? o->xfwd(NC, a, tw, st) : 0;
}
template <typename R> int fx_z1(int L, int T, ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st) {
  const auto *o = ops<R>(L);
  return o ? o->z1(T, a, Ly, tw, st) : 0;
}
template <typename R> int fx_z1_avail(int L, int T) {
  const auto *o = ops<R>(L);
  return o ? o->z1_avail(T) : 0;
}
template <typename R> int fx_xinv(int NH, XInvArgs a, const cplx_t<R> *tw, hipStream_t st) {
  const auto *o = ops<R>(NH);
  return o ? o->xinv(a, tw, st) : 0;
}

// lines per block of the one-component x passes, columns per block of the
// strided passes (0: length not specialised); whether L / P is even (the
// zero-half skipping of the register kernels applies)
template <typename R> int fx_tile_x1(int NH) {
  const auto *o = ops<R>(NH);
  return o ? o->tile_x1 : 0;
}
template <typename R> int fx_tile_cols(int L) {
  const auto *o = ops<R>(L);
  return o ? o->tile_cols : 0;
}
int fx_even_rows(int L) {
  const auto *o = ops<float>(L);
  return o ? o->even_rows : 0;
}
template <typename R> int fx_grp_xfwd(int NH, const GrpEntry<XFwdArgs> *t, int n, int64_t nb, hipStream_t st) {
  const auto *o = ops<R>(NH);
  return o ? o->grp_xfwd(t, n, nb, st) : 0;
}
template <typename R> int fx_grp_xinv(int NH, const GrpEntry<XInvArgs> *t, int n, int64_t nb, hipStream_t st) {
  const auto *o = ops<R>(NH);
  return o ? o->grp_xinv(t, n, nb, st) : 0;
}
template <typename R>
int fx_grp_strided(int L, int S, int kernel, bool half, const GrpEntry<StridedArgs> *t, int n, int64_t nb,
                   hipStream_t st) {
  const auto *o = ops<R>(L);
  return o ? o->grp_strided(S, kernel, half, t, n, nb, st) : 0;
}
template <typename R>
int fx_grp_z1(int L, int T, bool allper, bool half, const GrpEntry<ZArgs> *t, int n, int64_t nb, hipStream_t st) {
  const auto *o = ops<R>(L);
  return o ? o->grp_z1(T, allper, half, t, n, nb, st) : 0;
}

template int fx_tile_x1<float>(int);
template int fx_tile_x1<double>(int);
template int fx_tile_cols<float>(int);
template int fx_tile_cols<double>(int);
template int fx_grp_xfwd<float>(int, const GrpEntry<XFwdArgs> *, int, int64_t, hipStream_t);
template int fx_grp_xfwd<double>(int, const GrpEntry<XFwdArgs> *, int, int64_t, hipStream_t);
template int fx_grp_xinv<float>(int, const GrpEntry<XInvArgs> *, int, int64_t, hipStream_t);
template int fx_grp_xinv<double>(int, const GrpEntry<XInvArgs> *, int, int64_t, hipStream_t);
template int fx_grp_strided<float>(int, int, int, bool, const GrpEntry<StridedArgs> *, int, int64_t, hipStream_t);
template int fx_grp_strided<double>(int, int, int, bool, const GrpEntry<StridedArgs> *, int, int64_t, hipStream_t);
template int fx_grp_z1<float>(int, int, bool, bool, const GrpEntry<ZArgs> *, int, int64_t, hipStream_t);
template int fx_grp_z1<double>(int, int, bool, bool, const GrpEntry<ZArgs> *, int, int64_t, hipStream_t);
template int fx_strided<float>(int, int, StridedArgs, int, int, const float2 *, hipStream_t);
template int fx_strided<double>(int, int, StridedArgs, int, int, const double2 *, hipStream_t);
template int fx_z<float>(int, ZArgs, int, const float2 *, hipStream_t);
template int fx_z<double>(int, ZArgs, int, const double2 *, hipStream_t);
template int fx_xfwd<float>(int, int, XFwdArgs, const float2 *, hipStream_t);
template int fx_xfwd<double>(int, int, XFwdArgs, const double2 *, hipStream_t);
template int fx_xinv<float>(int, XInvArgs, const float2 *, hipStream_t);
template int fx_xinv<double>(int, XInvArgs, const double2 *, hipStream_t);
template int fx_z1<float>(int, int, ZArgs, int, const float2 *, hipStream_t);
template int fx_z1_avail<float>(int, int);
template int fx_z1_avail<double>(int, int);
template int fx_z1<double>(int, int, ZArgs, int, const double2 *, hipStream_t);

}  // namespace pmgpu
