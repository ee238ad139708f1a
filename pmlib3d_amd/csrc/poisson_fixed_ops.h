// Table of the length-specialised launchers of one transform length
// (instantiated per length in fx_len_*.hip; dispatched by poisson_fixed.hip).
#pragma once
#include "poisson_kernels.h"

namespace pmgpu {
namespace fx {

template <typename R> struct FxOps {
  int (*strided)(int S, StridedArgs a, int nb_y, int nb_z, const cplx_t<R> *tw, hipStream_t st);
  int (*z)(ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st);
  int (*xfwd)(int NC, XFwdArgs a, const cplx_t<R> *tw, hipStream_t st);
  int (*xinv)(XInvArgs a, const cplx_t<R> *tw, hipStream_t st);
  int (*z1)(int T, ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st);
  int (*z1_avail)(int T);
  int tile_x1;    // lines per block of the one-component x passes
  int tile_cols;  // columns per block of the strided passes
  int even_rows;  // L / P even: the register kernels skip the zero half
  int (*grp_xfwd)(const GrpEntry<XFwdArgs> *t, int n, int64_t nb, hipStream_t st);
  int (*grp_xinv)(const GrpEntry<XInvArgs> *t, int n, int64_t nb, hipStream_t st);
  int (*grp_strided)(int S, int kernel, bool half, const GrpEntry<StridedArgs> *t, int n, int64_t nb,
                     hipStream_t st);
  int (*grp_z1)(int T, bool allper, bool half, const GrpEntry<ZArgs> *t, int n, int64_t nb, hipStream_t st);
};

// defined in poisson_fixed_launch.h, instantiated for each length below
template <typename R, int L> const FxOps<R> *fx_ops();

}  // namespace fx
}  // namespace pmgpu

// the specialised lengths: factors 2, 3, 5 up to 1024
#define PM_FX_LENGTHS(X) \
  X(16) X(32) X(48) X(64) X(80) X(96) X(128) X(160) X(192) X(256) X(320) X(384) X(512) X(640) X(768) X(1024)
