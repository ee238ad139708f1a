// Host side of the Poisson solve: plans, buffers, launches, C ABI.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <cstdio>
#include <algorithm>
#include <map>
#include <type_traits>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pmgpu.h"
#include "poisson_kernels.h"

namespace pmgpu {

int device_lds_limit();  // capi.hip

// length-specialised launchers (poisson_fixed.hip): 1 launched, 0 no
// specialisation for this length, -1 launch error
template <typename R>
int fx_strided(int L, int S, StridedArgs a, int nb_y, int nb_z, const cplx_t<R> *tw, hipStream_t st);
template <typename R> int fx_z(int L, ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st);
template <typename R> int fx_xfwd(int NH, int NC, XFwdArgs a, const cplx_t<R> *tw, hipStream_t st);
template <typename R> int fx_xinv(int NH, XInvArgs a, const cplx_t<R> *tw, hipStream_t st);
template <typename R> int fx_z1(int L, int T, ZArgs a, int Ly, const cplx_t<R> *tw, hipStream_t st);
// 1 when fx_z1 has kernel variant T for length L (checked without launching)
template <typename R> int fx_z1_avail(int L, int T);
// group launches over several patches (same return convention) and the
// tile sizes their tables are built with (0: length not specialised)
template <typename R> int fx_tile_x1(int NH);
template <typename R> int fx_tile_cols(int L);
int fx_even_rows(int L);
template <typename R> int fx_grp_xfwd(int NH, const GrpEntry<XFwdArgs> *t, int n, int64_t nb, hipStream_t st);
template <typename R> int fx_grp_xinv(int NH, const GrpEntry<XInvArgs> *t, int n, int64_t nb, hipStream_t st);
template <typename R>
int fx_grp_strided(int L, int S, int kernel, bool half, const GrpEntry<StridedArgs> *t, int n, int64_t nb,
                   hipStream_t st);
template <typename R>
int fx_grp_z1(int L, int T, bool allper, bool half, const GrpEntry<ZArgs> *t, int n, int64_t nb, hipStream_t st);

// Tile widths chosen by timing, per (pass, lengths, precision), shared by all
// handles of the process and kept across runs in the file named by
// PMGPU_TUNE_CACHE (one "key value" line per entry) when that is set.
struct TuneCache {
  std::map<std::string, int> m;
  bool loaded = false;
  void load() {
    if (loaded) return;
    loaded = true;
    const char *f = getenv("PMGPU_TUNE_CACHE");
    if (!f) return;
    FILE *fp = fopen(f, "r");
    if (!fp) return;
    char key[256];
    int v;
    while (fscanf(fp, "%255s %d", key, &v) == 2) m[key] = v;
    fclose(fp);
  }
  int get(const std::string &k) {
    load();
    auto it = m.find(k);
    return it == m.end() ? 0 : it->second;
  }
  void put(const std::string &k, int v) {
    m[k] = v;
    const char *f = getenv("PMGPU_TUNE_CACHE");
    if (!f) return;
    FILE *fp = fopen(f, "a");
    if (!fp) return;
    fprintf(fp, "%s %d\n", k.c_str(), v);
    fclose(fp);
  }
};
static TuneCache &tune_cache() {
  static TuneCache c;
  return c;
}

// Factor L into the supported radices, largest first.
static bool make_plan(int L, FftPlan &pl) {
  pl.L = L;
  pl.P = (L + 7) / 8;
  pl.nstage = 0;
  int r = L;
  const int order[] = {8, 4, 2, 3, 5, 7, 11, 13};
  for (int rad : order) {
    while (r % rad == 0 && r > 1) {
      if (pl.nstage >= kMaxStages) return false;
      pl.radix[pl.nstage++] = rad;
      r /= rad;
    }
  }
  return r == 1;
}

// Twiddle tables exp(-2 pi i n / L) for a set of lengths, in one buffer.
template <typename R> struct TwTable {
  std::map<int, int> off;
  std::vector<cplx_t<R>> host;
  int add(int L) {
    auto it = off.find(L);
    if (it != off.end()) return it->second;
    int o = (int)host.size();
    off[L] = o;
    for (int n = 0; n < L; ++n) {
      long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)n / (long double)L;
      host.push_back(cmake<cplx_t<R>>((R)cosl(a), (R)sinl(a)));
    }
    return o;
  }
};

// The five passes of one handle's single-component solve as group-launch
// entries (arguments, twiddles, tiles) plus what selects their kernel
// instance: patches whose keys match share one launch per pass.
struct ScalarPasses {
  bool ok = false;   // every pass has a length-specialised kernel
  bool fp64 = false;
  bool ghost = false;  // periodic axes: ghost fill after the x inverse
  int NH = 0, Ly = 0, Lz = 0, tz1 = 0, skernel = 0;
  bool y_half = false, z_half = false, allper = false;
  GrpEntry<XFwdArgs> xf{};
  GrpEntry<StridedArgs> yf{};
  GrpEntry<ZArgs> z{};
  GrpEntry<StridedArgs> yi{};
  GrpEntry<XInvArgs> xi{};
  int64_t tiles[5] = {0, 0, 0, 0, 0};  // per pass: x fwd, y fwd, z, y inv, x inv
};

struct PoissonBase {
  virtual int scalar_passes(const void *w, void *psi, ScalarPasses &sp) = 0;
  virtual int scalar_ghost_fill(void *psi) = 0;
  virtual int setup() = 0;
  virtual int solve_device(const void *vort, void *vel, void *vort_out, void *vecpot, int reg, int reproj,
                           int psi) = 0;
  virtual int solve_host(const double *vort, double *vel, double *vort_out, double *vecpot, int reg,
                         int reproj, int psi) = 0;
  virtual int octant(void *host, int64_t cap, int32_t *dims) = 0;
  virtual int solve_scalar_device(const void *w, void *psi) = 0;
  virtual int solve_scalar_host(const double *w, double *psi) = 0;
  virtual int tune_scalar() = 0;
  virtual int precision() const = 0;
  pmgpu_geom g;
  pmgpu_params prm;
  // a = 2^(level - parent level) of a refinement patch (0: base level); the
  // kernel becomes G(eps) - G(a eps) at the next setup
  double refine = 0.0;
  hipStream_t stream = nullptr;
  // HIP-event timing of the fused z pass and of whole solves, recorded on
  // the handle's stream without host synchronisation; read back on demand
  bool timing = false;
  std::vector<hipEvent_t> evs;  // 4 per recorded solve
  int nrec = 0;
  static constexpr int kMaxRec = 512;
  int record(int slot) {
    if (!timing) return 0;
    if (nrec >= kMaxRec) return 0;
    size_t i = (size_t)nrec * 4 + slot;
    while (evs.size() <= i) {
      hipEvent_t e;
      PM_HIP(hipEventCreate(&e));
      evs.push_back(e);
    }
    PM_HIP(hipEventRecord(evs[i], stream));
    if (slot == 3) ++nrec;
    return 0;
  }
  int averages(double *z_ms, double *tot_ms) {
    *z_ms = -1;
    *tot_ms = -1;
    if (nrec == 0) return 0;
    PM_HIP(hipEventSynchronize(evs[(size_t)nrec * 4 - 1]));
    double za = 0, ta = 0;
    for (int r = 0; r < nrec; ++r) {
      float a = 0, b = 0;
      PM_HIP(hipEventElapsedTime(&a, evs[r * 4 + 1], evs[r * 4 + 2]));
      PM_HIP(hipEventElapsedTime(&b, evs[r * 4 + 0], evs[r * 4 + 3]));
      za += a;
      ta += b;
    }
    *z_ms = za / nrec;
    *tot_ms = ta / nrec;
    return 0;
  }
  virtual ~PoissonBase() {
    for (auto e : evs) (void)hipEventDestroy(e);
  }
};

template <typename R> struct PoissonGPU : PoissonBase {
  using V = cplx_t<R>;
  int M[3], L[3], E[3], lo[3], ngs[3];
  int Nh, Hx, Hy, Hz;
  bool allper = false;
  FftPlan pNh, pLy, pLz;
  int twLx_off = 0;
  TwTable<R> tw;
  V *d_tw = nullptr;
  V *A = nullptr, *B = nullptr, *Bpsi = nullptr, *Bvort = nullptr;
  R *Ghat = nullptr;
  V *Khat = nullptr;  // kernel 2: complex half-spectrum of grad G, 3 components
  bool ker2 = false;  // kernel 2 on a patch with a free axis (all-periodic: spectral)
  R *mesh_in = nullptr, *mesh_out = nullptr;  // staging for host calls
  size_t nA = 0, nB = 0, nmesh = 0;
  int Tx = 1, Ts = 1, Tz = 1;
  int64_t H3p = 0;   // padded row length (complex) of the interleaved spectra
  int tune_tc = 0;   // z-pass columns per block (0 = built-in default)
  int zkernel = 0;   // 0: register-resident velocity z pass where available, 1: LDS-resident
  int skernel = 0;   // the same choice for the strided (y and z) passes
  int lds_limit = 65536;

  int precision() const override { return sizeof(R) == 8 ? PMGPU_FP64 : PMGPU_FP32; }

  ~PoissonGPU() override {
    void *bufs[] = {d_tw, A, B, Bpsi, Bvort, Ghat, Khat, mesh_in, mesh_out};
    for (void *b : bufs)
      if (b) (void)hipFree(b);
  }

  int init(const pmgpu_geom &gg, const pmgpu_params &pp, hipStream_t s) {
    g = gg;
    prm = pp;
    stream = s;
    PM_CHECK(prm.poisson_kernel == 1 || prm.poisson_kernel == 2, "pmgpu: poisson_kernel must be 1 or 2");
    int o = prm.poisson_order;
    PM_CHECK(o == 0 || o == 2 || o == 4 || o == 6 || o == 8 || o == 10 || o == 100,
             "pmgpu: poisson_order must be 0, 2, 4, 6, 8, 10 or 100 (parameters/pmlib_parameters_check.f:52-63)");
    PM_CHECK(prm.poisson_order == 0 || prm.regularisation_radius >= 1.0,
             "pmgpu: regularisation_radius must be >= 1 (parameters/pmlib_parameters_check.f:67)");
    allper = g.bound_cond[0] == 1 && g.bound_cond[1] == 1 && g.bound_cond[2] == 1;
    // all-periodic kernel 2 is K = zeta i k / (2 pi |k|^2) = G 2 pi i k with
    // the kernel-1 spectral G: the same z pass serves both
    ker2 = prm.poisson_kernel == 2 && !allper;
    for (int d = 0; d < 3; ++d) {
      PM_CHECK(g.ncell[d] > 0 && g.dx[d] > 0, "pmgpu: bad geometry");
      int ng = g.bound_cond[d] == 0 ? g.nghost : (g.interp_order == 4 ? 3 : 2);
      ngs[d] = ng;
      E[d] = g.ncell[d] + 2 * ng;
      if (g.bound_cond[d] == 0) {
        M[d] = g.ncell[d] + 2 * g.nghost;
        L[d] = 2 * M[d];
        lo[d] = 0;
      } else {
        M[d] = g.ncell[d];
        L[d] = M[d];
        lo[d] = ng;
        PM_CHECK(g.ncell[d] >= ng, "pmgpu: periodic axis shorter than its ghost layer");
      }
      PM_CHECK(L[d] % 2 == 0, "pmgpu: nfft is an odd number (fourier/pmlib_fourier_fft.f:126)");
    }
    Nh = L[0] / 2;
    Hx = Nh + 1;
    Hy = L[1] / 2 + 1;
    Hz = L[2] / 2 + 1;
    PM_CHECK(make_plan(Nh, pNh) && make_plan(L[1], pLy) && make_plan(L[2], pLz),
             "pmgpu: FFT length with a prime factor > 13 is not supported");
    pNh.tw_off = tw.add(Nh);
    pLy.tw_off = tw.add(L[1]);
    pLz.tw_off = tw.add(L[2]);
    twLx_off = tw.add(L[0]);
    PM_HIP(hipMalloc(&d_tw, tw.host.size() * sizeof(V)));
    PM_HIP(hipMemcpy(d_tw, tw.host.data(), tw.host.size() * sizeof(V), hipMemcpyHostToDevice));
    // interleaved spectral rows padded to 128 bytes so tiles align with cache lines
    const int64_t align = 128 / (int64_t)sizeof(V);
    H3p = ((3 * (int64_t)Hx + align - 1) / align) * align;
    nA = (size_t)H3p * M[1] * M[2];
    nB = (size_t)H3p * L[1] * M[2];
    if (const char *env = getenv("PMGPU_Z_TC")) tune_tc = atoi(env);
    if (const char *env = getenv("PMGPU_Z_KERNEL")) zkernel = atoi(env);
    if (const char *env = getenv("PMGPU_STRIDED_KERNEL")) skernel = atoi(env);
    nmesh = (size_t)3 * E[0] * E[1] * E[2];
    PM_HIP(hipMalloc(&A, nA * sizeof(V)));
    PM_HIP(hipMalloc(&B, nB * sizeof(V)));
    lds_limit = device_lds_limit();
    // tile widths: 128-byte rows for the strided passes where LDS allows
    const int tmax = sizeof(R) == 4 ? 16 : 8;
    auto pick = [&](int len, int lines_per_unit, int tcap) {
      int t = tcap;
      while (t > 1 && ((size_t)t * lines_per_unit * lds_stride(len) * sizeof(V) > (size_t)lds_limit ||
                       t * lines_per_unit * ((len + 7) / 8) > 1024))
        t /= 2;
      return t;
    };
    Ts = pick(L[1] > L[2] ? L[1] : L[2], 1, tmax);
    Tz = pick(L[2], 1, tmax);
    Tx = pick(Nh, 3, 8);
    return 0;
  }

  size_t lds_bytes(int len, int lines) const { return (size_t)lines * lds_stride(len) * sizeof(V); }

  // ---- generic launches -------------------------------------------------
  template <typename RR>
  int launch_xfwd(const void *in, void *out, int64_t cs, int64_t xs, int64_t ys, int64_t zs, int64_t o_cs,
                  int64_t o_ks, int64_t o_ys, int64_t o_zs, int ncomp, int Min, int My, int Mz, int T,
                  const FftPlan &pl, int twL, const cplx_t<RR> *twp) {
    XFwdArgs a{in, out, cs, xs, ys, zs, o_cs, o_ks, o_ys, o_zs, ncomp, Min, My, Mz, pl.L + 1, T, pl, twL};
    int f = fx_xfwd<RR>(pl.L, ncomp, a, twp, stream);
    PM_CHECK(f >= 0, "pmgpu: x-pass launch failed");
    if (f == 1) return 0;
    dim3 grid((My + T - 1) / T, Mz);
    size_t sh = (size_t)T * ncomp * lds_stride(pl.L) * sizeof(cplx_t<RR>);
    hipLaunchKernelGGL(k_x_fwd<RR>, grid, dim3(pl.P * T * ncomp), sh, stream, a, twp);
    PM_HIP(hipGetLastError());
    return 0;
  }
  template <typename RR, int S>
  int launch_strided(const void *in, void *out, int64_t rs_in, int64_t rs_out, int64_t bs_in, int64_t bs_out,
                     int ncols, int nin, int nout, int nb_y, int nb_z, double scale, int T, const FftPlan &pl,
                     const cplx_t<RR> *twp) {
    StridedArgs a{in, out, rs_in, rs_out, bs_in, bs_out, ncols, nin, nout, T, scale, pl, skernel};
    int f = fx_strided<RR>(pl.L, S, a, nb_y, nb_z, twp, stream);
    PM_CHECK(f >= 0, "pmgpu: strided-pass launch failed");
    if (f == 1) return 0;
    dim3 grid((ncols + T - 1) / T, nb_y, nb_z);
    size_t sh = (size_t)T * lds_stride(pl.L) * sizeof(cplx_t<RR>);
    hipLaunchKernelGGL((k_strided<RR, S>), grid, dim3(pl.P * T), sh, stream, a, twp);
    PM_HIP(hipGetLastError());
    return 0;
  }

  // ---- pmlib_poisson_setup ---------------------------------------------
  int setup() override {
    PM_CHECK(refine == 0.0 || prm.poisson_order != 0,
             "pmgpu_poisson_setup: Regularisation must be used for multi-patch simulations "
             "(poisson/pmlib_poisson_setup.f:335-339)");
    PM_CHECK(refine == 0.0 || prm.poisson_order != 100,
             "pmgpu_poisson_setup: Gaussian regularisation must be used for multi-patch simulations "
             "(poisson/pmlib_poisson_setup.f:1514-1518)");
    if (allper) return tune_vector();  // G = zeta / (4 pi^2 k^2) is formed in the z pass
    if (ker2) return setup_kernel2();
    // FP64 kernel grid and its transform (x R2C, y and z C2C), then the octant
    double *Gr = nullptr;
    double2 *Gc = nullptr;
    size_t nG = (size_t)L[0] * L[1] * L[2];
    size_t nGc = (size_t)Hx * L[1] * L[2];
    PM_HIP(hipMalloc(&Gr, nG * sizeof(double)));
    PM_HIP(hipMalloc(&Gc, nGc * sizeof(double2)));
    GreenArgs ga;
    ga.out = Gr;
    double dmax = g.dx[0] > g.dx[1] ? g.dx[0] : g.dx[1];
    dmax = dmax > g.dx[2] ? dmax : g.dx[2];
    ga.eps = prm.regularisation_radius * dmax;
    ga.order = prm.poisson_order;
    ga.ca = refine > 0.0 ? 1.0 / refine : 0.0;
    ga.hmax = dmax;
    for (int d = 0; d < 3; ++d) {
      ga.L[d] = L[d];
      ga.bc[d] = g.bound_cond[d];
      ga.dx[d] = g.dx[d];
      ga.xk[d] = -(0.5 * (double)L[d] + 0.5) * g.dx[d];
      ga.len[d] = g.dx[d] * (double)L[d];
    }
    hipLaunchKernelGGL(k_green, dim3(4096), dim3(256), 0, stream, ga);
    PM_HIP(hipGetLastError());
    TwTable<double> t64;
    FftPlan q0, q1, q2;
    make_plan(Nh, q0);
    make_plan(L[1], q1);
    make_plan(L[2], q2);
    q0.tw_off = t64.add(Nh);
    q1.tw_off = t64.add(L[1]);
    q2.tw_off = t64.add(L[2]);
    int tl = t64.add(L[0]);
    double2 *dt = nullptr;
    PM_HIP(hipMalloc(&dt, t64.host.size() * sizeof(double2)));
    PM_HIP(hipMemcpyAsync(dt, t64.host.data(), t64.host.size() * sizeof(double2), hipMemcpyHostToDevice, stream));
    auto pick64 = [&](int len, int lines_per_unit, int tcap) {
      int t = tcap;
      while (t > 1 && ((size_t)t * lines_per_unit * lds_stride(len) * sizeof(double2) > (size_t)lds_limit ||
                       t * lines_per_unit * ((len + 7) / 8) > 1024))
        t /= 2;
      return t;
    };
    int tx = pick64(Nh, 1, 8), ty = pick64(L[1], 1, 8), tz = pick64(L[2], 1, 8);
    int e = 0;
    e |= launch_xfwd<double>(Gr, Gc, 0, 1, L[0], (int64_t)L[0] * L[1], 0, 1, Hx, (int64_t)Hx * L[1], 1, L[0],
                             L[1], L[2], tx, q0, tl, dt);
    e |= launch_strided<double, -1>(Gc, Gc, Hx, Hx, (int64_t)L[1] * Hx, (int64_t)L[1] * Hx, Hx, L[1], L[1], L[2],
                                    1, 1.0, ty, q1, dt);
    e |= launch_strided<double, -1>(Gc, Gc, (int64_t)L[1] * Hx, (int64_t)L[1] * Hx, Hx, Hx, Hx, L[2], L[2], L[1],
                                    1, 1.0, tz, q2, dt);
    if (e) return -1;
    if (!Ghat) PM_HIP(hipMalloc(&Ghat, (size_t)Hx * Hy * Hz * sizeof(R)));
    hipLaunchKernelGGL(k_g_octant<R>, dim3(2048), dim3(256), 0, stream, Gc, Ghat, Hx, L[1], Hy, Hz);
    PM_HIP(hipGetLastError());
    PM_HIP(hipStreamSynchronize(stream));
    (void)hipFree(Gr);
    (void)hipFree(Gc);
    (void)hipFree(dt);
    return tune_vector();
  }

  // Kernel 2: grad G on the kernel grid (three interleaved components), its
  // shifted forward transform (x R2C, y and z C2C) in FP64, kept as the
  // complex half-spectrum [((kz * Ly + ky) * Hx + kx) * 3 + c] in the
  // working precision (poisson/pmlib_poisson_setup.f:580-635, :1533-1535).
  int setup_kernel2() {
    double *Gr = nullptr;
    double2 *Gc = nullptr;
    const size_t nG = (size_t)L[0] * L[1] * L[2];
    const size_t nGc = (size_t)Hx * L[1] * L[2];
    PM_HIP(hipMalloc(&Gr, 3 * nG * sizeof(double)));
    PM_HIP(hipMalloc(&Gc, 3 * nGc * sizeof(double2)));
    GreenArgs ga;
    ga.out = Gr;
    double dmax = g.dx[0] > g.dx[1] ? g.dx[0] : g.dx[1];
    dmax = dmax > g.dx[2] ? dmax : g.dx[2];
    ga.eps = prm.regularisation_radius * dmax;
    ga.order = prm.poisson_order;
    ga.ca = refine > 0.0 ? 1.0 / refine : 0.0;
    ga.hmax = dmax;
    for (int d = 0; d < 3; ++d) {
      ga.L[d] = L[d];
      ga.bc[d] = g.bound_cond[d];
      ga.dx[d] = g.dx[d];
      ga.xk[d] = -(0.5 * (double)L[d] + 0.5) * g.dx[d];
      ga.len[d] = g.dx[d] * (double)L[d];
    }
    hipLaunchKernelGGL(k_green2, dim3(4096), dim3(256), 0, stream, ga);
    PM_HIP(hipGetLastError());
    TwTable<double> t64;
    FftPlan q0, q1, q2;
    make_plan(Nh, q0);
    make_plan(L[1], q1);
    make_plan(L[2], q2);
    q0.tw_off = t64.add(Nh);
    q1.tw_off = t64.add(L[1]);
    q2.tw_off = t64.add(L[2]);
    int tl = t64.add(L[0]);
    double2 *dt = nullptr;
    PM_HIP(hipMalloc(&dt, t64.host.size() * sizeof(double2)));
    PM_HIP(hipMemcpyAsync(dt, t64.host.data(), t64.host.size() * sizeof(double2), hipMemcpyHostToDevice, stream));
    auto pick64 = [&](int len, int lines_per_unit, int tcap) {
      int t = tcap;
      while (t > 1 && ((size_t)t * lines_per_unit * lds_stride(len) * sizeof(double2) > (size_t)lds_limit ||
                       t * lines_per_unit * ((len + 7) / 8) > 1024))
        t /= 2;
      return t;
    };
    const int tx = pick64(Nh, 3, 8), ty = pick64(L[1], 1, 8), tz = pick64(L[2], 1, 8);
    const int64_t H3 = 3 * (int64_t)Hx;
    int e = 0;
    e |= launch_xfwd<double>(Gr, Gc, 1, 3, 3 * (int64_t)L[0], 3 * (int64_t)L[0] * L[1], 1, 3, H3, H3 * L[1], 3, L[0],
                             L[1], L[2], tx, q0, tl, dt);
    e |= launch_strided<double, -1>(Gc, Gc, H3, H3, H3 * L[1], H3 * L[1], (int)H3, L[1], L[1], L[2], 1, 1.0, ty, q1,
                                    dt);
    e |= launch_strided<double, -1>(Gc, Gc, H3 * L[1], H3 * L[1], H3, H3, (int)H3, L[2], L[2], L[1], 1, 1.0, tz, q2,
                                    dt);
    if (e) return -1;
    if (!Khat) PM_HIP(hipMalloc(&Khat, 3 * nGc * sizeof(V)));
    hipLaunchKernelGGL(k_spec_convert<R>, dim3(2048), dim3(256), 0, stream, (const double2 *)Gc, Khat,
                       (int64_t)(3 * nGc));
    PM_HIP(hipGetLastError());
    PM_HIP(hipStreamSynchronize(stream));
    (void)hipFree(Gr);
    (void)hipFree(Gc);
    (void)hipFree(dt);
    return 0;
  }

  int octant(void *host, int64_t cap, int32_t *dims) override {
    PM_CHECK(!ker2, "pmgpu: kernel 2 keeps a complex spectrum, not a real octant");
    dims[0] = Hx;
    dims[1] = Hy;
    dims[2] = Hz;
    const size_t cnt = (size_t)Hx * Hy * Hz;
    PM_CHECK(Ghat != nullptr, "pmgpu: kernel not set up");
    PM_CHECK((int64_t)(cnt * sizeof(R)) <= cap, "pmgpu: buffer too small for the kernel octant");
    // device layout [(kz' * Hy + ky') * Hx + kx]; exported as [(ky' * Hx + kx) * Hz + kz']
    std::vector<R> dev(cnt);
    PM_HIP(hipMemcpy(dev.data(), Ghat, cnt * sizeof(R), hipMemcpyDeviceToHost));
    R *out = reinterpret_cast<R *>(host);
    for (int kz = 0; kz < Hz; ++kz)
      for (int ky = 0; ky < Hy; ++ky)
        for (int kx = 0; kx < Hx; ++kx)
          out[((size_t)ky * Hx + kx) * Hz + kz] = dev[((size_t)kz * Hy + ky) * Hx + kx];
    return 0;
  }

  // arguments of the three-component z pass
  ZArgs z3_args(bool want_psi, bool want_vort, int reproj) const {
    const int64_t H3 = H3p;
    ZArgs za;
    za.B = B;
    za.Bpsi = Bpsi;
    za.Bvort = Bvort;
    za.Ghat = ker2 ? (const void *)Khat : (const void *)Ghat;
    za.ker2 = ker2 ? 1 : 0;
    za.rs = H3 * L[1];
    za.cs = 1;
    za.kxs = 3;
    za.kys = H3;
    za.Hx = Hx;
    za.Ly = L[1];
    za.Lz = L[2];
    za.Lx = L[0];
    za.Mz = M[2];
    za.T = Tz;
    za.Hy = Hy;
    za.Hz = Hz;
    za.order = prm.poisson_order;
    za.allper = allper;
    za.want_psi = want_psi;
    za.want_vort = want_vort;
    za.reproj = reproj;
    za.tune_tc = tune_tc;
    za.zkernel = zkernel;
    double dmax = g.dx[0] > g.dx[1] ? g.dx[0] : g.dx[1];
    dmax = dmax > g.dx[2] ? dmax : g.dx[2];
    for (int d = 0; d < 3; ++d) {
      za.dk[d] = 1.0 / ((double)L[d] * g.dx[d]);
      za.ksmp[d] = 1.0 / (2.0 * g.dx[d]);
    }
    za.h3 = g.dx[0] * g.dx[1] * g.dx[2];
    za.eps2pi = prm.regularisation_radius * dmax * 2.0 * 3.14159265358979323846;
    za.plan = pLz;
    return za;
  }

  // Autotuner of the three-component velocity z pass: times the register
  // form at 2, 4 and 8 columns per block (two sweeps of 5 launches, a
  // candidate's time its faster sweep) on the handle's own spectra buffer and
  // keeps the fastest, unless the built-in 4 is within 3 % of it; cached per
  // transform shape like the scalar pass.  PMGPU_Z_TC fixes the width.
  int tune_vector() {
    if (tune_tc > 0 || ker2 || zkernel == 1) return 0;
    char k[128];
    snprintf(k, sizeof(k), "z3:L%d:Ly%d:Hx%d:%s", L[2], L[1], Hx, sizeof(R) == 4 ? "f32" : "f64");
    if (int v = tune_cache().get(k)) {
      tune_tc = v;
      return 0;
    }
    PM_HIP(hipMemsetAsync(B, 0, nB * sizeof(V), stream));
    hipEvent_t e0, e1;
    PM_HIP(hipEventCreate(&e0));
    PM_HIP(hipEventCreate(&e1));
    const int cands[3] = {2, 4, 8};
    float best[3] = {1e30f, 1e30f, 1e30f};
    for (int sweep = 0; sweep < 2; ++sweep)
      for (int c = 0; c < 3; ++c) {
        ZArgs za = z3_args(false, false, 0);
        za.tune_tc = cands[c];
        if (fx_z<R>(L[2], za, L[1], d_tw, stream) != 1) continue;   // warm-up / availability
        PM_HIP(hipEventRecord(e0, stream));
        for (int r = 0; r < 5; ++r) (void)fx_z<R>(L[2], za, L[1], d_tw, stream);
        PM_HIP(hipEventRecord(e1, stream));
        PM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        PM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best[c]) best[c] = ms;
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int pick = 1;
    for (int c = 0; c < 3; ++c)
      if (best[c] < best[pick] * 0.97f) pick = c;
    if (best[pick] >= 1e30f) return 0;   // no register form at this length: the built-in path
    tune_tc = cands[pick];
    tune_cache().put(k, tune_tc);
    return 0;
  }

  // ---- pmlib_poisson_solve ---------------------------------------------
  int solve_device(const void *vort, void *vel, void *vort_out, void *vecpot, int reg, int reproj,
                   int psi) override {
    const bool want_vort = reg || reproj;
    const bool want_psi = psi && prm.poisson_kernel == 1;
    PM_CHECK(allper || (ker2 ? Khat != nullptr : Ghat != nullptr), "pmgpu: pmgpu_poisson_setup was not called");
    PM_CHECK(vort && vel, "pmgpu: null mesh pointer");
    PM_CHECK(!want_vort || vort_out, "pmgpu: vort_out required with reg_vort/reproj");
    PM_CHECK(!want_vort || prm.poisson_order != 100,
             "pmgpu: order 100 has no regularised or reprojected vorticity output "
             "(poisson/pmlib_poisson_solve.f:311-329 defines no filter for it)");
    PM_CHECK(!want_psi || vecpot, "pmgpu: vecpot required with calc_vecpot");
    if (want_psi && !Bpsi) PM_HIP(hipMalloc(&Bpsi, nB * sizeof(V)));
    if (want_vort && !Bvort) PM_HIP(hipMalloc(&Bvort, nB * sizeof(V)));
    if (record(0)) return -1;
    const int64_t off = 3 * ((int64_t)lo[0] + (int64_t)E[0] * (lo[1] + (int64_t)E[1] * lo[2]));
    const int64_t cs = 1, xs = 3, ys = 3 * (int64_t)E[0], zs = 3 * (int64_t)E[0] * E[1];
    const R *vin = reinterpret_cast<const R *>(vort) + off;
    int e = 0;
    // component-interleaved spectra: A[((z*My + y)*Hx + kx)*3 + c],
    // B[((z*Ly + ky)*Hx + kx)*3 + c]
    const int64_t H3 = H3p;
    const int ncols = 3 * Hx;
    e |= launch_xfwd<R>(vin, A, cs, xs, ys, zs, 1, 3, H3, H3 * M[1], 3, M[0], M[1], M[2], Tx, pNh, twLx_off, d_tw);
    e |= launch_strided<R, -1>(A, B, H3, H3, H3 * M[1], H3 * L[1], ncols, M[1], L[1], M[2], 1, 1.0, Ts, pLy,
                               d_tw);
    if (e) return -1;
    ZArgs za = z3_args(want_psi, want_vort, reproj);
    if (record(1)) return -1;
    int fz = fx_z<R>(L[2], za, L[1], d_tw, stream);
    PM_CHECK(fz >= 0, "pmgpu: z-pass launch failed");
    if (fz == 0) {
      dim3 zg((Hx + Tz - 1) / Tz, L[1]);
      size_t zsh = lds_bytes(L[2], Tz);
      int Q = (L[2] + pLz.P - 1) / pLz.P;
      PM_CHECK(Q <= 8, "pmgpu: internal z-pass layout error");
      hipLaunchKernelGGL((k_z_convolve<R, 8>), zg, dim3(pLz.P * Tz), zsh, stream, za, d_tw);
      PM_HIP(hipGetLastError());
    }
    if (record(2)) return -1;
    V *fields[3] = {B, want_psi ? Bpsi : nullptr, want_vort ? Bvort : nullptr};
    void *outs[3] = {vel, vecpot, vort_out};
    for (int f = 0; f < 3; ++f) {
      if (!fields[f]) continue;
      e |= launch_strided<R, 1>(fields[f], A, H3, H3, H3 * L[1], H3 * M[1], ncols, L[1], M[1], M[2], 1,
                                1.0 / (double)L[1], Ts, pLy, d_tw);
      XInvArgs xa{A,  reinterpret_cast<R *>(outs[f]) + off, cs, xs, ys, zs, 1, 3, H3, H3 * M[1], 3, M[0], M[1], M[2],
                  Hx, Tx, 1.0 / (double)L[0], pNh, twLx_off};
      int fxi = fx_xinv<R>(Nh, xa, d_tw, stream);
      PM_CHECK(fxi >= 0, "pmgpu: x inverse launch failed");
      if (fxi == 0) {
        dim3 grid((M[1] + Tx - 1) / Tx, M[2]);
        hipLaunchKernelGGL(k_x_inv<R>, grid, dim3(pNh.P * Tx * 3), lds_bytes(Nh, 3 * Tx), stream, xa, d_tw);
        PM_HIP(hipGetLastError());
      }
      for (int d = 0; d < 3; ++d)
        if (g.bound_cond[d] == 1) {
          hipLaunchKernelGGL(k_ghost_fill<R>, dim3(1024), dim3(256), 0, stream, reinterpret_cast<R *>(outs[f]),
                             E[0], E[1], E[2], d, g.ncell[d], ngs[d], 3);
          PM_HIP(hipGetLastError());
        }
    }
    if (e) return -1;
    if (record(3)) return -1;
    return 0;
  }

  int solve_host(const double *vort, double *vel, double *vort_out, double *vecpot, int reg, int reproj,
                 int psi) override {
    const bool want_vort = reg || reproj;
    const bool want_psi = psi != 0;
    if (!mesh_in) PM_HIP(hipMalloc(&mesh_in, nmesh * sizeof(R)));
    size_t nout = 1 + (want_vort ? 1 : 0) + (want_psi ? 1 : 0);
    if (!mesh_out) PM_HIP(hipMalloc(&mesh_out, 3 * nmesh * sizeof(R)));
    std::vector<R> stage(nmesh);
    const R *src;
    if (sizeof(R) == sizeof(double)) {
      src = reinterpret_cast<const R *>(vort);
    } else {
      for (size_t i = 0; i < nmesh; ++i) stage[i] = (R)vort[i];
      src = stage.data();
    }
    PM_HIP(hipMemcpyAsync(mesh_in, src, nmesh * sizeof(R), hipMemcpyHostToDevice, stream));
    // cells outside the solve domain (periodic ghosts) are refilled by the solve;
    // zero the outputs first so untouched cells are defined
    PM_HIP(hipMemsetAsync(mesh_out, 0, 3 * nmesh * sizeof(R), stream));
    R *o_vel = mesh_out, *o_psi = mesh_out + nmesh, *o_vort = mesh_out + 2 * nmesh;
    if (solve_device(mesh_in, o_vel, want_vort ? o_vort : nullptr, want_psi ? o_psi : nullptr, reg, reproj, psi))
      return -1;
    PM_HIP(hipStreamSynchronize(stream));
    (void)nout;
    double *dst[3] = {vel, want_psi ? vecpot : nullptr, want_vort ? vort_out : nullptr};
    R *from[3] = {o_vel, o_psi, o_vort};
    for (int f = 0; f < 3; ++f) {
      if (!dst[f]) continue;
      if (sizeof(R) == sizeof(double)) {
        PM_HIP(hipMemcpy(dst[f], from[f], nmesh * sizeof(R), hipMemcpyDeviceToHost));
      } else {
        PM_HIP(hipMemcpy(stage.data(), from[f], nmesh * sizeof(R), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nmesh; ++i) dst[f][i] = (double)stage[i];
      }
    }
    return 0;
  }

  // ---- single-component solve: psi = G * w (Cin = Cout = 1) -----------
  int tz1 = 0;  // scalar z-pass columns per block (0: not tuned yet)
  std::string z1_key() const {
    char k[128];
    snprintf(k, sizeof(k), "z1:L%d:Ly%d:Hx%d:%s", L[2], L[1], Hx, sizeof(R) == 4 ? "f32" : "f64");
    return k;
  }
  int64_t row1() const {
    const int64_t align = 128 / (int64_t)sizeof(V);
    return ((int64_t)Hx + align - 1) / align * align;
  }
  ZArgs z1_args() const {
    ZArgs za{};
    const int64_t H1 = row1();
    za.B = B;
    za.Ghat = Ghat;
    za.rs = H1 * L[1];
    za.cs = 0;
    za.kxs = 1;
    za.kys = H1;
    za.Hx = Hx;
    za.Ly = L[1];
    za.Lz = L[2];
    za.Lx = L[0];
    za.Mz = M[2];
    za.Hy = Hy;
    za.Hz = Hz;
    za.allper = allper;
    for (int d = 0; d < 3; ++d) {
      za.dk[d] = 1.0 / ((double)L[d] * g.dx[d]);
      za.ksmp[d] = 1.0 / (2.0 * g.dx[d]);
    }
    za.h3 = g.dx[0] * g.dx[1] * g.dx[2];
    za.plan = pLz;
    return za;
  }

  // time the scalar z pass for each tile width on this handle's buffers
  int tune_scalar() override {
    if (tz1) return 0;
    if (const char *env = getenv("PMGPU_Z1_KERNEL")) {  // fixed choice, no timing
      const int t = atoi(env);
      PM_CHECK(t > 0, "pmgpu: bad PMGPU_Z1_KERNEL");
      PM_CHECK(fx_z1_avail<R>(L[2], t) == 1,
               "pmgpu: PMGPU_Z1_KERNEL=" + std::to_string(t) + " is not available for z length " + std::to_string(L[2]));
      tz1 = t;
      return 0;
    }
    const std::string key = z1_key();
    if (int v = tune_cache().get(key)) {
      tz1 = v;
      return 0;
    }
    PM_CHECK(prm.poisson_kernel == 1, "pmgpu: the single-component (stream function) solve needs poisson_kernel = 1");
    PM_CHECK(allper || Ghat != nullptr, "pmgpu: pmgpu_poisson_setup was not called");
    PM_HIP(hipMemsetAsync(B, 0, nB * sizeof(V), stream));
    hipEvent_t e0, e1;
    PM_HIP(hipEventCreate(&e0));
    PM_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int best_t = 0;
    const int cands[] = {1, 2, 4, 8, 16, kZ1Reg + 4, kZ1Reg + 8, kZ1Reg + 16};
    ZArgs za = z1_args();
    // two sweeps over the candidates, 5 launches each; a candidate's time is
    // its faster sweep (short launches are noisy)
    for (int sweep = 0; sweep < 2; ++sweep)
      for (int t : cands) {
        if (fx_z1<R>(L[2], t, za, L[1], d_tw, stream) != 1) continue;  // warm-up / availability
        PM_HIP(hipEventRecord(e0, stream));
        for (int r = 0; r < 5; ++r) fx_z1<R>(L[2], t, za, L[1], d_tw, stream);
        PM_HIP(hipEventRecord(e1, stream));
        PM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        PM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
          best = ms;
          best_t = t;
        }
      }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    PM_CHECK(best_t > 0, "pmgpu: no scalar z-pass kernel for this transform length");
    tz1 = best_t;
    tune_cache().put(key, best_t);
    return 0;
  }

  // Stream-ordered; no host synchronisation and no memset, so it can be
  // captured into a graph once tune_scalar has run.
  int solve_scalar_device(const void *w, void *psi) override {
    PM_CHECK(prm.poisson_kernel == 1, "pmgpu: the single-component (stream function) solve needs poisson_kernel = 1");
    PM_CHECK(allper || Ghat != nullptr, "pmgpu: pmgpu_poisson_setup was not called");
    PM_CHECK(w && psi, "pmgpu: null mesh pointer");
    PM_CHECK(tz1 > 0, "pmgpu: scalar solve not tuned (pmgpu_poisson_tune_scalar)");
    const int64_t H1 = row1();
    const int64_t off = (int64_t)lo[0] + (int64_t)E[0] * (lo[1] + (int64_t)E[1] * lo[2]);
    const int64_t ys = E[0], zs = (int64_t)E[0] * E[1];
    const R *win = reinterpret_cast<const R *>(w) + off;
    int e = 0;
    e |= launch_xfwd<R>(win, A, 0, 1, ys, zs, 0, 1, H1, H1 * M[1], 1, M[0], M[1], M[2], Tx, pNh, twLx_off, d_tw);
    e |= launch_strided<R, -1>(A, B, H1, H1, H1 * M[1], H1 * L[1], Hx, M[1], L[1], M[2], 1, 1.0, Ts, pLy, d_tw);
    if (e) return -1;
    ZArgs za = z1_args();
    PM_CHECK(fx_z1<R>(L[2], tz1, za, L[1], d_tw, stream) == 1, "pmgpu: scalar z-pass launch failed");
    e |= launch_strided<R, 1>(B, A, H1, H1, H1 * L[1], H1 * M[1], Hx, L[1], M[1], M[2], 1, 1.0 / (double)L[1], Ts,
                              pLy, d_tw);
    if (e) return -1;
    XInvArgs xa{A, reinterpret_cast<R *>(psi) + off, 0, 1, ys, zs, 0, 1, H1, H1 * M[1], 1, M[0], M[1], M[2],
                Hx, Tx, 1.0 / (double)L[0], pNh, twLx_off};
    PM_CHECK(fx_xinv<R>(Nh, xa, d_tw, stream) == 1, "pmgpu: scalar solve needs a specialised x length");
    return scalar_ghost_fill(psi);
  }

  int scalar_ghost_fill(void *psi) override {
    for (int d = 0; d < 3; ++d)
      if (g.bound_cond[d] == 1) {
        hipLaunchKernelGGL(k_ghost_fill<R>, dim3(1024), dim3(256), 0, stream, reinterpret_cast<R *>(psi), E[0],
                           E[1], E[2], d, g.ncell[d], ngs[d], 1);
        PM_HIP(hipGetLastError());
      }
    return 0;
  }

  // The passes of solve_scalar_device as group entries: the same arguments,
  // the same kernels (the tile sizes of the length-specialised launchers).
  int scalar_passes(const void *w, void *psi, ScalarPasses &sp) override {
    PM_CHECK(prm.poisson_kernel == 1, "pmgpu: the single-component (stream function) solve needs poisson_kernel = 1");
    PM_CHECK(allper || Ghat != nullptr, "pmgpu: pmgpu_poisson_setup was not called");
    PM_CHECK(w && psi, "pmgpu: null mesh pointer");
    PM_CHECK(tz1 > 0, "pmgpu: scalar solve not tuned (pmgpu_poisson_tune_scalar)");
    sp = ScalarPasses{};
    const int tx = fx_tile_x1<R>(Nh), ty = fx_tile_cols<R>(L[1]);
    const int tz = tz1 < kZ1Reg ? tz1 : tz1 - kZ1Reg;
    sp.ok = tx > 0 && ty > 0 && fx_tile_cols<R>(L[2]) > 0;
    if (!sp.ok) return 0;
    sp.fp64 = sizeof(R) == 8;
    sp.NH = Nh;
    sp.Ly = L[1];
    sp.Lz = L[2];
    sp.tz1 = tz1;
    sp.skernel = skernel;
    sp.allper = allper != 0;
    sp.y_half = fx_even_rows(L[1]) == 1 && 2 * M[1] == L[1];
    sp.z_half = !sp.allper && fx_even_rows(L[2]) == 1 && 2 * M[2] == L[2];
    for (int d = 0; d < 3; ++d) sp.ghost = sp.ghost || g.bound_cond[d] == 1;
    const int64_t H1 = row1();
    const int64_t off = (int64_t)lo[0] + (int64_t)E[0] * (lo[1] + (int64_t)E[1] * lo[2]);
    const int64_t ys = E[0], zs = (int64_t)E[0] * E[1];
    const R *win = reinterpret_cast<const R *>(w) + off;
    R *pout = reinterpret_cast<R *>(psi) + off;
    auto ceil_div = [](int64_t a, int64_t b) { return (a + b - 1) / b; };
    sp.xf.a = XFwdArgs{win, A, 0, 1, ys, zs, 0, 1, H1, H1 * M[1], 1, M[0], M[1], M[2], pNh.L + 1, tx, pNh, twLx_off};
    sp.xf.gx = ceil_div(M[1], tx);
    sp.tiles[0] = sp.xf.gx * M[2];
    sp.yf.a = StridedArgs{A, B, H1, H1, H1 * M[1], H1 * L[1], Hx, M[1], L[1], ty, 1.0, pLy, skernel};
    sp.yf.gx = ceil_div(Hx, ty);
    sp.tiles[1] = sp.yf.gx * M[2];
    sp.z.a = z1_args();
    sp.z.gx = ceil_div(Hx, tz);
    sp.tiles[2] = sp.z.gx * L[1];
    sp.yi.a = StridedArgs{B, A, H1, H1, H1 * L[1], H1 * M[1], Hx, L[1], M[1], ty, 1.0 / (double)L[1], pLy, skernel};
    sp.yi.gx = sp.yf.gx;
    sp.tiles[3] = sp.tiles[1];
    sp.xi.a = XInvArgs{A, pout, 0, 1, ys, zs, 0, 1, H1, H1 * M[1], 1, M[0], M[1], M[2], Hx, tx, 1.0 / (double)L[0], pNh,
                       twLx_off};
    sp.xi.gx = sp.xf.gx;
    sp.tiles[4] = sp.tiles[0];
    sp.xf.tw = sp.yf.tw = sp.z.tw = sp.yi.tw = sp.xi.tw = d_tw;
    return 0;
  }

  int solve_scalar_host(const double *w, double *psi) override {
    if (tune_scalar()) return -1;
    const size_t n1 = nmesh / 3;
    if (!mesh_in) PM_HIP(hipMalloc(&mesh_in, nmesh * sizeof(R)));
    if (!mesh_out) PM_HIP(hipMalloc(&mesh_out, 3 * nmesh * sizeof(R)));
    std::vector<R> stage(n1);
    for (size_t i = 0; i < n1; ++i) stage[i] = (R)w[i];
    PM_HIP(hipMemcpyAsync(mesh_in, stage.data(), n1 * sizeof(R), hipMemcpyHostToDevice, stream));
    PM_HIP(hipMemsetAsync(mesh_out, 0, n1 * sizeof(R), stream));
    if (solve_scalar_device(mesh_in, mesh_out)) return -1;
    PM_HIP(hipStreamSynchronize(stream));
    PM_HIP(hipMemcpy(stage.data(), mesh_out, n1 * sizeof(R), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n1; ++i) psi[i] = (double)stage[i];
    return 0;
  }
};

}  // namespace pmgpu

using namespace pmgpu;

struct pmgpu_poisson_s {
  std::unique_ptr<PoissonBase> impl;
};

// A set of solver handles (patches of any sizes) solved together: the
// single-component solves of all patches are captured once into a HIP graph
// and the graph is replayed while the mesh pointers stay the same.  Grouped
// form (every length specialised): per pass, the patches that share a
// kernel instance run as one group launch (a table of their arguments), the
// groups of a pass side by side on their own streams, the passes in order.
// Otherwise one branch per patch, each on its own stream.
struct BatchGroup {
  int pass;        // 0 x fwd, 1 y fwd, 2 z, 3 y inv, 4 x inv
  bool fp64;
  int len, tz1, skernel;
  bool half, allper;
  std::vector<int> members;
  size_t tab_off = 0;
  int64_t nb = 0;  // tiles (blocks) of the launch
};

struct pmgpu_poisson_batch_s {
  std::vector<pmgpu_poisson_t> h;
  hipStream_t stream = nullptr;   // the caller's stream (may be the null stream)
  hipStream_t cap = nullptr;      // the graph's own stream (the null stream cannot be captured)
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> ev;
  hipEvent_t fork = nullptr, in_done = nullptr, out_done = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<const void *> in;
  std::vector<void *> out;
  std::vector<BatchGroup> groups;
  std::vector<hipEvent_t> gev;
  void *d_tab = nullptr;  // the groups' entry tables
  ~pmgpu_poisson_batch_s() {
    if (cap) (void)hipStreamSynchronize(cap);
    if (d_tab) (void)hipFree(d_tab);
    for (auto e : gev) (void)hipEventDestroy(e);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : ev) (void)hipEventDestroy(e);
    hipEvent_t es[3] = {fork, in_done, out_done};
    for (auto e : es)
      if (e) (void)hipEventDestroy(e);
    for (auto st : side) (void)hipStreamDestroy(st);
    if (cap) (void)hipStreamDestroy(cap);
  }
};

extern "C" {

int pmgpu_poisson_create(const pmgpu_geom *g, const pmgpu_params *p, int32_t precision, void *stream,
                         pmgpu_poisson_t *out) {
  PM_CHECK(g && p && out, "pmgpu_poisson_create: null argument");
  PM_CHECK(precision == PMGPU_FP64 || precision == PMGPU_FP32, "pmgpu_poisson_create: bad precision");
  auto h = new pmgpu_poisson_s;
  int e;
  if (precision == PMGPU_FP64) {
    auto s = new PoissonGPU<double>;
    h->impl.reset(s);
    e = s->init(*g, *p, (hipStream_t)stream);
  } else {
    auto s = new PoissonGPU<float>;
    h->impl.reset(s);
    e = s->init(*g, *p, (hipStream_t)stream);
  }
  if (e) {
    delete h;
    return -1;
  }
  *out = h;
  return 0;
}

int pmgpu_poisson_destroy(pmgpu_poisson_t h) {
  delete h;
  return 0;
}

int pmgpu_poisson_setup(pmgpu_poisson_t h) {
  PM_CHECK(h, "pmgpu_poisson_setup: null handle");
  return h->impl->setup();
}

int pmgpu_poisson_set_level(pmgpu_poisson_t h, int32_t level, int32_t parent_level) {
  PM_CHECK(h, "pmgpu_poisson_set_level: null handle");
  if (level == 1) {
    h->impl->refine = 0.0;
    return 0;
  }
  PM_CHECK(level - parent_level >= 0 && level - parent_level <= 30,
           "pmgpu_poisson_set_level: the patch level must be 0 to 30 levels above its parent");
  h->impl->refine = std::ldexp(1.0, level - parent_level);
  return 0;
}

int pmgpu_poisson_solve(pmgpu_poisson_t h, const double *vort, double *vel, double *vort_out, double *vecpot,
                        int32_t reg_vort, int32_t reproj, int32_t calc_vecpot) {
  PM_CHECK(h && vort && vel, "pmgpu_poisson_solve: null argument");
  return h->impl->solve_host(vort, vel, vort_out, vecpot, reg_vort, reproj, calc_vecpot);
}

int pmgpu_poisson_solve_device(pmgpu_poisson_t h, const void *d_vort, void *d_vel, void *d_vort_out,
                               void *d_vecpot, int32_t reg_vort, int32_t reproj, int32_t calc_vecpot) {
  PM_CHECK(h, "pmgpu_poisson_solve_device: null handle");
  return h->impl->solve_device(d_vort, d_vel, d_vort_out, d_vecpot, reg_vort, reproj, calc_vecpot);
}

}  // extern "C"

namespace pmgpu {
int poisson_handle_info(pmgpu_poisson_t h, pmgpu_geom *g, int *precision, hipStream_t *stream) {
  PM_CHECK(h, "pmgpu: null solver handle");
  *g = h->impl->g;
  *precision = h->impl->precision();
  *stream = h->impl->stream;
  return 0;
}
}  // namespace pmgpu

extern "C" {

int pmgpu_poisson_kernel_octant(pmgpu_poisson_t h, void *host, int64_t cap_bytes, int32_t dims[3]) {
  PM_CHECK(h && host && dims, "pmgpu_poisson_kernel_octant: null argument");
  return h->impl->octant(host, cap_bytes, dims);
}

int pmgpu_poisson_last_timing(pmgpu_poisson_t h, double *zpass_ms, double *total_ms) {
  PM_CHECK(h, "pmgpu_poisson_last_timing: null handle");
  double a, b;
  if (h->impl->averages(&a, &b)) return -1;
  if (zpass_ms) *zpass_ms = a;
  if (total_ms) *total_ms = b;
  return 0;
}

int pmgpu_poisson_set_timing(pmgpu_poisson_t h, int32_t enable) {
  PM_CHECK(h, "pmgpu_poisson_set_timing: null handle");
  h->impl->timing = enable != 0;
  h->impl->nrec = 0;
  return 0;
}

int pmgpu_poisson_tune_scalar(pmgpu_poisson_t h) {
  PM_CHECK(h, "pmgpu_poisson_tune_scalar: null handle");
  return h->impl->tune_scalar();
}

int pmgpu_poisson_solve_scalar(pmgpu_poisson_t h, const double *w, double *psi) {
  PM_CHECK(h && w && psi, "pmgpu_poisson_solve_scalar: null argument");
  return h->impl->solve_scalar_host(w, psi);
}

int pmgpu_poisson_solve_scalar_device(pmgpu_poisson_t h, const void *d_w, void *d_psi) {
  PM_CHECK(h, "pmgpu_poisson_solve_scalar_device: null handle");
  if (h->impl->tune_scalar()) return -1;
  return h->impl->solve_scalar_device(d_w, d_psi);
}

int pmgpu_poisson_batch_create(const pmgpu_poisson_t *handles, int32_t n, void *stream,
                               pmgpu_poisson_batch_t *out) {
  PM_CHECK(handles && out && n > 0, "pmgpu_poisson_batch_create: bad argument");
  std::unique_ptr<pmgpu_poisson_batch_s> b(new pmgpu_poisson_batch_s);
  b->h.assign(handles, handles + n);
  b->stream = (hipStream_t)stream;
  for (int i = 0; i < n; ++i) {
    PM_CHECK(b->h[i], "pmgpu_poisson_batch_create: null handle");
    // each handle owns its FFT workspaces: two branches of one graph on the
    // same handle would write the same scratch concurrently
    for (int j = 0; j < i; ++j)
      PM_CHECK(b->h[j] != b->h[i], "pmgpu_poisson_batch_create: a solver handle appears twice in the batch "
                                   "(each patch needs its own handle)");
  }
  for (auto hh : b->h)
    if (hh->impl->tune_scalar()) return -1;  // tuning launches may not be captured
  // side streams of the captured graph (PMGPU_BATCH_STREAMS, default 8)
  int nside = 8;
  if (const char *env = getenv("PMGPU_BATCH_STREAMS")) nside = atoi(env) > 0 ? atoi(env) : 8;
  nside = n < nside ? n : nside;
  for (int i = 0; i < nside; ++i) {
    hipStream_t st;
    PM_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    b->side.push_back(st);
  }
  for (int i = 0; i < n; ++i) {
    hipEvent_t e;
    PM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->ev.push_back(e);
  }
  PM_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  PM_HIP(hipEventCreateWithFlags(&b->in_done, hipEventDisableTiming));
  PM_HIP(hipEventCreateWithFlags(&b->out_done, hipEventDisableTiming));
  PM_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  *out = b.release();
  return 0;
}

int pmgpu_poisson_batch_destroy(pmgpu_poisson_batch_t b) {
  delete b;
  return 0;
}

}  // extern "C"

namespace pmgpu {

// Group the patches' passes by kernel instance and upload the entry tables.
static int batch_build_groups(pmgpu_poisson_batch_s *b, const std::vector<ScalarPasses> &sp) {
  b->groups.clear();
  std::map<std::string, int> index;
  for (int pass = 0; pass < 5; ++pass)
    for (size_t i = 0; i < sp.size(); ++i) {
      const ScalarPasses &q = sp[i];
      BatchGroup G{};
      G.pass = pass;
      G.fp64 = q.fp64;
      G.len = (pass == 0 || pass == 4) ? q.NH : (pass == 2 ? q.Lz : q.Ly);
      G.tz1 = pass == 2 ? q.tz1 : 0;
      G.skernel = (pass == 1 || pass == 3) ? q.skernel : 0;
      G.half = (pass == 1 || pass == 3) ? q.y_half : (pass == 2 ? q.z_half : false);
      G.allper = pass == 2 && q.allper;
      char key[128];
      snprintf(key, sizeof(key), "%d:%d:%d:%d:%d:%d:%d", G.pass, (int)G.fp64, G.len, G.tz1, G.skernel, (int)G.half,
               (int)G.allper);
      auto it = index.find(key);
      if (it == index.end()) {
        index[key] = (int)b->groups.size();
        G.members.push_back((int)i);
        b->groups.push_back(G);
      } else {
        b->groups[it->second].members.push_back((int)i);
      }
    }
  // tables: one per group, 256-byte aligned, in one device buffer
  std::vector<unsigned char> host;
  auto put = [&](BatchGroup &G, auto entry_of) {
    using E = std::remove_reference_t<decltype(entry_of(sp[0]))>;
    G.tab_off = (host.size() + 255) / 256 * 256;
    host.resize(G.tab_off + G.members.size() * sizeof(E));
    int64_t off = 0;
    E *t = reinterpret_cast<E *>(host.data() + G.tab_off);
    for (size_t k = 0; k < G.members.size(); ++k) {
      const ScalarPasses &q = sp[G.members[k]];
      t[k] = entry_of(q);
      t[k].off = off;
      off += q.tiles[G.pass];
    }
    G.nb = off;
  };
  for (auto &G : b->groups) {
    switch (G.pass) {
      case 0: put(G, [](const ScalarPasses &q) -> GrpEntry<XFwdArgs> { return q.xf; }); break;
      case 1: put(G, [](const ScalarPasses &q) -> GrpEntry<StridedArgs> { return q.yf; }); break;
      case 2: put(G, [](const ScalarPasses &q) -> GrpEntry<ZArgs> { return q.z; }); break;
      case 3: put(G, [](const ScalarPasses &q) -> GrpEntry<StridedArgs> { return q.yi; }); break;
      default: put(G, [](const ScalarPasses &q) -> GrpEntry<XInvArgs> { return q.xi; }); break;
    }
    PM_CHECK(G.nb > 0 && G.nb < 2147483647LL, "pmgpu_poisson_batch_solve_scalar: group launch too large");
  }
  if (b->d_tab) PM_HIP(hipFree(b->d_tab));
  b->d_tab = nullptr;
  PM_HIP(hipMalloc(&b->d_tab, host.size()));
  PM_HIP(hipMemcpy(b->d_tab, host.data(), host.size(), hipMemcpyHostToDevice));
  while (b->gev.size() < b->groups.size()) {
    hipEvent_t e;
    PM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->gev.push_back(e);
  }
  return 0;
}

static int batch_launch_group(const pmgpu_poisson_batch_s *b, const BatchGroup &G, hipStream_t st) {
  const void *t = static_cast<const unsigned char *>(b->d_tab) + G.tab_off;
  const int n = (int)G.members.size();
  switch (G.pass) {
    case 0: {
      auto tt = static_cast<const GrpEntry<XFwdArgs> *>(t);
      return G.fp64 ? fx_grp_xfwd<double>(G.len, tt, n, G.nb, st) : fx_grp_xfwd<float>(G.len, tt, n, G.nb, st);
    }
    case 1:
    case 3: {
      auto tt = static_cast<const GrpEntry<StridedArgs> *>(t);
      const int S = G.pass == 1 ? -1 : 1;
      return G.fp64 ? fx_grp_strided<double>(G.len, S, G.skernel, G.half, tt, n, G.nb, st)
                    : fx_grp_strided<float>(G.len, S, G.skernel, G.half, tt, n, G.nb, st);
    }
    case 2: {
      auto tt = static_cast<const GrpEntry<ZArgs> *>(t);
      return G.fp64 ? fx_grp_z1<double>(G.len, G.tz1, G.allper, G.half, tt, n, G.nb, st)
                    : fx_grp_z1<float>(G.len, G.tz1, G.allper, G.half, tt, n, G.nb, st);
    }
    default: {
      auto tt = static_cast<const GrpEntry<XInvArgs> *>(t);
      return G.fp64 ? fx_grp_xinv<double>(G.len, tt, n, G.nb, st) : fx_grp_xinv<float>(G.len, tt, n, G.nb, st);
    }
  }
}

}  // namespace pmgpu

extern "C" {

int pmgpu_poisson_batch_solve_scalar(pmgpu_poisson_batch_t b, const void *const *d_in, void *const *d_out) {
  PM_CHECK(b && d_in && d_out, "pmgpu_poisson_batch_solve_scalar: null argument");
  const size_t n = b->h.size();
  bool same = b->exec != nullptr;
  for (size_t i = 0; same && i < n; ++i) same = b->in[i] == d_in[i] && b->out[i] == d_out[i];
  if (!same) {
    // a replay of the previous graph may still read its tables
    PM_HIP(hipStreamSynchronize(b->cap));
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->in.clear();
    b->out.clear();
    // grouped form (PMGPU_BATCH_GROUPED=1) when every patch's passes have
    // group kernels; per-patch branches otherwise
    std::vector<ScalarPasses> sp(n);
    bool grouped = false;
    if (const char *env = getenv("PMGPU_BATCH_GROUPED")) grouped = atoi(env) != 0;
    for (size_t i = 0; grouped && i < n; ++i) {
      if (b->h[i]->impl->scalar_passes(d_in[i], d_out[i], sp[i])) return -1;
      grouped = sp[i].ok;
    }
    if (grouped && batch_build_groups(b, sp)) return -1;
    // capture on the batch's own stream, forking onto the side streams
    PM_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
    int rc = hipEventRecord(b->fork, b->cap) == hipSuccess ? 0 : -1;
    if (grouped) {
      size_t k = 0;
      for (int pass = 0; rc == 0 && pass < 5; ++pass) {
        int gi = 0;
        for (const BatchGroup &G : b->groups) {
          if (G.pass != pass || rc != 0) continue;
          hipStream_t st = b->side[gi++ % b->side.size()];
          if (hipStreamWaitEvent(st, b->fork, 0) != hipSuccess) rc = -1;
          if (rc == 0 && batch_launch_group(b, G, st) != 1) rc = -1;
          if (rc == 0 && hipEventRecord(b->gev[k], st) != hipSuccess) rc = -1;
          if (rc == 0 && hipStreamWaitEvent(b->cap, b->gev[k], 0) != hipSuccess) rc = -1;
          ++k;
        }
        if (rc == 0 && hipEventRecord(b->fork, b->cap) != hipSuccess) rc = -1;  // the next pass starts here
      }
      for (size_t i = 0; rc == 0 && i < n; ++i) {
        if (!sp[i].ghost) continue;
        PoissonBase *p = b->h[i]->impl.get();
        hipStream_t own = p->stream;
        p->stream = b->cap;
        if (p->scalar_ghost_fill(d_out[i])) rc = -1;
        p->stream = own;
      }
    } else {
      // largest patches first, each onto the side stream with the least
      // work so far (mesh cells), so that the branches end together
      std::vector<double> cost(n), load(b->side.size(), 0.0);
      std::vector<size_t> order(n);
      for (size_t i = 0; i < n; ++i) {
        const pmgpu_geom &gg = b->h[i]->impl->g;
        cost[i] = 1.0;
        for (int d = 0; d < 3; ++d) cost[i] *= (double)(gg.ncell[d] + 2 * gg.nghost);
        order[i] = i;
      }
      std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
      for (size_t oi = 0; rc == 0 && oi < n; ++oi) {
        const size_t i = order[oi];
        const size_t j = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        load[j] += cost[i];
        hipStream_t st = b->side[j];
        PoissonBase *p = b->h[i]->impl.get();
        hipStream_t own = p->stream;
        if (hipStreamWaitEvent(st, b->fork, 0) != hipSuccess) rc = -1;
        p->stream = st;
        if (rc == 0 && p->solve_scalar_device(d_in[i], d_out[i])) rc = -1;
        p->stream = own;
        if (rc == 0 && hipEventRecord(b->ev[i], st) != hipSuccess) rc = -1;
        if (rc == 0 && hipStreamWaitEvent(b->cap, b->ev[i], 0) != hipSuccess) rc = -1;
      }
    }
    hipGraph_t graph = nullptr;
    hipError_t ec = hipStreamEndCapture(b->cap, &graph);
    (void)hipGetLastError();  // a failed capture leaves its error behind
    if (rc != 0 || ec != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      PM_CHECK(false, "pmgpu_poisson_batch_solve_scalar: graph capture failed");
    }
    b->graph = graph;
    PM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->in.assign(d_in, d_in + n);
    b->out.assign(d_out, d_out + n);
  }
  // order the replay after the caller's earlier work and before its later work
  PM_HIP(hipEventRecord(b->in_done, b->stream));
  PM_HIP(hipStreamWaitEvent(b->cap, b->in_done, 0));
  PM_HIP(hipGraphLaunch(b->exec, b->cap));
  PM_HIP(hipEventRecord(b->out_done, b->cap));
  PM_HIP(hipStreamWaitEvent(b->stream, b->out_done, 0));
  return 0;
}

}  // extern "C"
