// C ABI of the gfx950 Fourier-Galerkin operator: plans, passes, CG driver, energies.
// Declarations and reference citations: include/ffthom.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ffthom.h"
#include "fh_kernels.hpp"
#include "fh_launch.hpp"
#include "fh_batch.hpp"
#include "fh_forms.hpp"
#include "fh_ga.hpp"
#include "fh_iter.hpp"
#include "fh_micro.hpp"
#include "fh_pot.hpp"

using fh::AxisPlan;
using fh::CGState;
using fh::cpx;
using fh::GParams;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(FH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define RETCHK(call)        \
  do {                      \
    int r_ = (call);        \
    if (r_ != FH_OK) return r_; \
  } while (0)

constexpr size_t kLdsTarget = 48 * 1024;
// threads of the one-workgroup CG reductions (alpha, beta): each reads ~np/1024 partial sums
// (a few thousand at 243^3), so the cross-XCD load latency is paid a few times, not ~30
constexpr int kFinThreads = 1024;
constexpr size_t kLdsMax = 160 * 1024;
// doubles of a plan's small-result block (after its partial sums)
constexpr size_t kAuxDoubles = 64;

// Scratch of fh_energy_matrix_noalloc: the per-workgroup double-double partial sums of up to
// six fields (2 * 36 * 1024 doubles) and the 36 results. One copy per GPU (576.5 KiB), held until
// the process ends. Host code never uses this name as an address: ensure_device() takes the
// device address of the current GPU's copy with hipGetSymbolAddress.
constexpr size_t kNoAllocDoubles = 2 * 36 * 1024 + 64;
__device__ __align__(256) double scratch[kNoAllocDoubles];
constexpr int kMaxDevices = 64;
double* g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device
std::mutex g_noalloc_mu;                  // one fh_energy_matrix_noalloc at a time in this process

std::mutex g_init_mu;
int g_init_state = -1;  // -1 unknown, FH_OK, or an error code
std::string g_init_msg;
std::string g_dev_name;

int ensure_device() {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_init_state == FH_OK) return FH_OK;
  if (g_init_state > 0) return fail(g_init_state, g_init_msg);
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n <= 0) {
    g_init_state = FH_ERR_NODEV;
    g_init_msg = std::string("no HIP device: ") + (e != hipSuccess ? hipGetErrorString(e) : "count is 0");
    return fail(g_init_state, g_init_msg);
  }
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, 0);
  if (e != hipSuccess) {
    g_init_state = FH_ERR_HIP;
    g_init_msg = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
    return fail(g_init_state, g_init_msg);
  }
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    g_init_state = FH_ERR_UNSUPPORTED;
    g_init_msg = std::string("device arch ") + prop.gcnArchName + " is not gfx950";
    return fail(g_init_state, g_init_msg);
  }
  e = hipSetDevice(0);
  if (e != hipSuccess) {
    g_init_state = FH_ERR_HIP;
    g_init_msg = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return fail(g_init_state, g_init_msg);
  }
  g_dev_name = std::string(prop.name[0] ? prop.name : "AMD GPU") + " (" + prop.gcnArchName + ")";
  // the device's own copy of the no-alloc scratch (device 0 is current: set just above)
  void* sp = nullptr;
  e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !sp) {
    g_init_state = FH_ERR_HIP;
    g_init_msg = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return fail(g_init_state, g_init_msg);
  }
  g_scratch_dev[0] = static_cast<double*>(sp);
  g_init_state = FH_OK;
  return FH_OK;
}

// Radix schedule: 4s first, then 2, 3, 5, 7, 11, 13. False if L has a larger prime factor.
bool factorize(int L, AxisPlan* ap) {
  std::memset(ap, 0, sizeof(*ap));
  ap->L = L;
  ap->M = L;
  ap->blue = 0;
  int rem = L, st = 0, pre = 1;
  const int rads[] = {4, 2, 3, 5, 7, 11, 13};
  for (int R : rads) {
    while (rem % R == 0) {
      if (st >= fh::kMaxStages) return false;
      ap->rad[st] = R;
      ap->pre[st] = pre;
      pre *= R;
      rem /= R;
      ++st;
    }
  }
  ap->nst = st;
  return rem == 1;
}

bool smooth7(int n) {
  for (int p : {2, 3, 5, 7})
    while (n % p == 0) n /= p;
  return n == 1;
}

// Schedule of an axis of any length: the radix stages of L, or for a length with a prime factor
// above 13 Bluestein's convolution on the smallest 7-smooth M >= 2L - 1 (fh_fft.hpp, fft_lds).
// Bluestein lengths are capped at kMaxBluestein: a pencil of a longer convolution no longer fits
// one workgroup's LDS in double precision (2 or more pencils of 16-byte values in 160 KiB).
constexpr int kMaxBluestein = 4096;

bool plan_axis(int L, AxisPlan* ap) {
  if (factorize(L, ap)) return true;
  int M = 2 * L - 1;
  while (!smooth7(M)) ++M;
  if (M > kMaxBluestein) return false;
  if (!factorize(M, ap)) return false;
  ap->L = L;
  ap->M = M;
  ap->blue = 1;
  return true;
}

// Twiddle table of an axis (see fft_lds): W_M^m (m < M), and with Bluestein the chirp
// w_m = exp(-i pi m^2 / L) (m < L) and the spectra FFT_M(conj w)/M, FFT_M(w)/M of the periodically
// extended convolution kernel; complex values as (re, im) pairs, computed in long double.
std::vector<long double> axis_table(const AxisPlan& ap) {
  const long double pi = 3.14159265358979323846264338327950288L;
  const int L = ap.L, M = ap.M;
  std::vector<long double> t(2 * (size_t)(ap.blue ? 3 * M + L : M));
  for (int m = 0; m < M; ++m) {
    const long double ang = -2.0L * pi * (long double)m / (long double)M;
    t[2 * m] = cosl(ang);
    t[2 * m + 1] = sinl(ang);
  }
  if (!ap.blue) return t;
  std::vector<long double> wr(L), wi(L);
  for (int m = 0; m < L; ++m) {
    const long long r = ((long long)m * m) % (2LL * L);  // w_m depends on m^2 mod 2L only
    const long double ang = -pi * (long double)r / (long double)L;
    wr[m] = cosl(ang);
    wi[m] = sinl(ang);
    t[2 * (M + m)] = wr[m];
    t[2 * (M + m) + 1] = wi[m];
  }
  // kernel b_j = conj(w_|j|) for |j| < L, placed circularly on M points
  std::vector<long double> br(M, 0.0L), bi(M, 0.0L);
  for (int m = 0; m < L; ++m) {
    br[m] = wr[m];
    bi[m] = -wi[m];
    if (m > 0) {
      br[M - m] = wr[m];
      bi[M - m] = -wi[m];
    }
  }
  const size_t o1 = 2 * (size_t)(M + L), o2 = o1 + 2 * (size_t)M;
  for (int k = 0; k < M; ++k) {
    long double fr = 0.0L, fi = 0.0L, gr = 0.0L, gi = 0.0L;  // FFT_M(b), FFT_M(conj b)
    for (int j = 0; j < M; ++j) {
      const long long r = ((long long)j * k) % M;
      const long double c = t[2 * r], sn = t[2 * r + 1];  // W_M^{jk}
      fr += br[j] * c - bi[j] * sn;
      fi += br[j] * sn + bi[j] * c;
      gr += br[j] * c + bi[j] * sn;
      gi += br[j] * sn - bi[j] * c;
    }
    t[o1 + 2 * k] = fr / M;
    t[o1 + 2 * k + 1] = fi / M;
    t[o2 + 2 * k] = gr / M;
    t[o2 + 2 * k + 1] = gi / M;
  }
  return t;
}

int fft_threads(const AxisPlan& ap, int npen) {
  long long need = 64;
  for (int s = 0; s < ap.nst; ++s) {
    const int R = ap.rad[s];
    const long long bf = (long long)npen * (ap.M / R);
    const long long mb = fh::max_butterflies(R);
    need = std::max(need, (bf + mb - 1) / mb);
  }
  need = ((need + 63) / 64) * 64;
  return (int)need;
}

struct RowCfg {
  int rows_pb = 2, nthr = 64, nblocks = 0;
  size_t lds = 0;
};

struct ColCfg {
  int P = 1, nthr = 64, ntiles = 0;
  size_t lds = 0;
};

}  // namespace

struct fh_plan {
  int dim = 0, physics = 0, primaldual = 0, ncomp = 0;
  int Nu[3] = {1, 1, 1};
  double Y[3] = {1, 1, 1};
  int n0 = 1, n1 = 1, n2 = 1, n2h = 1;
  long long n = 0, nh = 0;
  AxisPlan ax[3];
  void* tw64[3] = {nullptr, nullptr, nullptr};
  void* tw32[3] = {nullptr, nullptr, nullptr};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int coef_kind = 0;
  double* coef64 = nullptr;
  float* coef32 = nullptr;
  void* spec = nullptr;          // ncomp*nh complex doubles (fp32 passes reuse it)
  double* partial = nullptr;     // per-workgroup partial sums
  size_t partial_cap = 0;
  double* aux = nullptr;         // kAuxDoubles after the partial sums: small results (energies, sums)
  CGState* st = nullptr;
  CGState* st_add = nullptr;  // alpha = -1: the inverse pass's epilogue 2 adds (x += G c)
  double* hist = nullptr;
  int hist_cap = 0;
  void* work = nullptr;          // CG vectors
  size_t work_bytes = 0;
  double last_matvec_ms = 0, last_solve_ms = 0;
  int last_updates = 0;
  int generic = 0;               // 1: never use the fixed-length kernels
  bool fused_cg = true;          // mixed CG: fused iteration where available
  bool defer_ok = true;          // fused iteration: alpha/beta inside the passes on small grids
  bool cg_defer = false;         // this solve runs the deferred alpha/beta
  int defer_nb2 = 0;             // workgroups of the fused inverse row pass (its partials)
  hipEvent_t prof_t0 = nullptr, prof_t1 = nullptr;  // matvec profile: events of the next launch
  int banded = 0;                // 1: projection restricted to |k_a| <= band[a] (view axes)
  int band[3] = {0, 0, 0};
  double op_scale = 1.0;         // factor of the operator G A (not of the bare projection)
  // fused mixed CG: a batch of iterations captured once as a HIP graph, replayed per batch
  hipGraphExec_t it_exec[9] = {};  // captured batches of 1, 2, 4, 8 fused iterations, by length
  const void* it_key[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  bool it_graph_off = false;     // capture failed on this plan: direct launches from then on
  bool use_graphs = true;        // fh_plan_use_graphs
  int last_graph_batches = 0;    // batches of the last solve replayed from the graph
  int proj_var[2] = {0, 0};      // schedule variants of the fixed-length passes: [fp64, fp32]
  int cols_var[2] = {0, 0};
  int rows_var[2] = {0, 0};
  int res_every = 0;             // fp64 CG: log ||b - A x|| every res_every iterations (0: off)
  fh_iterate_fn it_fn = nullptr; // host callback on the iterate every it_every iterations (0: off)
  void* it_user = nullptr;
  int it_every = 0;
  std::vector<int> res_kit;      // the log of the last solve: iteration counts and norms
  std::vector<double> res_norm;
};

struct fh_batch {
  int S = 0, ntasks = 0, nslots = 0;
  std::vector<fh::batch::Sample> smp;
  long long nvox = 0;              // voxels over all samples
  size_t lds = 0;
  fh::batch::Args args{};
  fh::batch::Sample* d_smp = nullptr;
  int* d_order = nullptr;
  int* d_counter = nullptr;
  double* d_a64 = nullptr;
  float* d_a32 = nullptr;
  cpx<double>* d_tw64 = nullptr;
  cpx<float>* d_tw32 = nullptr;
  double* d_ws64 = nullptr;
  float* d_ws32 = nullptr;
  double* d_energy = nullptr;
  int* d_kit = nullptr;
  int* d_nupd = nullptr;
  double* d_normres = nullptr;
  int* d_status = nullptr;
  fh::batch::Params* d_par = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool coef_set = false;
  double last_ms = 0.0;
  int variant = -1;                // radix schedule variant of the solve kernel (-1: autotune)
  int order_mode = 0;              // 0 largest grids first, 1 sample order
  std::vector<int> order_size, order_sample;
  double tune_ms[2 * fh::batch::NVAR] = {0};
};

// Ga post-processing: exactly integrated coefficient and interpolated solutions on a
// quadrature grid (see fh_ga.hpp for why any grid of at least 2N-1 points per axis is exact
// for odd N).
struct fh_ga {
  fh_plan* coarse = nullptr;  // grid of the solutions (not owned)
  fh_plan* fine = nullptr;    // quadrature grid (owned): FFT passes, partial sums
  fh::ga::SpecGrid g{};
  int coef_kind = 0, ncoef = 0;
  double* coef = nullptr;     // ncoef * fine->n, origin-0 voxel order
  double* fields = nullptr;   // interpolated solutions
  size_t fields_bytes = 0;
};

namespace {

GParams make_gparams(const fh_plan* pl, double scale, int ncomp_override = 0) {
  GParams g;
  g.dim = pl->dim;
  g.physics = pl->physics;
  g.dual = pl->primaldual;
  g.ncomp = ncomp_override ? ncomp_override : pl->ncomp;
  g.n0 = pl->n0;
  g.n1 = pl->n1;
  g.n2 = pl->n2;
  if (pl->dim == 3) {
    g.iY[0] = 1.0 / pl->Y[0];
    g.iY[1] = 1.0 / pl->Y[1];
    g.iY[2] = 1.0 / pl->Y[2];
  } else {
    g.iY[0] = 0.0;
    g.iY[1] = 1.0 / pl->Y[0];
    g.iY[2] = 1.0 / pl->Y[1];
  }
  g.scale = scale;
  g.banded = pl->banded;
  for (int a = 0; a < 3; ++a) g.band[a] = pl->band[a];
  return g;
}

template <typename T>
const cpx<T>* twiddles(const fh_plan* pl, int axis) {
  return reinterpret_cast<const cpx<T>*>(sizeof(T) == 8 ? pl->tw64[axis] : pl->tw32[axis]);
}

template <typename T>
fh::LaunchCtx make_ctx(const fh_plan* pl) {
  fh::LaunchCtx c;
  c.stream = pl->stream;
  c.n0 = pl->n0;
  c.n1 = pl->n1;
  c.n2 = pl->n2;
  c.n2h = pl->n2h;
  c.n = pl->n;
  c.nh = pl->nh;
  for (int a = 0; a < 3; ++a) c.tw[a] = sizeof(T) == 8 ? pl->tw64[a] : pl->tw32[a];
  c.t0 = pl->prof_t0;
  c.t1 = pl->prof_t1;
  c.proj_var = pl->proj_var[sizeof(T) == 8 ? 0 : 1];
  c.cols_var = pl->cols_var[sizeof(T) == 8 ? 0 : 1];
  c.rows_var = pl->rows_var[sizeof(T) == 8 ? 0 : 1];
  return c;
}

int fixed_result(int r, const char* what) {
  if (r < 0) return fail(FH_ERR_HIP, std::string(what) + ": " + hipGetErrorString((hipError_t)(-r)));
  return FH_OK;
}

// Launch of a runtime-radix kernel: LDS tiles above 64 KiB (long Bluestein pencils) need the
// kernel's dynamic-LDS limit raised first.
template <typename... KArgs, typename... Args>
void launch_gen(void (*kernel)(KArgs...), dim3 grid, dim3 block, size_t lds, hipStream_t stream, Args... args) {
  if (lds > 64 * 1024)
    (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(kernel, grid, block, lds, stream, args...);
}

int row_config(const AxisPlan& ap, int nc, size_t csz, long long nrows, RowCfg* c) {
  const int L = ap.L, W = ap.M;  // W: pencil length in LDS (Bluestein: the convolution length)
  int best = 2;
  for (int rp = 2; rp <= 32; rp += 2) {
    const size_t lds = (size_t)(rp / 2) * nc * W * csz;
    const int thr = fft_threads(ap, (rp / 2) * nc);
    if (rp > 2 && (lds > kLdsTarget || thr > 1024)) break;
    best = rp;
    if (rp >= nrows) break;
  }
  c->rows_pb = best;
  c->lds = (size_t)(best / 2) * nc * W * csz;
  c->nthr = std::max(fft_threads(ap, (best / 2) * nc), 128);
  c->nthr = std::min(c->nthr, 1024);
  if (fft_threads(ap, (best / 2) * nc) > 1024 || c->lds > kLdsMax)
    return fail(FH_ERR_UNSUPPORTED, "grid row of length " + std::to_string(L) + " does not fit one workgroup");
  c->nblocks = (int)((nrows + best - 1) / best);
  return FH_OK;
}

int col_config(const AxisPlan& ap, int unit, size_t csz, long long inner, ColCfg* c) {
  const int L = ap.L, W = ap.M;  // W: pencil length in LDS (Bluestein: the convolution length)
  long long P = (long long)(kLdsTarget / ((size_t)W * csz)) / unit * unit;
  if (P < unit) P = unit;
  const long long cap = (inner + unit - 1) / unit * unit;
  if (P > cap) P = cap;
  while (P > unit && fft_threads(ap, (int)P) > 1024) P -= unit;
  c->P = (int)P;
  c->lds = (size_t)W * P * csz;
  const int need = fft_threads(ap, (int)P);
  if (need > 1024 || c->lds > kLdsMax)
    return fail(FH_ERR_UNSUPPORTED, "grid axis of length " + std::to_string(L) + " does not fit one workgroup");
  c->nthr = std::max(need, 128);
  c->ntiles = (int)((inner + P - 1) / P);
  return FH_OK;
}

// rvec != null: `in` is the CG direction p, updated in place to r + beta p first.
template <typename T>
int launch_rows_fwd(fh_plan* pl, int nc, const T* in, const T* coef, int kind, cpx<T>* out,
                    const int* halt, const T* rvec = nullptr) {
  if (!pl->generic) {
    const int r = fh::fixed_rows_fwd<T>(make_ctx<T>(pl), nc, in, coef, kind, out, halt, rvec, pl->st, nullptr,
                                        nullptr, 0, nullptr);
    if (r != 0) return fixed_result(r, "rows forward (fixed)");
  }
  RowCfg c;
  const long long nrows = (long long)pl->n0 * pl->n1;
  RETCHK(row_config(pl->ax[2], nc, sizeof(cpx<T>), nrows, &c));
  dim3 grid(c.nblocks), block(c.nthr);
  const cpx<T>* tw = twiddles<T>(pl, 2);
  T* io = const_cast<T*>(in);
#define FH_GEN_RF(KIND)                                                                                        \
  do {                                                                                                         \
    if (rvec)                                                                                                  \
      launch_gen(fh::k_rows_fwd<T, KIND, 1>, grid, block, c.lds, pl->stream, io, coef, out, nrows, nc, \
                         pl->n, c.rows_pb, pl->ax[2], tw, halt, rvec, pl->st);                                 \
    else                                                                                                       \
      launch_gen(fh::k_rows_fwd<T, KIND, 0>, grid, block, c.lds, pl->stream, io, coef, out, nrows, nc, \
                         pl->n, c.rows_pb, pl->ax[2], tw, halt, rvec, pl->st);                                 \
  } while (0)
  switch (kind) {
    case fh::COEF_NONE: FH_GEN_RF(fh::COEF_NONE); break;
    case fh::COEF_ISO: FH_GEN_RF(fh::COEF_ISO); break;
    case fh::COEF_LAME: FH_GEN_RF(fh::COEF_LAME); break;
    case fh::COEF_FULL: FH_GEN_RF(fh::COEF_FULL); break;
    default:
      return fail(FH_ERR_ARG, "bad coefficient kind");
  }
#undef FH_GEN_RF
  HIPCHK(hipGetLastError());
  return FH_OK;
}

// Returns the number of workgroups through *nblocks (partial sums written when w != null).
template <typename T>
int launch_rows_inv(fh_plan* pl, int nc, const cpx<T>* in, T* out, T scale, const T* w,
                    double* partial, const int* halt, int* nblocks) {
  if (!pl->generic) {
    const int r = fh::fixed_rows_inv<T>(make_ctx<T>(pl), nc, in, out, scale, w ? 1 : 0, w, partial, pl->partial_cap, halt,
                                        nblocks, pl->st);
    if (r != 0) return fixed_result(r, "rows inverse (fixed)");
  }
  RowCfg c;
  const long long nrows = (long long)pl->n0 * pl->n1;
  RETCHK(row_config(pl->ax[2], nc, sizeof(cpx<T>), nrows, &c));
  if (w && (size_t)c.nblocks > pl->partial_cap) return fail(FH_ERR_STATE, "partial buffer too small");
  dim3 grid(c.nblocks), block(c.nthr);
  const cpx<T>* tw = twiddles<T>(pl, 2);
  if (w)
    launch_gen(fh::k_rows_inv<T, 1>, grid, block, c.lds, pl->stream, in, out, nrows, nc, pl->n,
                       c.rows_pb, pl->ax[2], tw, scale, w, partial, halt, pl->st);
  else
    launch_gen(fh::k_rows_inv<T, 0>, grid, block, c.lds, pl->stream, in, out, nrows, nc, pl->n,
                       c.rows_pb, pl->ax[2], tw, scale, w, partial, halt, pl->st);
  HIPCHK(hipGetLastError());
  if (nblocks) *nblocks = c.nblocks;
  return FH_OK;
}

template <typename T>
int launch_cols(fh_plan* pl, int mode, int axis, int nc, cpx<T>* data, const GParams& g,
                const int* halt) {
  if (!pl->generic) {
    const int r = mode == 2 ? fh::fixed_cols_g<T>(make_ctx<T>(pl), axis, nc, data, g, halt)
                            : fh::fixed_cols<T>(make_ctx<T>(pl), mode == 1 ? +1 : -1, axis, nc, data, halt);
    if (r != 0) return fixed_result(r, "strided pass (fixed)");
  }
  const AxisPlan& ap = pl->ax[axis];
  long long outer, ostride, inner;
  if (axis == 1) {
    outer = pl->n0;
    inner = (long long)pl->n2h * nc;
    ostride = (long long)pl->n1 * inner;
  } else {
    outer = 1;
    inner = (long long)pl->n1 * pl->n2h * nc;
    ostride = 0;
  }
  ColCfg c;
  RETCHK(col_config(ap, mode == 2 ? nc : 4, sizeof(cpx<T>), inner, &c));
  const long long nb = outer * c.ntiles;
  if (nb > 0x7fffffffLL) return fail(FH_ERR_UNSUPPORTED, "grid too large");
  dim3 grid((unsigned)nb), block(c.nthr);
  const cpx<T>* tw = twiddles<T>(pl, axis);
  if (mode == 0)
    launch_gen(fh::k_cols<T, 0>, grid, block, c.lds, pl->stream, data, ostride, inner, c.P, c.ntiles,
                       ap, tw, g, pl->n2h, halt);
  else if (mode == 1)
    launch_gen(fh::k_cols<T, 1>, grid, block, c.lds, pl->stream, data, ostride, inner, c.P, c.ntiles,
                       ap, tw, g, pl->n2h, halt);
  else
    launch_gen(fh::k_cols<T, 2>, grid, block, c.lds, pl->stream, data, ostride, inner, c.P, c.ntiles,
                       ap, tw, g, pl->n2h, halt);
  HIPCHK(hipGetLastError());
  return FH_OK;
}

// y = F^-1 G F (A x); kind COEF_NONE gives the bare projection. With w != null the
// last pass also writes per-workgroup partial sums of y.w into pl->partial.
// rvec != null: x is the CG direction, updated in place to rvec + beta x before use.
// passes 1 .. 4 of y = F^-1 G F (A x): the plan's spectrum then holds G F (A x) along all but the
// contiguous axis
template <typename T>
int run_operator_front(fh_plan* pl, const T* x, int kind, const T* coef, const int* halt, const T* rvec) {
  const int nc = pl->ncomp;
  cpx<T>* spec = reinterpret_cast<cpx<T>*>(pl->spec);
  const GParams g = make_gparams(pl, (kind != fh::COEF_NONE ? pl->op_scale : 1.0) / (double)pl->n);
  RETCHK(launch_rows_fwd<T>(pl, nc, x, coef, kind, spec, halt, rvec));
  if (pl->n0 > 1) {
    RETCHK(launch_cols<T>(pl, 0, 1, nc, spec, g, halt));
    RETCHK(launch_cols<T>(pl, 2, 0, nc, spec, g, halt));
    RETCHK(launch_cols<T>(pl, 1, 1, nc, spec, g, halt));
  } else {
    RETCHK(launch_cols<T>(pl, 2, 1, nc, spec, g, halt));
  }
  return FH_OK;
}

template <typename T>
int run_operator(fh_plan* pl, const T* x, T* y, int kind, const T* coef, const T* w, const int* halt,
                 int* nblocks_out, const T* rvec = nullptr) {
  RETCHK(run_operator_front<T>(pl, x, kind, coef, halt, rvec));
  cpx<T>* spec = reinterpret_cast<cpx<T>*>(pl->spec);
  RETCHK(launch_rows_inv<T>(pl, pl->ncomp, spec, y, T(1), w, pl->partial, halt, nblocks_out));
  return FH_OK;
}

// r = b - A x with the fp64 operator and a single-precision r. Where the fixed-length inverse
// row pass covers the plan it forms the residual in its epilogue (A x never goes to HBM; *fused
// = true, the partial sums of <r, r> come from its *np workgroups); otherwise Ax = A x is stored
// for the separate residual kernel (*fused = false).
int run_operator_residual(fh_plan* pl, const double* x, const double* b, double* Ax, float* r, bool* fused,
                          int* np) {
  RETCHK(run_operator_front<double>(pl, x, pl->coef_kind, reinterpret_cast<const double*>(pl->coef64), nullptr,
                                    nullptr));
  cpx<double>* spec = reinterpret_cast<cpx<double>*>(pl->spec);
  *fused = false;
  if (!pl->generic) {
    fh::LaunchCtx c = make_ctx<double>(pl);
    c.r32 = r;
    int nb = 0;
    const int rr = fh::fixed_rows_inv<double>(c, pl->ncomp, spec, nullptr, 1.0, 3, b, pl->partial, pl->partial_cap,
                                              nullptr, &nb, pl->st);
    if (rr != 0) {
      RETCHK(fixed_result(rr, "rows inverse with residual (fixed)"));
      *fused = true;
      *np = nb;
      return FH_OK;
    }
  }
  RETCHK(launch_rows_inv<double>(pl, pl->ncomp, spec, Ax, 1.0, nullptr, pl->partial, nullptr, nullptr));
  return FH_OK;
}

template <typename T>
const T* plan_coef(const fh_plan* pl) {
  if constexpr (sizeof(T) == 8)
    return reinterpret_cast<const T*>(pl->coef64);
  else
    return reinterpret_cast<const T*>(pl->coef32);
}

int grid_blocks(long long n, int per = 256, int cap = 2048) {
  long long b = (n + per - 1) / per;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (int)b;
}

// x += G c with the fp64 projection (c is overwritten where the fixed-length kernels do not cover
// the plan). The fixed-length inverse row pass adds G c into x in its epilogue: epilogue 2,
// out -= alpha y, with the constant state alpha = -1 (its partial sums land in pl->partial, which
// the caller's next residual overwrites); otherwise G c goes to c and a separate pass adds it.
static int run_project_add(fh_plan* pl, double* c, double* x) {
  RETCHK(run_operator_front<double>(pl, c, fh::COEF_NONE, nullptr, nullptr, nullptr));
  cpx<double>* spec = reinterpret_cast<cpx<double>*>(pl->spec);
  if (!pl->generic) {
    fh::LaunchCtx cx = make_ctx<double>(pl);
    cx.defer_beta = 0;
    int nb = 0;
    const int rr = fh::fixed_rows_inv<double>(cx, pl->ncomp, spec, x, 1.0, 2, nullptr, pl->partial, pl->partial_cap,
                                              nullptr, &nb, pl->st_add);
    if (rr != 0) return fixed_result(rr, "rows inverse adding the projection (fixed)");
  }
  RETCHK(launch_rows_inv<double>(pl, pl->ncomp, spec, c, 1.0, nullptr, pl->partial, nullptr, nullptr));
  const long long nv = (long long)pl->ncomp * pl->n;
  hipLaunchKernelGGL(fh::k_axpby<double>, dim3(grid_blocks(nv)), dim3(256), 0, pl->stream, nv, 1.0, (const double*)c,
                     1.0, x);
  HIPCHK(hipGetLastError());
  return FH_OK;
}

int check_plan(const fh_plan* pl) {
  if (!pl) return fail(FH_ERR_ARG, "null plan");
  return FH_OK;
}

int check_dtype(int dtype) {
  if (dtype != FH_F64 && dtype != FH_F32) return fail(FH_ERR_ARG, "dtype must be FH_F64 or FH_F32");
  return FH_OK;
}

// Partial-sum workgroups of the energy kernels for n voxels (at most 1024).
int energy_blocks(long long n) { return (int)std::min<long long>(1024, std::max<long long>(1, (n + 256 * 8 - 1) / (256 * 8))); }

// Enqueue AH[i][j] = <A e_i, e_j> (nf fields of nf components, n voxels, coefficient `coef` of
// layout `kind`) on `stream`: per-workgroup double-double partial sums into partial[0, 2 nf^2 nb)
// and the fixed-order final sum into AHdev[0, nf^2). Nothing is allocated.
int enqueue_energy(hipStream_t stream, const double* coef, int kind, const void* const* fields, int nf, long long n,
                   double* partial, double* AHdev) {
  if (nf < 1 || nf > 6) return fail(FH_ERR_ARG, "unsupported number of components");
  const int nn = nf * nf;
  const int nb = energy_blocks(n);
  fh::FieldPtrs F{};
  for (int i = 0; i < nf; ++i) F.p[i] = (const double*)fields[i];
  const dim3 grid(nb), block(256);
#define FH_ENERGY(NC)                                                                                \
  do {                                                                                               \
    if (kind == FH_COEF_ISO)                                                                         \
      hipLaunchKernelGGL((fh::k_energy<NC, fh::COEF_ISO>), grid, block, 0, stream, F, coef, n, partial);  \
    else if (kind == FH_COEF_LAME)                                                                   \
      hipLaunchKernelGGL((fh::k_energy<NC, fh::COEF_LAME>), grid, block, 0, stream, F, coef, n, partial); \
    else                                                                                             \
      hipLaunchKernelGGL((fh::k_energy<NC, fh::COEF_FULL>), grid, block, 0, stream, F, coef, n, partial); \
  } while (0)
  if (nf == 2)
    FH_ENERGY(2);
  else if (nf == 3)
    FH_ENERGY(3);
  else if (nf == 6)
    FH_ENERGY(6);
  else
    return fail(FH_ERR_ARG, "unsupported number of components");
#undef FH_ENERGY
  hipLaunchKernelGGL(fh::k_energy_final, dim3(1), dim3(256), 0, stream, partial, nb, nn, 1.0 / (double)n, AHdev);
  HIPCHK(hipGetLastError());
  return FH_OK;
}

// AH[i][j] = <A e_i, e_j> on the grid of pl with coefficient `coef` of layout `kind`; the plan's
// partial-sum buffer and its small result block hold the device side (no allocation).
int energy_impl(fh_plan* pl, const double* coef, int kind, const void* const* fields, int nf, double* AH) {
  const int nn = nf * nf;
  if ((size_t)2 * nn * energy_blocks(pl->n) > pl->partial_cap) return fail(FH_ERR_STATE, "partial buffer too small");
  RETCHK(enqueue_energy(pl->stream, coef, kind, fields, nf, pl->n, pl->partial, pl->aux));
  HIPCHK(hipMemcpyAsync(AH, pl->aux, nn * sizeof(double), hipMemcpyDeviceToHost, pl->stream));
  HIPCHK(hipStreamSynchronize(pl->stream));
  return FH_OK;
}

std::mutex g_scratch_mu;
double* g_scratch = nullptr;
constexpr int kScratchParts = 2048;

}  // namespace

namespace {

static int matvec_impl(fh_plan* pl, int dtype, const void* x, void* y, bool with_coef) {
  RETCHK(check_plan(pl));
  RETCHK(check_dtype(dtype));
  if (!x || !y) return fail(FH_ERR_ARG, "null field pointer");
  if (with_coef && !pl->coef_kind) return fail(FH_ERR_STATE, "fh_set_coeff has not been called");
  HIPCHK(hipEventRecord(pl->ev0, pl->stream));
  if (dtype == FH_F64) {
    RETCHK(run_operator<double>(pl, (const double*)x, (double*)y, with_coef ? pl->coef_kind : 0,
                                with_coef ? pl->coef64 : nullptr, nullptr, nullptr, nullptr));
  } else {
    RETCHK(run_operator<float>(pl, (const float*)x, (float*)y, with_coef ? pl->coef_kind : 0,
                               with_coef ? pl->coef32 : nullptr, nullptr, nullptr, nullptr));
  }
  HIPCHK(hipEventRecord(pl->ev1, pl->stream));
  HIPCHK(hipStreamSynchronize(pl->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
  pl->last_matvec_ms = ms;
  return FH_OK;
}

template <typename T>
static int rfftn_impl(fh_plan* pl, int ncomp, const T* x, cpx<T>* xhat) {
  const GParams g = make_gparams(pl, 1.0, 1);
  for (int c = 0; c < ncomp; ++c) {
    cpx<T>* out = xhat + (size_t)c * pl->nh;
    RETCHK(launch_rows_fwd<T>(pl, 1, x + (size_t)c * pl->n, nullptr, fh::COEF_NONE, out, nullptr));
    if (pl->n0 > 1) {
      RETCHK(launch_cols<T>(pl, 0, 1, 1, out, g, nullptr));
      RETCHK(launch_cols<T>(pl, 0, 0, 1, out, g, nullptr));
    } else {
      RETCHK(launch_cols<T>(pl, 0, 1, 1, out, g, nullptr));
    }
  }
  return FH_OK;
}

template <typename T>
static int irfftn_impl(fh_plan* pl, int ncomp, const cpx<T>* xhat, T* x) {
  const GParams g = make_gparams(pl, 1.0, 1);
  cpx<T>* tmp = reinterpret_cast<cpx<T>*>(pl->spec);
  for (int c = 0; c < ncomp; ++c) {
    HIPCHK(hipMemcpyAsync(tmp, xhat + (size_t)c * pl->nh, (size_t)pl->nh * sizeof(cpx<T>),
                          hipMemcpyDeviceToDevice, pl->stream));
    if (pl->n0 > 1) {
      RETCHK(launch_cols<T>(pl, 1, 0, 1, tmp, g, nullptr));
      RETCHK(launch_cols<T>(pl, 1, 1, 1, tmp, g, nullptr));
    } else {
      RETCHK(launch_cols<T>(pl, 1, 1, 1, tmp, g, nullptr));
    }
    RETCHK(launch_rows_inv<T>(pl, 1, tmp, x + (size_t)c * pl->n, T(1.0 / (double)pl->n), nullptr, nullptr,
                              nullptr, nullptr));
  }
  return FH_OK;
}

// ---- CG ---------------------------------------------------------------------------------

static int ensure_work(fh_plan* pl, size_t bytes) {
  if (pl->work_bytes >= bytes) return FH_OK;
  if (pl->work) HIPCHK(hipFree(pl->work));
  pl->work = nullptr;
  pl->work_bytes = 0;
  HIPCHK(hipMalloc(&pl->work, bytes));
  pl->work_bytes = bytes;
  return FH_OK;
}

static int ensure_hist(fh_plan* pl, int len) {
  if (pl->hist_cap >= len) return FH_OK;
  if (pl->hist) HIPCHK(hipFree(pl->hist));
  pl->hist = nullptr;
  pl->hist_cap = 0;
  HIPCHK(hipMalloc(&pl->hist, (size_t)len * sizeof(double)));
  pl->hist_cap = len;
  return FH_OK;
}

// Fused single-precision CG iteration (fixed-length kernels, isotropic or Lame coefficient):
//   rows_fwd : xc += alpha_prev p_old (pending step), p = r + beta p, y = A p, partial <p, A p>
//   alpha    : alpha = <r,r> / <p, A p>   (for p in the range of G, <p, A p> = <p, G(A p)>)
//   cols     : F, G, F^-1 along the strided axes
//   rows_inv : r -= alpha G(A p), partial <r, r>
//   beta     : beta, stop / reliable-update test
// Returns 0 when the fused kernels do not cover this plan (nothing was enqueued).
// Partials per pass that the deferred alpha/beta of a small grid sums (the forward pass's go to
// pl->partial[0..kDeferMax), the inverse pass's to pl->partial[kDeferMax..2 kDeferMax)).
constexpr int kDeferMax = 2048;

// Deferred alpha/beta: at most 6 * kDeferMax grid rows (every fixed-length row pass has at least 6
// rows per workgroup, so at most kDeferMax workgroups) and room for both partial runs.
static bool defer_applies(const fh_plan* pl) {
  return pl->defer_ok && (long long)pl->n0 * pl->n1 <= 6LL * kDeferMax && pl->partial_cap >= 2 * (size_t)kDeferMax;
}

// One fused fp32 CG iteration. With pl->cg_defer (small grids) alpha and beta have no kernels of
// their own: the forward row pass takes the previous beta from the inverse pass's partials, block 0
// of the first strided pass finalises beta and alpha, and the inverse row pass marks its beta as
// deferred; k_cg_flush finalises the last beta of a batch.
static int enqueue_fused_iteration(fh_plan* pl, float* xc, float* r, float* p, const float* coef, bool* ok) {
  *ok = false;
  if (pl->generic || (pl->coef_kind != fh::COEF_ISO && pl->coef_kind != fh::COEF_LAME)) return FH_OK;
  const int nc = pl->ncomp;
  const int* halt = &pl->st->halt;
  cpx<float>* spec = reinterpret_cast<cpx<float>*>(pl->spec);
  const bool defer = pl->cg_defer;
  double* part1 = pl->partial;
  double* part2 = defer ? pl->partial + kDeferMax : pl->partial;
  fh::LaunchCtx ctx = make_ctx<float>(pl);
  if (defer) {
    ctx.bpart = part2;
    ctx.bnp = pl->defer_nb2;
  }
  int nb = 0;
  int rr = fh::fixed_rows_fwd<float>(ctx, nc, p, coef, pl->coef_kind, spec, halt, r, pl->st, xc, part1,
                                     pl->partial_cap, &nb);
  if (rr == 0) return FH_OK;
  RETCHK(fixed_result(rr, "rows forward (fused)"));
  *ok = true;
  if (defer && nb > kDeferMax) return fail(FH_ERR_STATE, "deferred alpha: too many row workgroups");
  const GParams g = make_gparams(pl, 1.0 / (double)pl->n);
  // first strided pass: axis 1 forward (3-D) or the projection pass (2-D); its block 0 finalises
  // alpha (needed two passes later, by the inverse row pass) while the others transform, and on
  // small grids first the deferred beta (the partials of the last inverse pass)
  bool first_done = false;
  {
    fh::LaunchCtx cf = make_ctx<float>(pl);
    cf.fin = fh::CgFin{pl->st, part1, defer ? part2 : nullptr, nb, defer ? pl->defer_nb2 : 0};
    const int rf = pl->n0 > 1 ? fh::fixed_cols<float>(cf, -1, 1, nc, spec, halt)
                              : fh::fixed_cols_g<float>(cf, 1, nc, spec, g, halt);
    if (rf != 0) {
      RETCHK(fixed_result(rf, "strided pass (fused, finalising)"));
      first_done = true;
    } else {
      // no fixed kernel on this axis: finalise with the one-workgroup kernels
      if (defer)
        hipLaunchKernelGGL(fh::k_cg_flush, dim3(1), dim3(32), 0, pl->stream, pl->st, part2, pl->defer_nb2);
      hipLaunchKernelGGL(fh::k_cg_alpha, dim3(1), dim3(kFinThreads), 0, pl->stream, pl->st, part1, nb);
    }
  }
  if (pl->n0 > 1) {
    if (!first_done) RETCHK(launch_cols<float>(pl, 0, 1, nc, spec, g, halt));
    RETCHK(launch_cols<float>(pl, 2, 0, nc, spec, g, halt));
    RETCHK(launch_cols<float>(pl, 1, 1, nc, spec, g, halt));
  } else if (!first_done) {
    RETCHK(launch_cols<float>(pl, 2, 1, nc, spec, g, halt));
  }
  int nb2 = 0;
  fh::LaunchCtx ci = make_ctx<float>(pl);
  ci.defer_beta = defer ? 1 : 0;
  rr = fh::fixed_rows_inv<float>(ci, nc, spec, r, 1.0f, 2, nullptr, part2, defer ? (size_t)kDeferMax : pl->partial_cap,
                                 halt, &nb2, pl->st);
  if (rr == 0) return fail(FH_ERR_STATE, "fused rows inverse unavailable");
  RETCHK(fixed_result(rr, "rows inverse (fused)"));
  if (defer)
    pl->defer_nb2 = nb2;
  else
    hipLaunchKernelGGL(fh::k_cg_beta, dim3(1), dim3(kFinThreads), 0, pl->stream, pl->st, part2, nb2);
  HIPCHK(hipGetLastError());
  return FH_OK;
}

// End of a batch of fused iterations with deferred beta: finalise it for the host.
static void enqueue_flush(fh_plan* pl) {
  if (pl->cg_defer)
    hipLaunchKernelGGL(fh::k_cg_flush, dim3(1), dim3(32), 0, pl->stream, pl->st, pl->partial + kDeferMax,
                       pl->defer_nb2);
}

// One CG iteration in precision T on vectors (x, r, p, Ap); everything halts on st->halt.
// The direction update p = r + beta p of the previous iteration happens inside the first
// pass of the matvec (beta = 0 with p = r on the first iteration).
// With a Lame or full coefficient a voxel's components are spread over several threads of
// the first pass, so the in-place update runs as its own kernel there.
template <typename T>
static int enqueue_iteration(fh_plan* pl, T* x, T* r, T* p, T* Ap, const T* coef) {
  const long long nv = (long long)pl->ncomp * pl->n;
  const int* halt = &pl->st->halt;
  int nb = 0;
  const bool fused = pl->coef_kind == fh::COEF_ISO;
  if (!fused)
    hipLaunchKernelGGL((fh::k_cg_pupdate<T>), dim3(grid_blocks(nv)), dim3(256), 0, pl->stream, nv, p, r, pl->st);
  RETCHK(run_operator<T>(pl, p, Ap, pl->coef_kind, coef, p, halt, &nb, fused ? r : nullptr));
  hipLaunchKernelGGL(fh::k_cg_alpha, dim3(1), dim3(kFinThreads), 0, pl->stream, pl->st, pl->partial, nb);
  const int gb = grid_blocks(nv);
  hipLaunchKernelGGL((fh::k_cg_update<T>), dim3(gb), dim3(256), 0, pl->stream, nv, x, r, p, Ap, pl->st,
                     pl->partial);
  hipLaunchKernelGGL(fh::k_cg_beta, dim3(1), dim3(kFinThreads), 0, pl->stream, pl->st, pl->partial, gb);
  HIPCHK(hipGetLastError());
  return FH_OK;
}

// r = b - A x (fp64 operator); rr, beta, done via k_cg_restart
template <typename TR>
static int enqueue_true_residual(fh_plan* pl, const double* b, const double* x, double* Ax, TR* r,
                                 bool x_zero = false, bool newdir = false) {
  const long long nv = (long long)pl->ncomp * pl->n;
  int np = grid_blocks(nv);
  bool fused = false;
  // x = 0: A x = 0 exactly, so r = b without the operator (the same bits as b - A 0)
  if (!x_zero) {
    if constexpr (sizeof(TR) == 4)
      RETCHK(run_operator_residual(pl, x, b, Ax, r, &fused, &np));
    else
      RETCHK(run_operator<double>(pl, x, Ax, pl->coef_kind, pl->coef64, nullptr, nullptr, nullptr));
  }
  if (!fused)
    hipLaunchKernelGGL((fh::k_residual<TR>), dim3(np), dim3(256), 0, pl->stream, nv, b, x_zero ? nullptr : Ax, r,
                       pl->partial);
  hipLaunchKernelGGL(fh::k_cg_restart, dim3(1), dim3(256), 0, pl->stream, pl->st, pl->partial, np, pl->hist,
                     newdir);
  HIPCHK(hipGetLastError());
  return FH_OK;
}

// 1 if every value of x is +0 or -0 (the usual initial guess x0 = 0 of a solve).
static int field_is_zero(fh_plan* pl, const double* x, bool* zero) {
  const long long nv = (long long)pl->ncomp * pl->n;
  int* flag = reinterpret_cast<int*>(pl->partial);  // free at this point of a solve
  HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), pl->stream));
  hipLaunchKernelGGL(fh::micro::k_count_nonzero, dim3(grid_blocks(nv, 256, 4096)), dim3(256), 0, pl->stream, nv, x,
                     flag);
  HIPCHK(hipGetLastError());
  int h = 1;
  HIPCHK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, pl->stream));
  HIPCHK(hipStreamSynchronize(pl->stream));
  *zero = h == 0;
  return FH_OK;
}

static int scratch_parts(double** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch) HIPCHK(hipMalloc(&g_scratch, kScratchParts * sizeof(double) + 64));
  *out = g_scratch;
  return FH_OK;
}


// Quadrature grid length for the Ga energy on an axis of N points: 2N-1 (the reference's
// Nbar, ffthompy/postprocess.py:24) when its prime factors are <= 13, else for odd N the next
// such length (the energy is exact on any length >= 2N-1 there, and the radix kernels are faster
// than Bluestein's), and for even N 2N-1 itself (only that grid reproduces the reference's aliased
// quadrature; its transforms use Bluestein's algorithm).
int quadrature_length(int N) {
  const int nb = 2 * N - 1;
  AxisPlan ap;
  if (factorize(nb, &ap)) return nb;
  if ((N & 1) == 0) return nb;
  for (int L = nb + 1; L < 2 * nb + 16; ++L)
    if (factorize(L, &ap)) return L;
  return nb;
}

// Inverse real FFT of one spectrum component held in `spec` (the plan's half-spectrum layout,
// transformed in place) to `out`, scaled by `scale`.
int spec_to_real(fh_plan* pl, cpx<double>* spec, double* out, double scale) {
  const GParams g = make_gparams(pl, 1.0, 1);
  if (pl->n0 > 1) {
    RETCHK(launch_cols<double>(pl, 1, 0, 1, spec, g, nullptr));
    RETCHK(launch_cols<double>(pl, 1, 1, 1, spec, g, nullptr));
  } else {
    RETCHK(launch_cols<double>(pl, 1, 1, 1, spec, g, nullptr));
  }
  return launch_rows_inv<double>(pl, 1, spec, out, scale, nullptr, nullptr, nullptr, nullptr);
}

int ga_ncoef(const fh_ga* ga, int kind) {
  if (kind == FH_COEF_ISO) return 1;
  if (kind == FH_COEF_LAME) return ga->coarse->physics == FH_ELASTICITY ? 2 : -1;
  if (kind == FH_COEF_FULL) return ga->coarse->ncomp * ga->coarse->ncomp;
  return -1;
}

int ga_alloc_coef(fh_ga* ga, int kind, int ncoef) {
  if (ga->coef) (void)hipFree(ga->coef);
  ga->coef = nullptr;
  ga->coef_kind = 0;
  ga->ncoef = 0;
  HIPCHK(hipMalloc(&ga->coef, (size_t)ncoef * ga->fine->n * sizeof(double)));
  ga->ncoef = ncoef;
  (void)kind;
  return FH_OK;
}

// Trigonometric interpolation of one coarse field (ncomp components) onto the quadrature grid.
int ga_interpolate(fh_ga* ga, const double* x, double* out) {
  fh_plan* c = ga->coarse;
  fh_plan* f = ga->fine;
  cpx<double>* cs = reinterpret_cast<cpx<double>*>(c->spec);
  cpx<double>* fs = reinterpret_cast<cpx<double>*>(f->spec);
  const int nb = grid_blocks(f->nh, 256, 8192);
  for (int k = 0; k < c->ncomp; ++k) {
    RETCHK(rfftn_impl<double>(c, 1, x + (size_t)k * c->n, cs));
    HIPCHK(hipStreamSynchronize(c->stream));
    hipLaunchKernelGGL(fh::ga::k_ga_pad, dim3(nb), dim3(256), 0, f->stream, cs, c->n0, c->n1, c->n2,
                       1.0 / (double)c->n, ga->g, fs);
    HIPCHK(hipGetLastError());
    RETCHK(spec_to_real(f, fs, out + (size_t)k * f->n, 1.0));
    HIPCHK(hipStreamSynchronize(f->stream));
  }
  return FH_OK;
}

// ---- potential formulation -----------------------------------------------------------------

fh::pot::PotParams pot_params(const fh_plan* pl) {
  fh::pot::PotParams g;
  g.n0 = pl->n0;
  g.n1 = pl->n1;
  g.n2 = pl->n2;
  g.n2h = pl->n2h;
  g.dim = pl->dim;
  g.ax0 = 3 - pl->dim;
  g.Y[0] = pl->dim == 3 ? pl->Y[0] : 1.0;
  g.Y[1] = pl->Y[pl->dim == 3 ? 1 : 0];
  g.Y[2] = pl->Y[pl->dim - 1];
  g.banded = pl->banded;
  for (int a = 0; a < 3; ++a) {
    g.band[a] = pl->band[a];
    g.tpiY[a] = (double)(2.0L * 3.14159265358979323846264338327950288L / (long double)g.Y[a]);
  }
  return g;
}

// The potential operator out = P (-div) F A F^-1 grad (P p) on the plan's spectra, p first
// replaced by r + beta p when r != null; partial sums of the weighted <out, w> into pl->partial
// (*nb workgroups). Along the outermost axis the expansion into the gradient and the contraction
// of the divergence run inside that axis' strided passes, and along the contiguous axis C2R,
// a(x) and R2C run as one in-place kernel (the real field stays in LDS), where the fixed lengths
// cover the grid; otherwise separate elementwise kernels and the one-way passes (with the real
// field in `re`, ncomp * n doubles).
template <typename T>
int pot_operator(fh_plan* pl, cpx<T>* p, const cpx<T>* r, cpx<T>* out, T* re, const cpx<T>* w, const int* halt,
                 int* nb) {
  const int nc = pl->ncomp;
  const fh::pot::PotParams g = pot_params(pl);
  cpx<T>* spec = reinterpret_cast<cpx<T>*>(pl->spec);
  const GParams gp = make_gparams(pl, 1.0);
  const int gb = grid_blocks(pl->nh);
  const int ax = pl->n0 > 1 ? 0 : 1;  // outermost axis of the grid
  const fh::LaunchCtx ctx = make_ctx<T>(pl);
  const T* coef = plan_coef<T>(pl);
  int in_fused = 0;
  if (!pl->generic) {
    in_fused = fh::fixed_pot_cols<T>(ctx, ax, +1, nc, spec, p, r, pl->st, T(0), nullptr, nullptr, nullptr, 0, g, halt,
                                     nullptr);
    if (in_fused < 0) return fixed_result(in_fused, "potential strided pass (fixed)");
  }
  if (!in_fused) {
    if (r)
      hipLaunchKernelGGL((fh::k_cg_pupdate<T>), dim3(gb), dim3(256), 0, pl->stream, 2 * pl->nh, (T*)p, (const T*)r,
                         pl->st);
    hipLaunchKernelGGL((fh::pot::k_pot_grad<T>), dim3(gb), dim3(256), 0, pl->stream, p, spec, g, 1, halt);
    HIPCHK(hipGetLastError());
    RETCHK(launch_cols<T>(pl, 1, ax, nc, spec, gp, halt));
  }
  if (ax == 0) RETCHK(launch_cols<T>(pl, 1, 1, nc, spec, gp, halt));
  int rt_fused = 0;
  if (!pl->generic && pl->coef_kind == fh::COEF_ISO) {
    rt_fused = fh::fixed_rows_rt<T>(ctx, nc, spec, coef, T(1.0 / (double)pl->n), halt);
    if (rt_fused < 0) return fixed_result(rt_fused, "rows round trip (fixed)");
  }
  if (!rt_fused) {
    RETCHK(launch_rows_inv<T>(pl, nc, spec, re, T(1.0 / (double)pl->n), nullptr, nullptr, halt, nullptr));
    RETCHK(launch_rows_fwd<T>(pl, nc, re, coef, pl->coef_kind, spec, halt));
  }
  if (ax == 0) RETCHK(launch_cols<T>(pl, 0, 1, nc, spec, gp, halt));
  int out_fused = 0;
  if (!pl->generic) {
    out_fused = fh::fixed_pot_cols<T>(ctx, ax, -1, nc, spec, nullptr, nullptr, nullptr, T(-1), out, w, pl->partial,
                                      pl->partial_cap, g, halt, nb);
    if (out_fused < 0) return fixed_result(out_fused, "potential strided pass (fixed)");
  }
  if (!out_fused) {
    RETCHK(launch_cols<T>(pl, 0, ax, nc, spec, gp, halt));
    hipLaunchKernelGGL((fh::pot::k_pot_div<T>), dim3(gb), dim3(256), 0, pl->stream, spec, out, g, T(-1), w,
                       pl->partial, halt);
    HIPCHK(hipGetLastError());
    if (nb) *nb = gb;
  }
  return FH_OK;
}

// Forget the captured iteration graph of a plan: its kernels and arguments were chosen for the
// plan's projection, band, kernel family and coefficient at capture time. Called by every entry
// point that changes one of those (all entry points have finished their GPU work on return).
void drop_graphs(fh_plan* pl) {
  for (auto& e : pl->it_exec) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
}

// Length of the next batch of CG iterations between two host synchronisations: the iterations
// that the convergence rate since the last synchronisation predicts until the residual norm falls
// below `target` (the next stop or reliable-update threshold), plus one, rounded up to a power of
// two within [1, cap] (so that at most four captured batch graphs exist per plan). An iteration
// enqueued past the stop returns at once on the halt flag but still dispatches its grids (the
// waste of a fixed batch of 8 is up to 7 iterations at every stop and reliable update); a batch
// one short costs one more synchronisation.
struct BatchPredictor {
  double nr_prev = -1.0;
  int kit_prev = 0;
  double rate = -1.0;
  int next(const CGState& s, double target, int cap) {
    const double nr = std::sqrt(s.rr);
    if (nr_prev > 0.0 && s.kit > kit_prev && nr > 0.0)
      rate = nr < nr_prev ? std::pow(nr / nr_prev, 1.0 / (double)(s.kit - kit_prev)) : -1.0;
    nr_prev = nr;
    kit_prev = s.kit;
    if (!(rate > 0.0 && rate < 1.0) || !(target > 0.0) || !(nr > target)) return cap;
    const double n = std::log(target / nr) / std::log(rate);
    if (!(n < (double)cap)) return cap;
    const int want = (int)std::ceil(n) + 1;
    int len = 1;
    while (len < want && len < cap) len *= 2;
    return std::min(len, cap);
  }
};

}  // namespace

extern "C" {

const char* fh_last_error(void) { return g_err.c_str(); }

int fh_init(char* name, int name_len) {
  RETCHK(ensure_device());
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s", g_dev_name.c_str());
  }
  return FH_OK;
}

int fh_malloc(void** dptr, size_t bytes) {
  if (!dptr) return fail(FH_ERR_ARG, "null pointer argument");
  RETCHK(ensure_device());
  *dptr = nullptr;
  if (bytes == 0) bytes = 256;
  HIPCHK(hipMalloc(dptr, bytes));
  return FH_OK;
}

int fh_free(void* dptr) {
  if (!dptr) return FH_OK;
  RETCHK(ensure_device());
  HIPCHK(hipFree(dptr));
  return FH_OK;
}

int fh_memcpy_htod(void* dst, const void* src, size_t bytes) {
  RETCHK(ensure_device());
  if (bytes == 0) return FH_OK;
  if (!dst || !src) return fail(FH_ERR_ARG, "null pointer argument");
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return FH_OK;
}

int fh_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
  RETCHK(ensure_device());
  if (bytes == 0) return FH_OK;
  if (!dst || !src) return fail(FH_ERR_ARG, "null pointer argument");
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return FH_OK;
}

int fh_memcpy_dtod(void* dst, const void* src, size_t bytes) {
  RETCHK(ensure_device());
  if (bytes == 0) return FH_OK;
  if (!dst || !src) return fail(FH_ERR_ARG, "null pointer argument");
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
  return FH_OK;
}

int fh_memset_zero(void* dst, size_t bytes) {
  RETCHK(ensure_device());
  if (bytes == 0) return FH_OK;
  if (!dst) return fail(FH_ERR_ARG, "null pointer argument");
  HIPCHK(hipMemset(dst, 0, bytes));
  HIPCHK(hipDeviceSynchronize());
  return FH_OK;
}

int fh_plan_destroy(fh_plan* pl) {
  if (!pl) return FH_OK;
  for (int a = 0; a < 3; ++a) {
    if (pl->tw64[a]) (void)hipFree(pl->tw64[a]);
    if (pl->tw32[a]) (void)hipFree(pl->tw32[a]);
  }
  if (pl->coef64) (void)hipFree(pl->coef64);
  if (pl->coef32) (void)hipFree(pl->coef32);
  if (pl->spec) (void)hipFree(pl->spec);
  if (pl->partial) (void)hipFree(pl->partial);
  if (pl->st) (void)hipFree(pl->st);
  if (pl->hist) (void)hipFree(pl->hist);
  if (pl->work) (void)hipFree(pl->work);
  drop_graphs(pl);
  if (pl->ev0) (void)hipEventDestroy(pl->ev0);
  if (pl->ev1) (void)hipEventDestroy(pl->ev1);
  if (pl->stream) (void)hipStreamDestroy(pl->stream);
  delete pl;
  return FH_OK;
}

int fh_plan_create(int dim, const int* N, const double* Y, int physics, int primaldual, fh_plan** out) {
  if (!out || !N || !Y) return fail(FH_ERR_ARG, "null pointer argument");
  *out = nullptr;
  if (dim != 2 && dim != 3) return fail(FH_ERR_ARG, "dim must be 2 or 3");
  if (physics != FH_SCALAR && physics != FH_ELASTICITY) return fail(FH_ERR_ARG, "bad physics");
  if (primaldual != FH_PRIMAL && primaldual != FH_DUAL) return fail(FH_ERR_ARG, "bad primaldual");
  for (int a = 0; a < dim; ++a) {
    if (N[a] < 1 || N[a] > (1 << 20)) return fail(FH_ERR_ARG, "grid size out of range");
    if (!(Y[a] > 0.0) || !std::isfinite(Y[a])) return fail(FH_ERR_ARG, "cell size must be positive");
  }
  RETCHK(ensure_device());
  fh_plan* pl = new fh_plan();
  pl->dim = dim;
  pl->physics = physics;
  pl->primaldual = primaldual;
  pl->ncomp = physics == FH_SCALAR ? dim : dim * (dim + 1) / 2;
  for (int a = 0; a < dim; ++a) {
    pl->Nu[a] = N[a];
    pl->Y[a] = Y[a];
  }
  if (dim == 3) {
    pl->n0 = N[0];
    pl->n1 = N[1];
    pl->n2 = N[2];
  } else {
    pl->n0 = 1;
    pl->n1 = N[0];
    pl->n2 = N[1];
  }
  pl->n2h = pl->n2 / 2 + 1;
  pl->n = (long long)pl->n0 * pl->n1 * pl->n2;
  pl->nh = (long long)pl->n0 * pl->n1 * pl->n2h;
  const int lens[3] = {pl->n0, pl->n1, pl->n2};
  auto bail = [&](int code) {
    fh_plan_destroy(pl);
    return code;
  };
  for (int a = 0; a < 3; ++a) {
    if (!plan_axis(lens[a], &pl->ax[a])) {
      fh_plan_destroy(pl);
      return fail(FH_ERR_UNSUPPORTED, "grid length " + std::to_string(lens[a]) +
                                          ": its Bluestein convolution does not fit one workgroup");
    }
    const std::vector<long double> tl = axis_table(pl->ax[a]);
    std::vector<double> t64(tl.size());
    std::vector<float> t32(tl.size());
    for (size_t m = 0; m < tl.size(); ++m) {
      t64[m] = (double)tl[m];
      t32[m] = (float)tl[m];
    }
    if (hipMalloc(&pl->tw64[a], t64.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&pl->tw32[a], t32.size() * sizeof(float)) != hipSuccess)
      return bail(fail(FH_ERR_HIP, "hipMalloc twiddles failed"));
    if (hipMemcpy(pl->tw64[a], t64.data(), t64.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(pl->tw32[a], t32.data(), t32.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(FH_ERR_HIP, "hipMemcpy twiddles failed"));
  }
  // validate launch shapes once (fp64 is the largest)
  {
    RowCfg rc;
    ColCfg cc;
    int r = row_config(pl->ax[2], pl->ncomp, sizeof(cpx<double>), (long long)pl->n0 * pl->n1, &rc);
    if (r != FH_OK) return bail(r);
    if (pl->n0 > 1) {
      r = col_config(pl->ax[1], 4, sizeof(cpx<double>), (long long)pl->n2h * pl->ncomp, &cc);
      if (r != FH_OK) return bail(r);
      r = col_config(pl->ax[0], pl->ncomp, sizeof(cpx<double>), (long long)pl->n1 * pl->n2h * pl->ncomp, &cc);
      if (r != FH_OK) return bail(r);
    } else {
      r = col_config(pl->ax[1], pl->ncomp, sizeof(cpx<double>), (long long)pl->n2h * pl->ncomp, &cc);
      if (r != FH_OK) return bail(r);
    }
    RowCfg rc32;
    row_config(pl->ax[2], pl->ncomp, sizeof(cpx<float>), (long long)pl->n0 * pl->n1, &rc32);
    pl->partial_cap = (size_t)std::max<long long>({(long long)rc.nblocks, (long long)rc32.nblocks, (long long)kScratchParts,
                                                       2LL * 36 * 1024, (long long)pl->n0 * pl->n1});
  }
  if (hipStreamCreateWithFlags(&pl->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&pl->ev0) != hipSuccess || hipEventCreate(&pl->ev1) != hipSuccess)
    return bail(fail(FH_ERR_HIP, "stream/event creation failed"));
  const size_t spec_bytes = (size_t)pl->ncomp * pl->nh * sizeof(cpx<double>);
  if (hipMalloc(&pl->spec, spec_bytes) != hipSuccess ||
      hipMalloc(&pl->partial, (pl->partial_cap + kAuxDoubles) * sizeof(double)) != hipSuccess ||
      hipMalloc(&pl->st, 2 * sizeof(CGState)) != hipSuccess)
    return bail(fail(FH_ERR_HIP, "hipMalloc plan buffers failed (out of device memory?)"));
  pl->aux = pl->partial + pl->partial_cap;
  pl->st_add = pl->st + 1;
  {
    CGState u;
    std::memset(&u, 0, sizeof(u));
    u.alpha = -1.0;
    if (hipMemcpy(pl->st_add, &u, sizeof(u), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(FH_ERR_HIP, "plan state initialisation failed"));
  }
  *out = pl;
  return FH_OK;
}

int fh_plan_ncomp(const fh_plan* pl) { return pl ? pl->ncomp : -1; }

int fh_plan_use_generic(fh_plan* pl, int generic) {
  RETCHK(check_plan(pl));
  drop_graphs(pl);
  pl->generic = generic ? 1 : 0;
  return FH_OK;
}

int fh_plan_set_cg_variant(fh_plan* pl, int fused) {
  RETCHK(check_plan(pl));
  if (fused < 0 || fused > 2) return fail(FH_ERR_ARG, "fh_plan_set_cg_variant: 0, 1 or 2");
  drop_graphs(pl);
  pl->fused_cg = fused != 0;
  pl->defer_ok = fused == 1;
  return FH_OK;
}

int fh_plan_set_primaldual(fh_plan* pl, int primaldual) {
  RETCHK(check_plan(pl));
  drop_graphs(pl);
  if (primaldual != FH_PRIMAL && primaldual != FH_DUAL) return fail(FH_ERR_ARG, "bad primaldual");
  pl->primaldual = primaldual;
  return FH_OK;
}

int fh_plan_use_graphs(fh_plan* pl, int enable) {
  RETCHK(check_plan(pl));
  pl->use_graphs = enable != 0;
  return FH_OK;
}

int fh_plan_set_residual_log(fh_plan* pl, int every) {
  RETCHK(check_plan(pl));
  if (every < 0) return fail(FH_ERR_ARG, "every must be >= 0");
  pl->res_every = every;
  return FH_OK;
}

int fh_plan_set_iterate_callback(fh_plan* pl, fh_iterate_fn fn, void* user, int every) {
  RETCHK(check_plan(pl));
  if (every < 0) return fail(FH_ERR_ARG, "every must be >= 0");
  pl->it_fn = every > 0 ? fn : nullptr;
  pl->it_user = user;
  pl->it_every = fn ? every : 0;
  return FH_OK;
}

int fh_last_residual_log(const fh_plan* pl, int* kits, double* norms, int cap, int* count) {
  RETCHK(check_plan(pl));
  if (!count || cap < 0 || (cap > 0 && (!kits || !norms))) return fail(FH_ERR_ARG, "bad arguments");
  const int n = (int)pl->res_norm.size();
  *count = n;
  for (int i = 0; i < std::min(n, cap); ++i) {
    kits[i] = pl->res_kit[i];
    norms[i] = pl->res_norm[i];
  }
  return FH_OK;
}

int fh_last_solve_graph_batches(const fh_plan* pl, int* n) {
  RETCHK(check_plan(pl));
  if (!n) return fail(FH_ERR_ARG, "null pointer argument");
  *n = pl->last_graph_batches;
  return FH_OK;
}

int fh_plan_set_band(fh_plan* pl, const int* band, double scale) {
  RETCHK(check_plan(pl));
  drop_graphs(pl);
  if (!(scale > 0.0) || !std::isfinite(scale)) return fail(FH_ERR_ARG, "scale must be positive");
  if (!band) {
    pl->banded = 0;
    pl->op_scale = 1.0;
    return FH_OK;
  }
  const int off = 3 - pl->dim;
  int v[3] = {0, 0, 0};
  for (int a = 0; a < pl->dim; ++a) {
    if (band[a] < 0) return fail(FH_ERR_ARG, "band must be >= 0");
    v[a + off] = band[a];
  }
  for (int a = 0; a < 3; ++a) pl->band[a] = v[a];
  pl->banded = 1;
  pl->op_scale = scale;
  return FH_OK;
}

int fh_set_coeff(fh_plan* pl, int kind, const double* host_vals) {
  RETCHK(check_plan(pl));
  drop_graphs(pl);
  if (!host_vals) return fail(FH_ERR_ARG, "null coefficient array");
  int per;
  if (kind == FH_COEF_ISO) {
    per = 1;
  } else if (kind == FH_COEF_LAME) {
    if (pl->physics != FH_ELASTICITY) return fail(FH_ERR_ARG, "Lame coefficients need elasticity");
    per = 2;
  } else if (kind == FH_COEF_FULL) {
    per = pl->ncomp * pl->ncomp;
  } else {
    return fail(FH_ERR_ARG, "bad coefficient kind");
  }
  const size_t cnt = (size_t)per * pl->n;
  for (size_t i = 0; i < cnt; ++i)
    if (!std::isfinite(host_vals[i])) return fail(FH_ERR_ARG, "coefficient has a non-finite value");
  if (pl->coef64) (void)hipFree(pl->coef64);
  if (pl->coef32) (void)hipFree(pl->coef32);
  pl->coef64 = nullptr;
  pl->coef32 = nullptr;
  pl->coef_kind = 0;
  HIPCHK(hipMalloc(&pl->coef64, cnt * sizeof(double)));
  HIPCHK(hipMalloc(&pl->coef32, cnt * sizeof(float)));
  HIPCHK(hipMemcpy(pl->coef64, host_vals, cnt * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL((fh::k_convert<double, float>), dim3(grid_blocks((long long)cnt)), dim3(256), 0,
                     pl->stream, (long long)cnt, pl->coef64, pl->coef32);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pl->stream));
  pl->coef_kind = kind;
  return FH_OK;
}

int fh_matvec(fh_plan* pl, int dtype, const void* x, void* y) { return matvec_impl(pl, dtype, x, y, true); }

int fh_project(fh_plan* pl, int dtype, const void* x, void* y) { return matvec_impl(pl, dtype, x, y, false); }

int fh_rfftn(fh_plan* pl, int dtype, int ncomp, const void* x, void* xhat) {
  RETCHK(check_plan(pl));
  RETCHK(check_dtype(dtype));
  if (!x || !xhat || ncomp < 1) return fail(FH_ERR_ARG, "bad field arguments");
  if (dtype == FH_F64)
    RETCHK(rfftn_impl<double>(pl, ncomp, (const double*)x, (cpx<double>*)xhat));
  else
    RETCHK(rfftn_impl<float>(pl, ncomp, (const float*)x, (cpx<float>*)xhat));
  HIPCHK(hipStreamSynchronize(pl->stream));
  return FH_OK;
}

int fh_irfftn(fh_plan* pl, int dtype, int ncomp, const void* xhat, void* x) {
  RETCHK(check_plan(pl));
  RETCHK(check_dtype(dtype));
  if (!x || !xhat || ncomp < 1) return fail(FH_ERR_ARG, "bad field arguments");
  if (dtype == FH_F64)
    RETCHK(irfftn_impl<double>(pl, ncomp, (const cpx<double>*)xhat, (double*)x));
  else
    RETCHK(irfftn_impl<float>(pl, ncomp, (const cpx<float>*)xhat, (float*)x));
  HIPCHK(hipStreamSynchronize(pl->stream));
  return FH_OK;
}

int fh_apply_projection_hat(fh_plan* pl, int dtype, const void* in_hat, void* out_hat) {
  RETCHK(check_plan(pl));
  RETCHK(check_dtype(dtype));
  if (!in_hat || !out_hat) return fail(FH_ERR_ARG, "null spectrum pointer");
  const GParams g = make_gparams(pl, 1.0);
  const dim3 grid((unsigned)((pl->nh + 255) / 256)), block(256);
  if (dtype == FH_F64)
    hipLaunchKernelGGL(fh::k_apply_G_soa<double>, grid, block, 0, pl->stream, (const cpx<double>*)in_hat,
                       (cpx<double>*)out_hat, pl->nh, pl->n2h, g);
  else
    hipLaunchKernelGGL(fh::k_apply_G_soa<float>, grid, block, 0, pl->stream, (const cpx<float>*)in_hat,
                       (cpx<float>*)out_hat, pl->nh, pl->n2h, g);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pl->stream));
  return FH_OK;
}

int fh_cg_solve(fh_plan* pl, int mode, const void* b_, void* x_, double tol, int maxiter, double delta,
                int* kit, double* norm_res, double* res_hist, int hist_len, int* cg_status) {
  RETCHK(check_plan(pl));
  if (!b_ || !x_) return fail(FH_ERR_ARG, "null field pointer");
  if (mode != FH_SOLVE_FP64 && mode != FH_SOLVE_MIXED) return fail(FH_ERR_ARG, "bad solve mode");
  if (!pl->coef_kind) return fail(FH_ERR_STATE, "fh_set_coeff has not been called");
  if (!(tol >= 0.0) || maxiter < 0) return fail(FH_ERR_ARG, "bad tol/maxiter");
  if (mode == FH_SOLVE_MIXED && !(delta > 0.0 && delta < 1.0)) return fail(FH_ERR_ARG, "delta must be in (0,1)");
  if (mode == FH_SOLVE_MIXED && pl->res_every > 0)
    return fail(FH_ERR_UNSUPPORTED, "the residual log (fh_plan_set_residual_log) is kept by fp64 solves only");
  if (mode == FH_SOLVE_MIXED && pl->it_every > 0)
    return fail(FH_ERR_UNSUPPORTED, "the iterate callback (fh_plan_set_iterate_callback) runs in fp64 solves only");
  pl->res_kit.clear();
  pl->res_norm.clear();
  const double* b = (const double*)b_;
  double* x = (double*)x_;
  const long long nv = (long long)pl->ncomp * pl->n;
  const int hl = std::max(hist_len, 1);
  RETCHK(ensure_hist(pl, hl));
  CGState s;
  std::memset(&s, 0, sizeof(s));
  s.tol = tol;
  s.maxiter = maxiter;
  s.mixed = mode == FH_SOLVE_MIXED;
  s.delta = delta;
  s.inv_nvox = 1.0 / (double)pl->n;
  s.hist_len = hist_len > 0 ? hist_len : 0;
  s.hist = pl->hist;
  HIPCHK(hipMemcpyAsync(pl->st, &s, sizeof(s), hipMemcpyHostToDevice, pl->stream));
  HIPCHK(hipEventRecord(pl->ev0, pl->stream));
  const int per_check = 8;
  bool x0_zero = false;
  RETCHK(field_is_zero(pl, x, &x0_zero));
  if (mode == FH_SOLVE_FP64) {
    const long long nvp = (nv + 63) / 64 * 64;  // 512-byte aligned sub-buffers
    RETCHK(ensure_work(pl, 3 * (size_t)nvp * sizeof(double)));
    double* r = (double*)pl->work;
    double* p = r + nvp;
    double* Ap = p + nvp;
    RETCHK(enqueue_true_residual<double>(pl, b, x, Ap, r, x0_zero));
    HIPCHK(hipMemcpyAsync(p, r, (size_t)nv * sizeof(double), hipMemcpyDeviceToDevice, pl->stream));
    // residual log: ||b - A x|| in fp64 (the reference's CallBack, solver_pp.py:13-24) at x0 and
    // after every res_every-th iteration and the last; Ap is free between iterations
    const int every = pl->res_every;
    int logged = -1;
    auto log_true = [&](int kit_now) -> int {
      RETCHK(run_operator<double>(pl, x, Ap, pl->coef_kind, pl->coef64, nullptr, nullptr, nullptr));
      const int nb = grid_blocks(nv);
      hipLaunchKernelGGL(fh::k_resid_norm_partial, dim3(nb), dim3(256), 0, pl->stream, nv, b, Ap, pl->partial);
      double* dsum = pl->partial + nb;
      hipLaunchKernelGGL(fh::k_sum_partials, dim3(1), dim3(256), 0, pl->stream, pl->partial, nb, 1.0 / (double)pl->n,
                         dsum);
      HIPCHK(hipGetLastError());
      double h = 0.0;
      HIPCHK(hipMemcpyAsync(&h, dsum, sizeof(double), hipMemcpyDeviceToHost, pl->stream));
      HIPCHK(hipStreamSynchronize(pl->stream));
      pl->res_kit.push_back(kit_now);
      pl->res_norm.push_back(std::sqrt(h));
      logged = kit_now;
      return FH_OK;
    };
    if (every > 0) RETCHK(log_true(0));
    // host callback on the iterate (the reference's callback(xCG) after every iteration,
    // ffthompy/general/solver.py:134-136): x is current after each synchronisation
    const int cb_every = pl->it_fn ? pl->it_every : 0;
    int cb_last = 0;
    int kit_host = 0;
    BatchPredictor pred;
    int next_len = per_check;
    for (;;) {
      // batches end on the logged / called-back iterations (multiples of every, cb_every), at
      // most next_len long
      int batch = next_len;
      if (every > 0) batch = std::min(batch, every - kit_host % every);
      if (cb_every > 0) batch = std::min(batch, cb_every - kit_host % cb_every);
      for (int k = 0; k < batch; ++k) RETCHK(enqueue_iteration<double>(pl, x, r, p, Ap, pl->coef64));
      HIPCHK(hipMemcpyAsync(&s, pl->st, sizeof(s), hipMemcpyDeviceToHost, pl->stream));
      HIPCHK(hipStreamSynchronize(pl->stream));
      kit_host = s.kit;
      next_len = pred.next(s, s.tol, per_check);
      if (s.done && x0_zero) {
        // x <- G x once at the end: the recurrence leaves x outside the range of G by its accumulated
        // rounding, which enters the energy (and the bound ordering) at first order; the projection
        // removes it (x changes by that rounding only). With a nonzero x0 its component outside the
        // range of G belongs to the reference's result and is kept.
        // (in place: the forward row pass has read all of x before the inverse one writes it)
        RETCHK(run_operator<double>(pl, x, x, fh::COEF_NONE, nullptr, nullptr, nullptr, nullptr));
        HIPCHK(hipStreamSynchronize(pl->stream));
      }
      if (every > 0 && s.kit > logged && (s.kit % every == 0 || s.done)) RETCHK(log_true(s.kit));
      if (cb_every > 0 && s.kit > cb_last && (s.kit % cb_every == 0 || s.done)) {
        cb_last = s.kit;
        if (pl->it_fn(pl->it_user, s.kit, x) != 0) return fail(FH_ERR_STATE, "solve stopped by the iterate callback");
      }
      if (s.done) break;
    }
  } else {
    const long long nvp = (nv + 63) / 64 * 64;  // 256-byte aligned sub-buffers
    RETCHK(ensure_work(pl, 4 * (size_t)nvp * sizeof(float) + (size_t)nvp * sizeof(double)));
    double* Ax = (double*)pl->work;
    float* r = (float*)(Ax + nvp);
    float* p = r + nvp;
    float* Ap = p + nvp;
    float* xc = Ap + nvp;
    HIPCHK(hipMemsetAsync(xc, 0, (size_t)nv * sizeof(float), pl->stream));
    RETCHK(enqueue_true_residual<float>(pl, b, x, Ax, r, x0_zero));
    HIPCHK(hipMemcpyAsync(p, r, (size_t)nv * sizeof(float), hipMemcpyDeviceToDevice, pl->stream));
    // fused iterations where the fixed-length kernels cover the plan (isotropic coefficient)
    bool fused = pl->fused_cg && pl->op_scale == 1.0;  // the fused step assumes an unscaled G
    pl->cg_defer = fused && defer_applies(pl);
    pl->defer_nb2 = 0;
    const int gb = grid_blocks(nv);
    // The single-precision correction c is folded as x += G c, c projected in fp64 (in Ax, free
    // until the true residual overwrites it): rounding in the fp32 recurrence leaves c outside the
    // range of G by about 1e-5 of its size. Folded as it is, that part would enter every later
    // true residual, the iteration would offset it with an error inside the range, and a final
    // projection of x would leave that error in the solution (its residual about |A| times the
    // drift); left in x instead, it would enter the energies at first order. The true residual
    // then differs from the recurrence by more than rounding, so the iteration restarts its search
    // direction (p = r).
    auto fold = [&]() -> int {
      if (fused)
        hipLaunchKernelGGL(fh::k_fold_pending, dim3(gb), dim3(256), 0, pl->stream, nv, Ax, xc, p, pl->st, true);
      else
        hipLaunchKernelGGL(fh::k_fold_correction, dim3(gb), dim3(256), 0, pl->stream, nv, Ax, xc, true);
      HIPCHK(hipGetLastError());
      return run_project_add(pl, Ax, x);
    };
    auto update = [&]() -> int {
      RETCHK(fold());
      return enqueue_true_residual<float>(pl, b, x, Ax, r, false, true);
    };
    // After the first batch (which establishes that the fused kernels cover the plan), a batch of
    // fused iterations is captured once as a HIP graph (kernel nodes only) and replayed: one
    // launch per batch instead of seven per iteration. The graph is rebuilt when a buffer moves.
    int batch = 0;
    pl->last_graph_batches = 0;
    BatchPredictor pred;
    int len = per_check;
    // after a synchronisation: the next threshold is the stop test or the reliable update
    // (nr < delta * rmax, rmax the last true residual: after a pause, about the current norm)
    auto next_len = [&]() {
      const double rmax = s.pause ? std::sqrt(s.rr) : s.rmax;
      len = pred.next(s, std::max(s.tol, s.delta * rmax), per_check);
    };
    for (;; ++batch) {
      const void* key[5] = {xc, r, p, pl->coef32, pl->spec};
      if (fused && batch > 0 && pl->use_graphs && !pl->it_graph_off) {
        if (std::memcmp(key, pl->it_key, sizeof(key)) != 0) drop_graphs(pl);
        hipGraphExec_t& ex = pl->it_exec[len];
        if (!ex) {
          hipGraph_t graph = nullptr;
          bool all = true;
          int rc = FH_OK;
          if (hipStreamBeginCapture(pl->stream, hipStreamCaptureModeRelaxed) != hipSuccess) all = false;
          for (int k = 0; all && k < len; ++k) {
            bool ok = false;
            rc = enqueue_fused_iteration(pl, xc, r, p, pl->coef32, &ok);
            if (rc != FH_OK || !ok) all = false;
          }
          if (all) enqueue_flush(pl);
          const hipError_t ec = hipStreamEndCapture(pl->stream, &graph);
          if (all && rc == FH_OK && ec == hipSuccess && graph &&
              hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0) == hipSuccess) {
            std::memcpy(pl->it_key, key, sizeof(key));
          } else {
            ex = nullptr;
            pl->it_graph_off = true;  // capture not available here: enqueue the kernels directly
          }
          if (graph) (void)hipGraphDestroy(graph);
          (void)hipGetLastError();
        }
        if (ex) {
          HIPCHK(hipGraphLaunch(ex, pl->stream));
          ++pl->last_graph_batches;
          HIPCHK(hipMemcpyAsync(&s, pl->st, sizeof(s), hipMemcpyDeviceToHost, pl->stream));
          HIPCHK(hipStreamSynchronize(pl->stream));
          if (s.done) break;
          next_len();
          if (s.pause) RETCHK(update());
          continue;
        }
      }
      for (int k = 0; k < len; ++k) {
        if (fused) {
          bool ok = false;
          RETCHK(enqueue_fused_iteration(pl, xc, r, p, pl->coef32, &ok));
          if (ok) continue;
          fused = false;  // not covered: nothing was enqueued, use the general iteration
          pl->cg_defer = false;
        }
        RETCHK(enqueue_iteration<float>(pl, xc, r, p, Ap, pl->coef32));
      }
      if (fused) enqueue_flush(pl);
      HIPCHK(hipMemcpyAsync(&s, pl->st, sizeof(s), hipMemcpyDeviceToHost, pl->stream));
      HIPCHK(hipStreamSynchronize(pl->stream));
      if (s.done) break;
      next_len();
      if (s.pause) RETCHK(update());
    }
    // the correction since the last true residual (none when the solve stopped at one): x0 keeps
    // any part outside the range of G, as in the double-precision solve
    if (!(s.xfresh && !s.pend)) RETCHK(fold());
  }
  HIPCHK(hipEventRecord(pl->ev1, pl->stream));
  HIPCHK(hipMemcpyAsync(&s, pl->st, sizeof(s), hipMemcpyDeviceToHost, pl->stream));
  HIPCHK(hipStreamSynchronize(pl->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
  pl->last_solve_ms = ms;
  pl->last_updates = s.nupdates;
  if (kit) *kit = s.kit;
  if (norm_res) *norm_res = (s.kit == 0) ? 0.0 : std::sqrt(s.rr);
  if (cg_status) *cg_status = s.status;
  if (res_hist && hist_len > 0) {
    const int cnt = std::min(hist_len, s.kit + 1);
    HIPCHK(hipMemcpy(res_hist, pl->hist, (size_t)cnt * sizeof(double), hipMemcpyDeviceToHost));
  }
  return FH_OK;
}

int fh_energy_matrix_noalloc(long long nvox, int kind, const void* coef, const void* const* fields, int nf,
                             double* AH) {
  RETCHK(ensure_device());
  if (!coef || !fields || !AH) return fail(FH_ERR_ARG, "null pointer argument");
  if (nvox < 1) return fail(FH_ERR_ARG, "nvox must be positive");
  if (kind != FH_COEF_ISO && kind != FH_COEF_LAME && kind != FH_COEF_FULL) return fail(FH_ERR_ARG, "bad coefficient kind");
  if (nf != 2 && nf != 3 && nf != 6)
    return fail(FH_ERR_ARG, "fh_energy_matrix_noalloc: nf must be 2, 3 or 6 (the scratch holds at most 6 fields)");
  if (kind == FH_COEF_LAME && nf == 2) return fail(FH_ERR_ARG, "a Lame coefficient needs 3 or 6 Mandel components");
  for (int i = 0; i < nf; ++i)
    if (!fields[i]) return fail(FH_ERR_ARG, "null field pointer");
  const size_t need = (size_t)2 * nf * nf * energy_blocks(nvox) + (size_t)nf * nf;
  if (need > kNoAllocDoubles) return fail(FH_ERR_ARG, "fh_energy_matrix_noalloc: scratch too small");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices || !g_scratch_dev[dev])
    return fail(FH_ERR_STATE, "no scratch address for the current device (fh_init selects device 0)");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  double* part = g_scratch_dev[dev];
  double* res = part + (kNoAllocDoubles - 64);
  // the kernels write every partial sum they read: nothing has to start at zero
  RETCHK(enqueue_energy(nullptr, (const double*)coef, kind, fields, nf, nvox, part, res));
  HIPCHK(hipMemcpy(AH, res, (size_t)nf * nf * sizeof(double), hipMemcpyDeviceToHost));
  return FH_OK;
}

int fh_energy_matrix(fh_plan* pl, const void* const* fields, int nf, double* AH) {
  RETCHK(check_plan(pl));
  if (!fields || !AH) return fail(FH_ERR_ARG, "null pointer argument");
  if (nf != pl->ncomp) return fail(FH_ERR_ARG, "need one field per unit load (nf == ncomp)");
  if (!pl->coef_kind) return fail(FH_ERR_STATE, "fh_set_coeff has not been called");
  for (int i = 0; i < nf; ++i)
    if (!fields[i]) return fail(FH_ERR_ARG, "null field pointer");
  return energy_impl(pl, pl->coef64, pl->coef_kind, fields, nf, AH);
}

int fh_energy_gradient(fh_plan* pl, const void* const* fields, int nf, const double* W, const void* chain,
                       void* grad) {
  RETCHK(check_plan(pl));
  if (!fields || !W || !grad) return fail(FH_ERR_ARG, "null pointer argument");
  if (nf != pl->ncomp) return fail(FH_ERR_ARG, "need one field per unit load (nf == ncomp)");
  if (pl->coef_kind != FH_COEF_ISO && pl->coef_kind != FH_COEF_LAME)
    return fail(FH_ERR_STATE, "the gradient needs an iso or Lame coefficient");
  for (int i = 0; i < nf; ++i)
    if (!fields[i]) return fail(FH_ERR_ARG, "null field pointer");
  fh::FieldPtrs F{};
  for (int i = 0; i < nf; ++i) F.p[i] = (const double*)fields[i];
  fh::Mat6 Wm{};
  for (int k = 0; k < nf * nf; ++k) Wm.w[k] = W[k];
  const dim3 grid(grid_blocks(pl->n, 256, 4096)), block(256);
  const double inv_n = 1.0 / (double)pl->n;
  const double* ch = (const double*)chain;
  double* g = (double*)grad;
  const bool iso = pl->coef_kind == FH_COEF_ISO;
#define FH_EGRAD(NC)                                                                                            \
  do {                                                                                                          \
    if (iso)                                                                                                    \
      hipLaunchKernelGGL((fh::k_energy_grad<NC, fh::COEF_ISO>), grid, block, 0, pl->stream, F, pl->n, Wm, inv_n, \
                         ch, g);                                                                                \
    else                                                                                                        \
      hipLaunchKernelGGL((fh::k_energy_grad<NC, fh::COEF_LAME>), grid, block, 0, pl->stream, F, pl->n, Wm,      \
                         inv_n, ch, g);                                                                         \
  } while (0)
  if (nf == 2)
    FH_EGRAD(2);
  else if (nf == 3)
    FH_EGRAD(3);
  else if (nf == 6)
    FH_EGRAD(6);
  else
    return fail(FH_ERR_ARG, "unsupported number of components");
#undef FH_EGRAD
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pl->stream));
  return FH_OK;
}

int fh_field_gradient(fh_plan* pl, const void* const* fields, const void* const* adjoints, int nf,
                      const void* chain, void* grad) {
  RETCHK(check_plan(pl));
  if (!fields || !adjoints || !grad || nf < 1 || nf > 6) return fail(FH_ERR_ARG, "bad field gradient arguments");
  if (pl->coef_kind != FH_COEF_ISO) return fail(FH_ERR_STATE, "the field gradient needs an iso coefficient");
  for (int i = 0; i < nf; ++i)
    if (!fields[i] || !adjoints[i]) return fail(FH_ERR_ARG, "null field pointer");
  fh::FieldPtrs F{}, Lm{};
  for (int i = 0; i < nf; ++i) {
    F.p[i] = (const double*)fields[i];
    Lm.p[i] = (const double*)adjoints[i];
  }
  const dim3 grid(grid_blocks(pl->n, 256, 4096)), block(256);
  const double* ch = (const double*)chain;
  double* g = (double*)grad;
  switch (pl->ncomp) {
    case 2: hipLaunchKernelGGL((fh::k_field_grad<2>), grid, block, 0, pl->stream, F, Lm, nf, pl->n, ch, g); break;
    case 3: hipLaunchKernelGGL((fh::k_field_grad<3>), grid, block, 0, pl->stream, F, Lm, nf, pl->n, ch, g); break;
    default: return fail(FH_ERR_ARG, "unsupported number of components");
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pl->stream));
  return FH_OK;
}

int fh_axpby(int dtype, size_t n, double a, const void* x, double b, void* y) {
  RETCHK(check_dtype(dtype));
  RETCHK(ensure_device());
  if (n == 0) return FH_OK;
  if (!x || !y) return fail(FH_ERR_ARG, "null pointer argument");
  const dim3 grid(grid_blocks((long long)n)), block(256);
  if (dtype == FH_F64)
    hipLaunchKernelGGL(fh::k_axpby<double>, grid, block, 0, 0, (long long)n, a, (const double*)x, b, (double*)y);
  else
    hipLaunchKernelGGL(fh::k_axpby<float>, grid, block, 0, 0, (long long)n, a, (const float*)x, b, (float*)y);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return FH_OK;
}

int fh_dot(int dtype, size_t n, const void* x, const void* y, double* out) {
  RETCHK(check_dtype(dtype));
  RETCHK(ensure_device());
  if (!out) return fail(FH_ERR_ARG, "null pointer argument");
  if (n == 0) {
    *out = 0.0;
    return FH_OK;
  }
  if (!x || !y) return fail(FH_ERR_ARG, "null pointer argument");
  double* part = nullptr;
  RETCHK(scratch_parts(&part));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  const int gb = std::min(grid_blocks((long long)n), kScratchParts);
  if (dtype == FH_F64)
    hipLaunchKernelGGL(fh::k_dot_partial<double>, dim3(gb), dim3(256), 0, 0, (long long)n, (const double*)x,
                       (const double*)y, part);
  else
    hipLaunchKernelGGL(fh::k_dot_partial<float>, dim3(gb), dim3(256), 0, 0, (long long)n, (const float*)x,
                       (const float*)y, part);
  hipLaunchKernelGGL(fh::k_sum_partials, dim3(1), dim3(256), 0, 0, part, gb, 1.0, part + kScratchParts);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(out, part + kScratchParts, sizeof(double), hipMemcpyDeviceToHost));
  return FH_OK;
}

int fh_add_constant(int dtype, size_t nvox, int ncomp, const double* E, void* y) {
  RETCHK(check_dtype(dtype));
  RETCHK(ensure_device());
  if (!E || !y || ncomp < 1 || ncomp > 36) return fail(FH_ERR_ARG, "bad arguments (1 <= ncomp <= 36)");
  fh::Mat6 Ev{};
  for (int c = 0; c < ncomp; ++c) Ev.w[c] = E[c];
  const dim3 grid(grid_blocks((long long)nvox)), block(256);
  if (dtype == FH_F64)
    hipLaunchKernelGGL(fh::k_add_const<double>, grid, block, 0, 0, (long long)nvox, ncomp, Ev, (double*)y);
  else
    hipLaunchKernelGGL(fh::k_add_const<float>, grid, block, 0, 0, (long long)nvox, ncomp, Ev, (float*)y);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return FH_OK;
}

// ---- spectrum forms and grid changes (fh_forms.hpp) ------------------------------------------

namespace {

// 3-D view of a grid of dimension 1..3 (leading axes of length 1); false on bad sizes
bool view3(int dim, const int* N, fh::forms::Grid3& g) {
  if (dim < 1 || dim > 3 || !N) return false;
  int v[3] = {1, 1, 1};
  for (int a = 0; a < dim; ++a) {
    if (N[a] < 1) return false;
    v[3 - dim + a] = N[a];
  }
  if ((long long)v[0] * v[1] * v[2] > (1LL << 31)) return false;
  g = {v[0], v[1], v[2], v[2] / 2 + 1};
  return true;
}

int forms_done(hipError_t e, const char* what) {
  if (e != hipSuccess) return fail(FH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  HIPCHK(hipDeviceSynchronize());
  return FH_OK;
}

}  // namespace

int fh_spectrum_half_to_full(int dtype, int dim, const int* N, long long count, const void* half, void* full,
                             double scale, int centred) {
  RETCHK(check_dtype(dtype));
  fh::forms::Grid3 g;
  if (!view3(dim, N, g) || count < 0) return fail(FH_ERR_ARG, "fh_spectrum_half_to_full: bad grid or count");
  if (count == 0) return FH_OK;
  if (!half || !full || half == full) return fail(FH_ERR_ARG, "fh_spectrum_half_to_full: bad pointers");
  RETCHK(ensure_device());
  return forms_done(fh::forms::half_to_full(dtype == FH_F64, g, count, half, full, scale, centred ? 1 : 0),
                    "fh_spectrum_half_to_full");
}

int fh_spectrum_full_to_half(int dtype, int dim, const int* N, long long count, const void* full, void* half,
                             double scale, int centred, int hermitian) {
  RETCHK(check_dtype(dtype));
  fh::forms::Grid3 g;
  if (!view3(dim, N, g) || count < 0) return fail(FH_ERR_ARG, "fh_spectrum_full_to_half: bad grid or count");
  if (count == 0) return FH_OK;
  if (!half || !full || half == full) return fail(FH_ERR_ARG, "fh_spectrum_full_to_half: bad pointers");
  RETCHK(ensure_device());
  return forms_done(
      fh::forms::full_to_half(dtype == FH_F64, g, count, full, half, scale, centred ? 1 : 0, hermitian ? 1 : 0),
      "fh_spectrum_full_to_half");
}

int fh_fftshift(int elem_bytes, int dim, const int* N, long long count, const void* in, void* out, int inverse) {
  fh::forms::Grid3 g;
  if (!view3(dim, N, g) || count < 0) return fail(FH_ERR_ARG, "fh_fftshift: bad grid or count");
  if (elem_bytes != 4 && elem_bytes != 8 && elem_bytes != 16) return fail(FH_ERR_ARG, "fh_fftshift: elem_bytes");
  if (count == 0) return FH_OK;
  if (!in || !out || in == out) return fail(FH_ERR_ARG, "fh_fftshift: bad pointers");
  RETCHK(ensure_device());
  return forms_done(fh::forms::shift(g, count, elem_bytes / 4, in, out, inverse ? 1 : 0), "fh_fftshift");
}

int fh_spectrum_resize(int dtype, int dim, const int* N, const int* M, long long count, const void* in, void* out,
                       int enlarge) {
  RETCHK(check_dtype(dtype));
  fh::forms::Grid3 a, b;
  if (!view3(dim, N, a) || !view3(dim, M, b) || count < 0)
    return fail(FH_ERR_ARG, "fh_spectrum_resize: bad grids or count");
  for (int ax = 0; ax < dim; ++ax) {
    const bool ok = enlarge ? M[ax] >= N[ax] + (N[ax] % 2 == 0 ? 1 : 0) : M[ax] <= N[ax];
    if (!ok)
      return fail(FH_ERR_ARG, enlarge ? "fh_spectrum_resize: enlarging needs M >= N + 1 (even N) or M >= N (odd N)"
                                      : "fh_spectrum_resize: decreasing needs M <= N");
  }
  if (count == 0) return FH_OK;
  if (!in || !out || in == out) return fail(FH_ERR_ARG, "fh_spectrum_resize: bad pointers");
  RETCHK(ensure_device());
  return forms_done(fh::forms::resize_centred(dtype == FH_F64, a, b, count, in, out, enlarge ? 1 : 0),
                    "fh_spectrum_resize");
}

int fh_project_spectrum(int dtype, int dim, const int* N, const double* Y, int physics, int which, int form,
                        int null_nyq, const void* in, void* out) {
  RETCHK(check_dtype(dtype));
  fh::forms::Grid3 v;
  if (dim < 2 || !view3(dim, N, v)) return fail(FH_ERR_ARG, "fh_project_spectrum: dim must be 2 or 3");
  if (physics != FH_SCALAR && physics != FH_ELASTICITY) return fail(FH_ERR_ARG, "fh_project_spectrum: physics");
  if (which != 1 && which != 2) return fail(FH_ERR_ARG, "fh_project_spectrum: which must be 1 or 2");
  if (form != FH_FORM_R && form != FH_FORM_FULL && form != FH_FORM_CENTRED)
    return fail(FH_ERR_ARG, "fh_project_spectrum: form");
  if (!Y) return fail(FH_ERR_ARG, "fh_project_spectrum: null Y");
  for (int a = 0; a < dim; ++a)
    if (!(Y[a] > 0.0)) return fail(FH_ERR_ARG, "fh_project_spectrum: Y must be positive");
  if (!in || !out || in == out) return fail(FH_ERR_ARG, "fh_project_spectrum: bad pointers");
  RETCHK(ensure_device());
  GParams g{};
  g.dim = dim;
  g.physics = physics;
  g.dual = which == 2;
  g.ncomp = physics == FH_SCALAR ? dim : dim * (dim + 1) / 2;
  g.n0 = v.n0;
  g.n1 = v.n1;
  g.n2 = v.n2;
  for (int a = 0; a < 3; ++a) g.iY[a] = 0.0;
  for (int a = 0; a < dim; ++a) g.iY[3 - dim + a] = 1.0 / Y[a];
  g.scale = 1.0;
  g.banded = 0;
  return forms_done(fh::forms::project_form(dtype == FH_F64, g, form, null_nyq ? 1 : 0, in, out),
                    "fh_project_spectrum");
}

int fh_spectrum_grad(int dtype, int dim, const int* N, const double* Y, const void* in, void* out) {
  RETCHK(check_dtype(dtype));
  fh::forms::Grid3 v;
  if (!view3(dim, N, v)) return fail(FH_ERR_ARG, "fh_spectrum_grad: bad grid");
  if (!Y) return fail(FH_ERR_ARG, "fh_spectrum_grad: null Y");
  double tp[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < dim; ++a) {
    if (!(Y[a] > 0.0)) return fail(FH_ERR_ARG, "fh_spectrum_grad: Y must be positive");
    tp[3 - dim + a] = (double)(2.0L * 3.14159265358979323846264338327950288L / (long double)Y[a]);
  }
  if (!in || !out || in == out) return fail(FH_ERR_ARG, "fh_spectrum_grad: bad pointers");
  RETCHK(ensure_device());
  return forms_done(fh::forms::spectrum_grad(dtype == FH_F64, v, dim, tp, in, out), "fh_spectrum_grad");
}

int fh_pointwise_product(int dtype, int is_complex, int mode, int A, int B, long long nvox, const void* X,
                         const void* Y, void* out) {
  RETCHK(check_dtype(dtype));
  if ((mode != 0 && mode != 1) || A < 1 || A > 4096 || (mode == 0 && (B < 1 || B > 4096)) || nvox < 0)
    return fail(FH_ERR_ARG, "fh_pointwise_product: bad mode or sizes");
  if (nvox == 0) return FH_OK;
  if (!X || !Y || !out || out == X || out == Y) return fail(FH_ERR_ARG, "fh_pointwise_product: bad pointers");
  RETCHK(ensure_device());
  return forms_done(fh::forms::pointwise_product(dtype == FH_F64, is_complex ? 1 : 0, mode, A, mode == 0 ? B : 1,
                                                 nvox, X, Y, out),
                    "fh_pointwise_product");
}

int fh_pointwise_inverse(int dtype, int k, long long nvox, const void* in, void* out) {
  RETCHK(check_dtype(dtype));
  if (k < 1 || k > 6 || nvox < 0) return fail(FH_ERR_ARG, "fh_pointwise_inverse: 1 <= k <= 6, nvox >= 0");
  if (nvox == 0) return FH_OK;
  if (!in || !out || in == out) return fail(FH_ERR_ARG, "fh_pointwise_inverse: bad pointers");
  RETCHK(ensure_device());
  return forms_done(fh::forms::pointwise_inverse(dtype == FH_F64, k, nvox, in, out), "fh_pointwise_inverse");
}

int fh_last_solve_updates(const fh_plan* pl, int* n_updates) {
  if (!pl || !n_updates) return fail(FH_ERR_ARG, "null argument");
  *n_updates = pl->last_updates;
  return FH_OK;
}

int fh_synchronize(void) {
  RETCHK(ensure_device());
  HIPCHK(hipDeviceSynchronize());
  return FH_OK;
}

int fh_matvec_profile(fh_plan* pl, int dtype, const void* x, void* y, int nrep, int max_kernels,
                      double* kernel_ms, double* kernel_bytes, int* nkernels) {
  RETCHK(check_plan(pl));
  RETCHK(check_dtype(dtype));
  if (!x || !y || !kernel_ms || !kernel_bytes || !nkernels || nrep < 1) return fail(FH_ERR_ARG, "bad arguments");
  if (!pl->coef_kind) return fail(FH_ERR_STATE, "fh_set_coeff has not been called");
  const int nk = pl->n0 > 1 ? 5 : 3;
  if (max_kernels < nk) return fail(FH_ERR_ARG, "max_kernels too small");
  // Each launch of the fixed-length kernels records its own start and stop (hipExtLaunchKernel
  // events: the dispatch's timestamps, no extra packets between kernels). nrep matvecs run
  // back to back; pass k's duration is averaged over them.
  std::vector<hipEvent_t> evs(2 * nk);
  for (auto& e : evs) HIPCHK(hipEventCreate(&e));
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::vector<double> acc(nk, 0.0);
  const int nc = pl->ncomp;
  const size_t sr = dtype == FH_F64 ? 8 : 4, sc = 2 * sr;
  const double cvals = pl->coef_kind == FH_COEF_ISO ? 1 : (pl->coef_kind == FH_COEF_LAME ? 2 : nc * nc);
  const double nb_real = (double)nc * pl->n * sr, nb_spec = (double)nc * pl->nh * sc;
  int rc = FH_OK;
  auto pass = [&](auto tag, int k) -> int {
    using T = decltype(tag);
    cpx<T>* spec = reinterpret_cast<cpx<T>*>(pl->spec);
    const T* coef = sizeof(T) == 8 ? (const T*)pl->coef64 : (const T*)pl->coef32;
    const GParams g = make_gparams(pl, pl->op_scale / (double)pl->n);
    if (k == 0) return launch_rows_fwd<T>(pl, nc, (const T*)x, coef, pl->coef_kind, spec, nullptr);
    if (k == nk - 1) return launch_rows_inv<T>(pl, nc, spec, (T*)y, T(1), nullptr, nullptr, nullptr, nullptr);
    if (pl->n0 > 1) {
      if (k == 1) return launch_cols<T>(pl, 0, 1, nc, spec, g, nullptr);
      if (k == 2) return launch_cols<T>(pl, 2, 0, nc, spec, g, nullptr);
      return launch_cols<T>(pl, 1, 1, nc, spec, g, nullptr);
    }
    return launch_cols<T>(pl, 2, 1, nc, spec, g, nullptr);
  };
  for (int it = -1; it < nrep && rc == FH_OK; ++it) {  // it = -1: warm-up
    for (int k = 0; k < nk && rc == FH_OK; ++k) {
      pl->prof_t0 = evs[2 * k];
      pl->prof_t1 = evs[2 * k + 1];
      rc = dtype == FH_F64 ? pass(double(0), k) : pass(float(0), k);
      pl->prof_t0 = pl->prof_t1 = nullptr;
    }
    if (rc == FH_OK && hipStreamSynchronize(pl->stream) != hipSuccess) rc = fail(FH_ERR_HIP, "hipStreamSynchronize");
    for (int k = 0; k < nk && rc == FH_OK && it >= 0; ++k) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, evs[2 * k], evs[2 * k + 1]) != hipSuccess)
        rc = fail(FH_ERR_UNSUPPORTED, "kernel timing needs the fixed-length kernels");
      acc[k] += (double)ms;
    }
  }
  for (auto& e : evs) (void)hipEventDestroy(e);
  (void)ev;
  if (rc != FH_OK) return rc;
  for (int k = 0; k < nk; ++k) {
    kernel_ms[k] = acc[k] / nrep;
    if (k == 0)
      kernel_bytes[k] = nb_real + cvals * pl->n * sr + nb_spec;
    else if (k == nk - 1)
      kernel_bytes[k] = nb_spec + nb_real;
    else
      kernel_bytes[k] = 2.0 * nb_spec;
  }
  *nkernels = nk;
  return FH_OK;
}

int fh_plan_set_schedule(fh_plan* pl, int dtype, int proj_variant, int cols_variant, int rows_variant) {
  RETCHK(check_plan(pl));
  RETCHK(check_dtype(dtype));
  if (proj_variant < 0 || proj_variant >= fh::kProjVariants || cols_variant < 0 || cols_variant >= fh::kColsVariants ||
      rows_variant < 0 || rows_variant >= fh::kRowsVariants)
    return fail(FH_ERR_ARG, "schedule variant out of range");
  const int i = dtype == FH_F64 ? 0 : 1;
  if (pl->proj_var[i] != proj_variant || pl->cols_var[i] != cols_variant || pl->rows_var[i] != rows_variant)
    drop_graphs(pl);
  pl->proj_var[i] = proj_variant;
  pl->cols_var[i] = cols_variant;
  pl->rows_var[i] = rows_variant;
  return FH_OK;
}

int fh_plan_get_schedule(const fh_plan* pl, int dtype, int* proj_variant, int* cols_variant, int* rows_variant) {
  RETCHK(check_plan(pl));
  RETCHK(check_dtype(dtype));
  if (!proj_variant || !cols_variant || !rows_variant) return fail(FH_ERR_ARG, "null pointer argument");
  const int i = dtype == FH_F64 ? 0 : 1;
  *proj_variant = pl->proj_var[i];
  *cols_variant = pl->cols_var[i];
  *rows_variant = pl->rows_var[i];
  return FH_OK;
}

// Average time (ms) of the two row passes of the fused fp32 CG iteration (forward with the direction
// update and the pending step, inverse with the residual update), nrep launches each after one
// warm-up, on the plan's current schedule: what the mixed solve runs, where the matvec's row passes
// move other vectors. The vectors are the mixed solve's own (plan work memory, p and r copied from
// x and y); the CG state is a benign one (no halt, one pending step) that the next solve replaces.
// *ms < 0 where the fused kernels do not cover the plan.
static int time_fused_rows(fh_plan* pl, const float* x, const float* y, int nrep, double* ms) {
  *ms = -1.0;
  if (pl->generic || (pl->coef_kind != fh::COEF_ISO && pl->coef_kind != fh::COEF_LAME)) return FH_OK;
  const long long nv = (long long)pl->ncomp * pl->n;
  const long long nvp = (nv + 63) / 64 * 64;  // the mixed solve's layout of its work vectors
  RETCHK(ensure_work(pl, 4 * (size_t)nvp * sizeof(float) + (size_t)nvp * sizeof(double)));
  float* r = (float*)((double*)pl->work + nvp);
  float* p = r + nvp;
  float* xc = p + 2 * nvp;
  HIPCHK(hipMemcpyAsync(p, x, (size_t)nv * sizeof(float), hipMemcpyDeviceToDevice, pl->stream));
  HIPCHK(hipMemcpyAsync(r, y, (size_t)nv * sizeof(float), hipMemcpyDeviceToDevice, pl->stream));
  HIPCHK(hipMemsetAsync(xc, 0, (size_t)nv * sizeof(float), pl->stream));
  CGState s;
  std::memset(&s, 0, sizeof(s));
  s.rr = 1.0;
  s.alpha = 1e-3;
  s.beta = 0.5;
  s.inv_nvox = 1.0 / (double)pl->n;
  s.maxiter = 1 << 30;
  s.mixed = 1;
  s.pend = 1;
  HIPCHK(hipMemcpyAsync(pl->st, &s, sizeof(s), hipMemcpyHostToDevice, pl->stream));
  hipEvent_t ev[4] = {};
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  const int nc = pl->ncomp;
  cpx<float>* spec = reinterpret_cast<cpx<float>*>(pl->spec);
  int rc = FH_OK;
  double acc = 0.0;
  bool covered = true;
  for (int it = -1; it < nrep && rc == FH_OK && covered; ++it) {
    fh::LaunchCtx cf = make_ctx<float>(pl);
    cf.t0 = ev[0];
    cf.t1 = ev[1];
    int nb = 0, nb2 = 0;
    int k = fh::fixed_rows_fwd<float>(cf, nc, p, pl->coef32, pl->coef_kind, spec, &pl->st->halt, r, pl->st, xc,
                                      pl->partial, pl->partial_cap, &nb);
    if (k == 0) {
      covered = false;
      break;
    }
    rc = fixed_result(k, "rows forward (fused, timed)");
    if (rc != FH_OK) break;
    fh::LaunchCtx ci = make_ctx<float>(pl);
    ci.t0 = ev[2];
    ci.t1 = ev[3];
    k = fh::fixed_rows_inv<float>(ci, nc, spec, r, 1.0f, 2, nullptr, pl->partial, pl->partial_cap, &pl->st->halt,
                                  &nb2, pl->st);
    if (k == 0) {
      covered = false;
      break;
    }
    rc = fixed_result(k, "rows inverse (fused, timed)");
    if (rc != FH_OK) break;
    if (hipStreamSynchronize(pl->stream) != hipSuccess) {
      rc = fail(FH_ERR_HIP, "hipStreamSynchronize");
      break;
    }
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, ev[0], ev[1]) != hipSuccess || hipEventElapsedTime(&b, ev[2], ev[3]) != hipSuccess) {
      rc = fail(FH_ERR_HIP, "hipEventElapsedTime");
      break;
    }
    if (it >= 0) acc += (double)a + (double)b;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc != FH_OK) return rc;
  if (covered) *ms = acc / nrep;
  return FH_OK;
}

int fh_plan_autotune(fh_plan* pl, int dtype, const void* x, void* y, int nrep, double* proj_ms, double* cols_ms,
                     double* rows_ms, int* proj_variant, int* cols_variant, int* rows_variant) {
  RETCHK(check_plan(pl));
  RETCHK(check_dtype(dtype));
  if (!x || !y || nrep < 1) return fail(FH_ERR_ARG, "bad arguments");
  const int i = dtype == FH_F64 ? 0 : 1;
  const int nk = pl->n0 > 1 ? 5 : 3;
  const int kproj = pl->n0 > 1 ? 2 : 1;
  double kms[8], kby[8];
  int nkr = 0;
  const int keep_p = pl->proj_var[i], keep_c = pl->cols_var[i], keep_r = pl->rows_var[i];
  double tp[fh::kProjVariants], tc[fh::kColsVariants], tr[fh::kRowsVariants];
  // each variant of one pass with the other passes at their current choice; the faster wins.
  // kRounds interleaved rounds over the variants, each variant keeping its best round, so a
  // slow spell of the device does not fall on one variant only
  constexpr int kRounds = 3;
  int rc = FH_OK;
  for (int v = 0; v < fh::kProjVariants; ++v) tp[v] = 1e300;
  for (int r = 0; r < kRounds; ++r) {
    for (int v = 0; v < fh::kProjVariants && rc == FH_OK; ++v) {
      pl->proj_var[i] = v;
      rc = fh_matvec_profile(pl, dtype, x, y, nrep, 8, kms, kby, &nkr);
      if (rc == FH_OK) tp[v] = std::min(tp[v], kms[kproj]);
    }
  }
  int bp = 0;
  for (int v = 1; v < fh::kProjVariants; ++v)
    if (tp[v] < tp[bp]) bp = v;
  pl->proj_var[i] = bp;
  int bc = keep_c;
  for (int v = 0; v < fh::kColsVariants; ++v) tc[v] = nk == 3 ? 0.0 : 1e300;
  for (int r = 0; r < kRounds && nk == 5; ++r) {
    for (int v = 0; v < fh::kColsVariants && rc == FH_OK; ++v) {
      pl->cols_var[i] = v;
      rc = fh_matvec_profile(pl, dtype, x, y, nrep, 8, kms, kby, &nkr);
      if (rc == FH_OK) tc[v] = std::min(tc[v], kms[1] + kms[3]);
    }
  }
  if (nk == 5) {
    bc = 0;
    for (int v = 1; v < fh::kColsVariants; ++v)
      if (tc[v] < tc[bc]) bc = v;
  }
  pl->cols_var[i] = bc;
  // contiguous-axis passes: forward and inverse together; in single precision, where the mixed
  // solve runs the fused iteration, its row passes (which also stream r, p and the correction)
  for (int v = 0; v < fh::kRowsVariants; ++v) tr[v] = 1e300;
  bool fused_rows = dtype == FH_F32 && pl->fused_cg && pl->op_scale == 1.0;
  for (int r = 0; r < kRounds; ++r) {
    for (int v = 0; v < fh::kRowsVariants && rc == FH_OK; ++v) {
      pl->rows_var[i] = v;
      if (fused_rows) {
        double fms = -1.0;
        rc = time_fused_rows(pl, (const float*)x, (const float*)y, nrep, &fms);
        if (rc == FH_OK && fms >= 0.0) {
          tr[v] = std::min(tr[v], fms);
          continue;
        }
        // not covered: every variant by the matvec's row passes
        fused_rows = false;
        for (int u = 0; u < fh::kRowsVariants; ++u) tr[u] = 1e300;
        r = -1;
        break;
      }
      rc = fh_matvec_profile(pl, dtype, x, y, nrep, 8, kms, kby, &nkr);
      if (rc == FH_OK) tr[v] = std::min(tr[v], kms[0] + kms[nk - 1]);
    }
  }
  int br = 0;
  for (int v = 1; v < fh::kRowsVariants; ++v)
    if (tr[v] < tr[br]) br = v;
  pl->proj_var[i] = keep_p;
  pl->cols_var[i] = keep_c;
  pl->rows_var[i] = keep_r;
  if (rc != FH_OK) return rc;
  RETCHK(fh_plan_set_schedule(pl, dtype, bp, bc, br));
  if (proj_ms)
    for (int v = 0; v < fh::kProjVariants; ++v) proj_ms[v] = tp[v];
  if (cols_ms)
    for (int v = 0; v < fh::kColsVariants; ++v) cols_ms[v] = tc[v];
  if (rows_ms)
    for (int v = 0; v < fh::kRowsVariants; ++v) rows_ms[v] = tr[v];
  if (proj_variant) *proj_variant = bp;
  if (cols_variant) *cols_variant = bc;
  if (rows_variant) *rows_variant = br;
  return FH_OK;
}

int fh_last_timing(const fh_plan* pl, double* matvec_ms, double* solve_ms) {
  if (!pl) return fail(FH_ERR_ARG, "null plan");
  if (matvec_ms) *matvec_ms = pl->last_matvec_ms;
  if (solve_ms) *solve_ms = pl->last_solve_ms;
  return FH_OK;
}

// ---- batched 2-D scalar solves -------------------------------------------------------------

int fh_batch_destroy(fh_batch* b) {
  if (!b) return FH_OK;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  void* bufs[] = {b->d_smp, b->d_order, b->d_counter, b->d_a64, b->d_a32, b->d_tw64, b->d_tw32, b->d_ws64,
                  b->d_ws32, b->d_energy, b->d_kit, b->d_nupd, b->d_normres, b->d_status, b->d_par};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return FH_OK;
}

int fh_batch_create(int nsamples, const int* dims, const double* Y, fh_batch** out) {
  if (!out || !dims || nsamples <= 0) return fail(FH_ERR_ARG, "fh_batch_create: bad arguments");
  *out = nullptr;
  RETCHK(ensure_device());
  namespace B = fh::batch;
  fh_batch* b = new fh_batch();
  auto bail = [&](int code) {
    fh_batch_destroy(b);
    return code;
  };
  b->S = nsamples;
  b->ntasks = 2 * nsamples;
  b->smp.resize(nsamples);
  long long nvec = 0, spec = 0, nvox = 0;
  for (int s = 0; s < nsamples; ++s) {
    const int n0 = dims[2 * s], n1 = dims[2 * s + 1];
    if (B::len_id(n0) < 0 || B::len_id(n1) < 0)
      return bail(fail(FH_ERR_UNSUPPORTED, "sample " + std::to_string(s) + ": grid " + std::to_string(n0) + "x" +
                                              std::to_string(n1) + " (axis lengths 27, 45, 63, 81, 105, 135)"));
    const double y0 = Y ? Y[2 * s] : 1.0, y1 = Y ? Y[2 * s + 1] : 1.0;
    if (!(y0 > 0.0) || !(y1 > 0.0)) return bail(fail(FH_ERR_ARG, "cell sizes must be positive"));
    B::Sample& sm = b->smp[s];
    sm.n0 = n0;
    sm.n1 = n1;
    sm.off = nvox;
    sm.iY0 = 1.0 / y0;
    sm.iY1 = 1.0 / y1;
    nvox += (long long)n0 * n1;
    nvec = std::max(nvec, 2LL * n0 * n1);
    spec = std::max(spec, B::spec_elems(n0, n1));
  }
  b->nvox = nvox;
  nvec = (nvec + 63) / 64 * 64;
  const long long spec64 = (2 * spec + 63) / 64 * 64;  // doubles
  b->lds = (size_t)B::OFF_SPEC + (size_t)spec * sizeof(cpx<float>);
  if (b->lds > kLdsMax) return bail(fail(FH_ERR_UNSUPPORTED, "half spectrum exceeds LDS"));
  // tasks: largest grids first, primal and dual of a sample adjacent
  std::vector<int> order(b->ntasks);
  for (int t = 0; t < b->ntasks; ++t) order[t] = t;
  b->order_sample = order;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    const B::Sample &sa = b->smp[a >> 1], &sc = b->smp[c >> 1];
    return (long long)sa.n0 * sa.n1 > (long long)sc.n0 * sc.n1;
  });
  b->order_size = order;
  int ncu = 0;
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
  b->nslots = std::max(1, std::min(b->ntasks, ncu));
  // twiddle tables exp(-2 pi i m / L) of every supported length
  std::vector<cpx<double>> t64((size_t)B::NLEN * B::TWS);
  std::vector<cpx<float>> t32((size_t)B::NLEN * B::TWS);
  for (int id = 0; id < B::NLEN; ++id) {
    const int L = B::kLens[id];
    for (int m = 0; m < B::TWS; ++m) {
      const long double ang = -2.0L * 3.14159265358979323846264338327950288L * (long double)(m % L) / (long double)L;
      t64[id * B::TWS + m] = fh::mk((double)cosl(ang), (double)sinl(ang));
      t32[id * B::TWS + m] = fh::mk((float)(double)cosl(ang), (float)(double)sinl(ang));
    }
  }
  const size_t ws64 = (size_t)b->nslots * (6 * nvec + spec64), ws32 = (size_t)b->nslots * 4 * nvec;
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&b->ev0);
  if (e == hipSuccess) e = hipEventCreate(&b->ev1);
  if (e == hipSuccess) e = hipMalloc(&b->d_smp, nsamples * sizeof(B::Sample));
  if (e == hipSuccess) e = hipMalloc(&b->d_order, b->ntasks * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&b->d_counter, 64 * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&b->d_a64, std::max<long long>(nvox, 1) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&b->d_a32, std::max<long long>(nvox, 1) * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&b->d_tw64, t64.size() * sizeof(cpx<double>));
  if (e == hipSuccess) e = hipMalloc(&b->d_tw32, t32.size() * sizeof(cpx<float>));
  if (e == hipSuccess) e = hipMalloc(&b->d_ws64, ws64 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&b->d_ws32, ws32 * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&b->d_energy, (size_t)nsamples * 8 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&b->d_kit, (size_t)nsamples * 4 * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&b->d_nupd, (size_t)nsamples * 4 * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&b->d_normres, (size_t)nsamples * 4 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&b->d_status, (size_t)nsamples * 2 * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&b->d_par, sizeof(B::Params));
  if (e == hipSuccess)
    e = hipMemcpy(b->d_smp, b->smp.data(), nsamples * sizeof(B::Sample), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_order, order.data(), b->ntasks * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_tw64, t64.data(), t64.size() * sizeof(cpx<double>), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(b->d_tw32, t32.data(), t32.size() * sizeof(cpx<float>), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = B::allow_lds(b->lds);
  if (e != hipSuccess) return bail(fail(FH_ERR_HIP, std::string("fh_batch_create: ") + hipGetErrorString(e)));
  B::Args& A = b->args;
  A.smp = b->d_smp;
  A.order = b->d_order;
  A.ntasks = b->ntasks;
  A.counter = b->d_counter;
  A.a64 = b->d_a64;
  A.a32 = b->d_a32;
  A.tw64 = b->d_tw64;
  A.tw32 = b->d_tw32;
  A.ws64 = b->d_ws64;
  A.ws32 = b->d_ws32;
  A.ws64_stride = 6 * nvec + spec64;
  A.ws32_stride = 4 * nvec;
  A.nvec = nvec;
  A.energy = b->d_energy;
  A.kit = b->d_kit;
  A.nupd = b->d_nupd;
  A.normres = b->d_normres;
  A.status = b->d_status;
  A.par = b->d_par;
  *out = b;
  return FH_OK;
}

int fh_batch_set_coeff(fh_batch* b, const double* a) {
  if (!b || !a) return fail(FH_ERR_ARG, "fh_batch_set_coeff: null argument");
  for (long long i = 0; i < b->nvox; ++i)
    if (!(a[i] > 0.0) || !std::isfinite(a[i]))
      return fail(FH_ERR_ARG, "conductivity must be positive and finite (voxel " + std::to_string(i) + ")");
  std::vector<float> a32((size_t)b->nvox);
  for (long long i = 0; i < b->nvox; ++i) a32[i] = (float)a[i];
  HIPCHK(hipMemcpy(b->d_a64, a, (size_t)b->nvox * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->d_a32, a32.data(), (size_t)b->nvox * sizeof(float), hipMemcpyHostToDevice));
  b->coef_set = true;
  return FH_OK;
}

int fh_batch_solve(fh_batch* b, int mode, double tol, int maxiter, double delta, double* upper, double* lower,
                   int* kit, int* status) {
  if (!b) return fail(FH_ERR_ARG, "null batch");
  if (!b->coef_set) return fail(FH_ERR_STATE, "fh_batch_set_coeff has not been called");
  if (mode != FH_SOLVE_FP64 && mode != FH_SOLVE_MIXED) return fail(FH_ERR_ARG, "bad solve mode");
  if (!(tol >= 0.0) || maxiter < 0) return fail(FH_ERR_ARG, "bad tol/maxiter");
  if (mode == FH_SOLVE_MIXED && !(delta > 0.0 && delta < 1.0)) return fail(FH_ERR_ARG, "delta must be in (0,1)");
  fh::batch::Params P;
  P.tol = tol;
  P.delta = delta;
  P.maxiter = maxiter;
  P.mixed = mode == FH_SOLVE_MIXED;
  if (b->variant < 0) RETCHK(fh_batch_autotune(b, mode, delta, 12, nullptr, nullptr, nullptr));
  HIPCHK(hipMemcpyAsync(b->d_par, &P, sizeof(P), hipMemcpyHostToDevice, b->stream));
  if (!b->exec) {
    // one graph: counter reset + the persistent solve kernel; replayed by every call
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = fh::batch::enqueue(b->args, b->nslots, b->lds, b->stream, b->variant);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(FH_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e != hipSuccess ? e : e2));
    }
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  HIPCHK(hipEventRecord(b->ev0, b->stream));
  HIPCHK(hipGraphLaunch(b->exec, b->stream));
  HIPCHK(hipEventRecord(b->ev1, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  const int S = b->S;
  std::vector<double> E((size_t)S * 8);
  std::vector<int> K((size_t)S * 4), St((size_t)S * 2);
  HIPCHK(hipMemcpy(E.data(), b->d_energy, E.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(K.data(), b->d_kit, K.size() * sizeof(int), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(St.data(), b->d_status, St.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (int s = 0; s < S; ++s) {
    const double* ep = &E[(size_t)s * 8];      // primal energy = upper bound
    const double* ed = &E[(size_t)s * 8 + 4];  // dual energy; lower bound = its inverse
    if (upper)
      for (int k = 0; k < 4; ++k) upper[4 * s + k] = ep[k];
    if (lower) {
      const double det = ed[0] * ed[3] - ed[1] * ed[2];
      lower[4 * s + 0] = ed[3] / det;
      lower[4 * s + 1] = -ed[1] / det;
      lower[4 * s + 2] = -ed[2] / det;
      lower[4 * s + 3] = ed[0] / det;
    }
    if (kit)
      for (int k = 0; k < 4; ++k) kit[4 * s + k] = K[(size_t)s * 4 + k];
    if (status) status[s] = St[2 * s] | St[2 * s + 1];
  }
  return FH_OK;
}

static int batch_set_schedule(fh_batch* b, int variant, int order_mode) {
  const std::vector<int>& ord = order_mode ? b->order_sample : b->order_size;
  HIPCHK(hipMemcpy(b->d_order, ord.data(), ord.size() * sizeof(int), hipMemcpyHostToDevice));
  if (variant != b->variant && b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  b->variant = variant;
  b->order_mode = order_mode;
  return FH_OK;
}

int fh_batch_set_schedule(fh_batch* b, int variant, int order_mode) {
  if (!b) return fail(FH_ERR_ARG, "null batch");
  if (variant < 0 || variant >= fh::batch::NVAR || order_mode < 0 || order_mode > 1)
    return fail(FH_ERR_ARG, "bad schedule");
  return batch_set_schedule(b, variant, order_mode);
}

int fh_batch_autotune(fh_batch* b, int mode, double delta, int iters, double* times_ms, int* variant,
                      int* order_mode) {
  if (!b) return fail(FH_ERR_ARG, "null batch");
  if (!b->coef_set) return fail(FH_ERR_STATE, "fh_batch_set_coeff has not been called");
  if (iters < 1) return fail(FH_ERR_ARG, "iters must be positive");
  namespace B = fh::batch;
  // every task runs `iters` CG iterations per load (tolerance 0) with each schedule
  B::Params P;
  P.tol = 0.0;
  P.delta = (delta > 0.0 && delta < 1.0) ? delta : 0.01;
  P.maxiter = iters;
  P.mixed = mode == FH_SOLVE_MIXED;
  HIPCHK(hipMemcpy(b->d_par, &P, sizeof(P), hipMemcpyHostToDevice));
  double best = 1e300;
  int bv = 0, bo = 0;
  for (int o = 0; o < 2; ++o) {
    for (int v = 0; v < B::NVAR; ++v) {
      RETCHK(batch_set_schedule(b, v, o));
      if (o == 0) {  // first launch of each kernel: load its code before timing
        HIPCHK(B::enqueue(b->args, b->nslots, b->lds, b->stream, v));
      }
      HIPCHK(hipEventRecord(b->ev0, b->stream));
      HIPCHK(B::enqueue(b->args, b->nslots, b->lds, b->stream, v));
      HIPCHK(hipEventRecord(b->ev1, b->stream));
      HIPCHK(hipStreamSynchronize(b->stream));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, b->ev0, b->ev1));
      b->tune_ms[o * B::NVAR + v] = ms;
      if (times_ms) times_ms[o * B::NVAR + v] = ms;
      if (ms < best) {
        best = ms;
        bv = v;
        bo = o;
      }
    }
  }
  RETCHK(batch_set_schedule(b, bv, bo));
  if (variant) *variant = bv;
  if (order_mode) *order_mode = bo;
  return FH_OK;
}

int fh_batch_stats(const fh_batch* b, int* n_updates, double* norm_res, double* solve_ms) {
  if (!b) return fail(FH_ERR_ARG, "null batch");
  if (n_updates) HIPCHK(hipMemcpy(n_updates, b->d_nupd, (size_t)b->S * 4 * sizeof(int), hipMemcpyDeviceToHost));
  if (norm_res) HIPCHK(hipMemcpy(norm_res, b->d_normres, (size_t)b->S * 4 * sizeof(double), hipMemcpyDeviceToHost));
  if (solve_ms) *solve_ms = b->last_ms;
  return FH_OK;
}


// ---- Ga post-processing ------------------------------------------------------------------

int fh_ga_create(fh_plan* coarse, const int* grid, fh_ga** out) {
  RETCHK(check_plan(coarse));
  if (!out) return fail(FH_ERR_ARG, "null pointer argument");
  *out = nullptr;
  const int dim = coarse->dim;
  int M[3] = {1, 1, 1};
  for (int a = 0; a < dim; ++a) {
    if (grid) {
      M[a] = grid[a];
      if (M[a] < 2 * coarse->Nu[a] - 1) return fail(FH_ERR_ARG, "quadrature grid smaller than 2N-1");
      continue;
    }
    M[a] = quadrature_length(coarse->Nu[a]);
    if (M[a] < 0)
      return fail(FH_ERR_UNSUPPORTED, "Ga quadrature grid: 2N-1 = " + std::to_string(2 * coarse->Nu[a] - 1) +
                                          " has a prime factor above 13 and N is even");
  }
  fh_plan* fine = nullptr;
  RETCHK(fh_plan_create(dim, M, coarse->Y, coarse->physics, coarse->primaldual, &fine));
  fh_ga* ga = new fh_ga();
  ga->coarse = coarse;
  ga->fine = fine;
  fh::ga::SpecGrid& g = ga->g;
  g.m0 = fine->n0;
  g.m1 = fine->n1;
  g.m2 = fine->n2;
  g.m2h = fine->n2h;
  g.vol = 1.0;
  for (int a = 0; a < dim; ++a) g.vol *= coarse->Y[a];
  const int off = 3 - dim;  // view axis of problem axis 0
  for (int v = 0; v < 3; ++v) {
    const int a = v - off;
    const int N = a >= 0 ? coarse->Nu[a] : 1;
    g.band[v] = N - 1;
    g.nbar[v] = 2 * N - 1;
    g.Y[v] = a >= 0 ? coarse->Y[a] : 1.0;
  }
  g.scale = 1.0;  // spec_to_real is the unnormalised inverse transform
  *out = ga;
  return FH_OK;
}

int fh_ga_destroy(fh_ga* ga) {
  if (!ga) return FH_OK;
  if (ga->coef) (void)hipFree(ga->coef);
  if (ga->fields) (void)hipFree(ga->fields);
  fh_plan_destroy(ga->fine);
  delete ga;
  return FH_OK;
}

int fh_ga_grid(const fh_ga* ga, int* M) {
  if (!ga || !M) return fail(FH_ERR_ARG, "null pointer argument");
  for (int a = 0; a < ga->coarse->dim; ++a) M[a] = ga->fine->Nu[a];
  return FH_OK;
}

int fh_ga_set_inclusions(fh_ga* ga, int nincl, const int* kinds, const double* params, const double* positions,
                         int kind, const double* vals) {
  if (!ga || !kinds || !params || !positions || !vals) return fail(FH_ERR_ARG, "null pointer argument");
  const int ncoef = ga_ncoef(ga, kind);
  if (ncoef < 0) return fail(FH_ERR_ARG, "bad coefficient kind");
  if (nincl < 1 || nincl > 4096) return fail(FH_ERR_ARG, "number of inclusions out of range");
  const int dim = ga->coarse->dim, off = 3 - dim;
  std::vector<fh::ga::Incl> h(nincl);
  for (int ii = 0; ii < nincl; ++ii) {
    fh::ga::Incl& in = h[ii];
    std::memset(&in, 0, sizeof(in));
    in.kind = kinds[ii];
    if (in.kind < fh::ga::INCL_CUBE || in.kind > fh::ga::INCL_OTHERWISE)
      return fail(FH_ERR_ARG, "unknown inclusion kind");
    if (in.kind == fh::ga::INCL_OTHERWISE && ii != nincl - 1)
      return fail(FH_ERR_ARG, "'otherwise' must be the last inclusion");
    for (int v = 0; v < 3; ++v) {
      const int a = v - off;
      in.p[v] = a >= 0 ? params[3 * ii + a] : 1.0;
      in.pos[v] = a >= 0 ? positions[3 * ii + a] : 0.0;
      if (!std::isfinite(in.p[v]) || !std::isfinite(in.pos[v])) return fail(FH_ERR_ARG, "non-finite inclusion");
    }
    if (in.kind == fh::ga::INCL_BALL) in.p[0] = params[3 * ii];
    for (int c = 0; c < ncoef; ++c) {
      in.v[c] = vals[(size_t)ii * ncoef + c];
      if (!std::isfinite(in.v[c])) return fail(FH_ERR_ARG, "coefficient has a non-finite value");
    }
  }
  fh::ga::Incl* d = nullptr;
  HIPCHK(hipMalloc(&d, nincl * sizeof(fh::ga::Incl)));
  int rc = FH_OK;
  hipError_t e = hipMemcpy(d, h.data(), nincl * sizeof(fh::ga::Incl), hipMemcpyHostToDevice);
  if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
  if (rc == FH_OK) rc = ga_alloc_coef(ga, kind, ncoef);
  fh_plan* f = ga->fine;
  cpx<double>* fs = reinterpret_cast<cpx<double>*>(f->spec);
  const int nb = grid_blocks(f->nh, 256, 8192);
  for (int c = 0; c < ncoef && rc == FH_OK; ++c) {
    hipLaunchKernelGGL(fh::ga::k_ga_incl_spec, dim3(nb), dim3(256), 0, f->stream, d, nincl, c, ga->g, fs);
    e = hipGetLastError();
    if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("Ga coefficient: ") + hipGetErrorString(e));
    if (rc == FH_OK) rc = spec_to_real(f, fs, ga->coef + (size_t)c * f->n, 1.0);
  }
  e = hipStreamSynchronize(f->stream);
  if (rc == FH_OK && e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("Ga coefficient: ") + hipGetErrorString(e));
  (void)hipFree(d);
  if (rc == FH_OK) ga->coef_kind = kind;
  return rc;
}

int fh_ga_set_grid_coeff(fh_ga* ga, int order, const int* P, int kind, const double* vals) {
  if (!ga || !P || !vals) return fail(FH_ERR_ARG, "null pointer argument");
  if (order != 0 && order != 1) return fail(FH_ERR_ARG, "order must be 0 or 1");
  const int ncoef = ga_ncoef(ga, kind);
  if (ncoef < 0) return fail(FH_ERR_ARG, "bad coefficient kind");
  const int dim = ga->coarse->dim;
  fh_plan* pp = nullptr;
  RETCHK(fh_plan_create(dim, P, ga->coarse->Y, FH_SCALAR, FH_PRIMAL, &pp));
  const size_t np = (size_t)pp->n;
  for (size_t i = 0; i < (size_t)ncoef * np; ++i)
    if (!std::isfinite(vals[i])) {
      fh_plan_destroy(pp);
      return fail(FH_ERR_ARG, "coefficient has a non-finite value");
    }
  double* dv = nullptr;
  cpx<double>* dh = nullptr;
  int rc = FH_OK;
  hipError_t e = hipMalloc(&dv, np * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dh, (size_t)pp->nh * sizeof(cpx<double>));
  if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  if (rc == FH_OK) rc = ga_alloc_coef(ga, kind, ncoef);
  fh_plan* f = ga->fine;
  cpx<double>* fs = reinterpret_cast<cpx<double>*>(f->spec);
  const int nb = grid_blocks(f->nh, 256, 8192);
  for (int c = 0; c < ncoef && rc == FH_OK; ++c) {
    e = hipMemcpy(dv, vals + (size_t)c * np, np * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
    if (rc == FH_OK) rc = rfftn_impl<double>(pp, 1, dv, dh);
    if (rc == FH_OK) {
      e = hipStreamSynchronize(pp->stream);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(fh::ga::k_ga_grid_spec, dim3(nb), dim3(256), 0, f->stream, dh, pp->n0, pp->n1, pp->n2,
                           order + 1, ga->g, fs);
        e = hipGetLastError();
      }
      if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("Ga coefficient: ") + hipGetErrorString(e));
    }
    if (rc == FH_OK) rc = spec_to_real(f, fs, ga->coef + (size_t)c * f->n, 1.0);
    if (rc == FH_OK) {
      e = hipStreamSynchronize(f->stream);
      if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("Ga coefficient: ") + hipGetErrorString(e));
    }
  }
  if (dv) (void)hipFree(dv);
  if (dh) (void)hipFree(dh);
  fh_plan_destroy(pp);
  if (rc == FH_OK) ga->coef_kind = kind;
  return rc;
}

int fh_ga_set_plan_coeff(const fh_ga* ga, fh_plan* pl) {
  if (!ga || !pl) return fail(FH_ERR_ARG, "null pointer argument");
  drop_graphs(pl);
  if (!ga->coef_kind) return fail(FH_ERR_STATE, "Ga coefficient not set");
  const fh_plan* f = ga->fine;
  if (pl->dim != f->dim || pl->n0 != f->n0 || pl->n1 != f->n1 || pl->n2 != f->n2)
    return fail(FH_ERR_ARG, "plan grid differs from the Ga quadrature grid");
  if (pl->ncomp != f->ncomp || (ga->coef_kind == FH_COEF_LAME && pl->physics != FH_ELASTICITY))
    return fail(FH_ERR_ARG, "plan physics differs from the Ga coefficient");
  const size_t cnt = (size_t)ga->ncoef * pl->n;
  if (pl->coef64) (void)hipFree(pl->coef64);
  if (pl->coef32) (void)hipFree(pl->coef32);
  pl->coef64 = nullptr;
  pl->coef32 = nullptr;
  pl->coef_kind = 0;
  HIPCHK(hipMalloc(&pl->coef64, cnt * sizeof(double)));
  HIPCHK(hipMalloc(&pl->coef32, cnt * sizeof(float)));
  HIPCHK(hipMemcpy(pl->coef64, ga->coef, cnt * sizeof(double), hipMemcpyDeviceToDevice));
  hipLaunchKernelGGL((fh::k_convert<double, float>), dim3(grid_blocks((long long)cnt)), dim3(256), 0,
                     pl->stream, (long long)cnt, pl->coef64, pl->coef32);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(pl->stream));
  pl->coef_kind = ga->coef_kind;
  return FH_OK;
}

int fh_ga_coefficient(const fh_ga* ga, double* host_out) {
  if (!ga || !host_out) return fail(FH_ERR_ARG, "null pointer argument");
  if (!ga->coef_kind) return fail(FH_ERR_STATE, "Ga coefficient not set");
  HIPCHK(hipMemcpy(host_out, ga->coef, (size_t)ga->ncoef * ga->fine->n * sizeof(double), hipMemcpyDeviceToHost));
  return FH_OK;
}

int fh_ga_interpolate(fh_ga* ga, const void* x, void* out) {
  if (!ga || !x || !out) return fail(FH_ERR_ARG, "null pointer argument");
  return ga_interpolate(ga, (const double*)x, (double*)out);
}

int fh_ga_restrict(fh_ga* ga, const void* x_, void* out_) {
  if (!ga || !x_ || !out_) return fail(FH_ERR_ARG, "null pointer argument");
  fh_plan* c = ga->coarse;
  fh_plan* f = ga->fine;
  const double* x = (const double*)x_;
  double* out = (double*)out_;
  cpx<double>* fs = reinterpret_cast<cpx<double>*>(f->spec);
  cpx<double>* cs = reinterpret_cast<cpx<double>*>(c->spec);
  const int nb = grid_blocks(c->nh, 256, 8192);
  for (int k = 0; k < c->ncomp; ++k) {
    RETCHK(rfftn_impl<double>(f, 1, x + (size_t)k * f->n, fs));
    HIPCHK(hipStreamSynchronize(f->stream));
    hipLaunchKernelGGL(fh::ga::k_ga_crop, dim3(nb), dim3(256), 0, c->stream, fs, c->n0, c->n1, c->n2,
                       1.0 / (double)f->n, ga->g, cs);
    HIPCHK(hipGetLastError());
    RETCHK(spec_to_real(c, cs, out + (size_t)k * c->n, 1.0));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return FH_OK;
}

int fh_ga_energy_on_grid(fh_ga* ga, const void* const* fields, int nf, double* AH) {
  if (!ga || !fields || !AH) return fail(FH_ERR_ARG, "null pointer argument");
  if (nf != ga->coarse->ncomp) return fail(FH_ERR_ARG, "need one field per unit load (nf == ncomp)");
  if (!ga->coef_kind) return fail(FH_ERR_STATE, "Ga coefficient not set");
  for (int i = 0; i < nf; ++i)
    if (!fields[i]) return fail(FH_ERR_ARG, "null field pointer");
  return energy_impl(ga->fine, ga->coef, ga->coef_kind, fields, nf, AH);
}

int fh_ga_energy(fh_ga* ga, const void* const* fields, int nf, double* AH) {
  if (!ga || !fields || !AH) return fail(FH_ERR_ARG, "null pointer argument");
  const int nc = ga->coarse->ncomp;
  if (nf != nc) return fail(FH_ERR_ARG, "need one field per unit load (nf == ncomp)");
  if (!ga->coef_kind) return fail(FH_ERR_STATE, "Ga coefficient not set");
  for (int i = 0; i < nf; ++i)
    if (!fields[i]) return fail(FH_ERR_ARG, "null field pointer");
  const size_t per = (size_t)nc * ga->fine->n;
  const size_t bytes = (size_t)nf * per * sizeof(double);
  if (ga->fields_bytes < bytes) {
    if (ga->fields) (void)hipFree(ga->fields);
    ga->fields = nullptr;
    ga->fields_bytes = 0;
    HIPCHK(hipMalloc(&ga->fields, bytes));
    ga->fields_bytes = bytes;
  }
  std::vector<const void*> ptr(nf);
  for (int i = 0; i < nf; ++i) {
    RETCHK(ga_interpolate(ga, (const double*)fields[i], ga->fields + i * per));
    ptr[i] = ga->fields + i * per;
  }
  return energy_impl(ga->fine, ga->coef, ga->coef_kind, ptr.data(), nf, AH);
}


// ---- Richardson and Chebyshev iterations ---------------------------------------------------

int fh_iter_solve(fh_plan* pl, int method, const void* b_, void* x_, const double* coefs, double tol, int maxiter,
                  int* kit, double* norm_res, double* res_hist, int hist_len) {
  RETCHK(check_plan(pl));
  if (!b_ || !x_ || !coefs) return fail(FH_ERR_ARG, "null pointer argument");
  if (method != FH_ITER_RICHARDSON && method != FH_ITER_CHEBYSHEV) return fail(FH_ERR_ARG, "bad method");
  if (!pl->coef_kind) return fail(FH_ERR_STATE, "fh_set_coeff has not been called");
  if (!(tol >= 0.0) || maxiter < 0) return fail(FH_ERR_ARG, "bad tol/maxiter");
  const double* b = (const double*)b_;
  double* x = (double*)x_;
  const long long nv = (long long)pl->ncomp * pl->n;
  const long long nvp = (nv + 63) / 64 * 64;
  RETCHK(ensure_work(pl, 3 * (size_t)nvp * sizeof(double)));
  double* Ax = (double*)pl->work;
  double* r = Ax + nvp;
  double* v = r + nvp;
  RETCHK(ensure_hist(pl, std::max(hist_len, 1)));
  fh::it::IterState s;
  std::memset(&s, 0, sizeof(s));
  s.tol = tol;
  s.inv_nvox = 1.0 / (double)pl->n;
  s.maxiter = maxiter;
  s.hist_len = hist_len > 0 ? hist_len : 0;
  s.hist = pl->hist;
  s.norm = 1e15;  // the reference's initial value (richardson, solver.py:65)
  fh::it::IterState* st = nullptr;
  HIPCHK(hipMalloc(&st, sizeof(s)));
  int rc = FH_OK;
  double norm_init = s.norm;  // before the first iteration: 1e15 (Richardson) or |r0| / |b| (Chebyshev)
  // host callback on the iterate (the reference's callback(x) after every step, solver.py:75-76,
  // 277-278): the batches end on its iterations
  const int cb_every = pl->it_fn ? pl->it_every : 0;
  int cb_last = 0;
  auto batch_len = [&](int done_iters) {
    return cb_every > 0 ? std::min(8, cb_every - done_iters % cb_every) : 8;
  };
  auto call_back = [&](int limit) -> int {
    if (cb_every > 0 && s.kit > cb_last && (s.kit % cb_every == 0 || s.done || s.kit >= limit)) {
      cb_last = s.kit;
      if (pl->it_fn(pl->it_user, s.kit, x) != 0) return fail(FH_ERR_STATE, "solve stopped by the iterate callback");
    }
    return FH_OK;
  };
  auto run = [&]() -> int {
    const int gb = grid_blocks(nv);
    int* halt = &st->halt;
    if (method == FH_ITER_RICHARDSON) {
      if (!(coefs[0] != 0.0)) return fail(FH_ERR_ARG, "alpha must be nonzero");
      const double omega = 1.0 / coefs[0];
      s.done = s.halt = maxiter == 0;
      HIPCHK(hipMemcpy(st, &s, sizeof(s), hipMemcpyHostToDevice));
      for (int k = 0; !s.done && k < maxiter;) {
        const int per_check = batch_len(k);
        for (int j = 0; j < per_check && k < maxiter; ++j, ++k) {
          RETCHK(run_operator<double>(pl, x, Ax, pl->coef_kind, pl->coef64, nullptr, halt, nullptr));
          hipLaunchKernelGGL(fh::it::k_it_residual, dim3(gb), dim3(256), 0, pl->stream, nv, b, Ax, r, pl->partial,
                             halt);
          hipLaunchKernelGGL(fh::it::k_it_richardson, dim3(gb), dim3(256), 0, pl->stream, nv, x, r, omega, halt);
          hipLaunchKernelGGL(fh::it::k_it_finish, dim3(1), dim3(256), 0, pl->stream, st, pl->partial, gb);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&s, st, sizeof(s), hipMemcpyDeviceToHost, pl->stream));
        HIPCHK(hipStreamSynchronize(pl->stream));
        RETCHK(call_back(maxiter));
      }
      return FH_OK;
    }
    // Chebyshev: initial residual, its norm r0 and the norm of b (solver.py:250-262)
    const double e0 = coefs[0], e1 = coefs[1];
    if (!(e1 > e0) || !(e0 > 0.0)) return fail(FH_ERR_ARG, "eigenvalue range must satisfy 0 < min < max");
    s.halt = 0;
    HIPCHK(hipMemcpy(st, &s, sizeof(s), hipMemcpyHostToDevice));
    RETCHK(run_operator<double>(pl, x, Ax, pl->coef_kind, pl->coef64, nullptr, nullptr, nullptr));
    hipLaunchKernelGGL(fh::it::k_it_residual, dim3(gb), dim3(256), 0, pl->stream, nv, b, Ax, r, pl->partial, halt);
    HIPCHK(hipGetLastError());
    std::vector<double> part(gb);
    HIPCHK(hipMemcpyAsync(part.data(), pl->partial, gb * sizeof(double), hipMemcpyDeviceToHost, pl->stream));
    HIPCHK(hipStreamSynchronize(pl->stream));
    double rr = 0.0;
    for (double q : part) rr += q;
    const double r0 = std::sqrt(rr * s.inv_nvox);
    hipLaunchKernelGGL((fh::k_residual<double>), dim3(gb), dim3(256), 0, pl->stream, nv, b, nullptr, v, pl->partial);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(part.data(), pl->partial, gb * sizeof(double), hipMemcpyDeviceToHost, pl->stream));
    HIPCHK(hipStreamSynchronize(pl->stream));
    double bb = 0.0;
    for (double q : part) bb += q;
    const double bn = std::sqrt(bb * s.inv_nvox);
    s.norm = r0 / bn;  // the reference's Ib * r0
    norm_init = s.norm;
    s.r0 = r0;
    s.relative = 1;
    if (s.norm < tol || maxiter == 0) {
      s.done = 1;
      if (maxiter == 0 && !(s.norm < tol)) s.norm = 0.0;  // kit == 0 at the end: norm_res = 0 (solver.py:282-283)
      return FH_OK;
    }
    HIPCHK(hipMemsetAsync(v, 0, (size_t)nv * sizeof(double), pl->stream));
    HIPCHK(hipMemcpyAsync(st, &s, sizeof(s), hipMemcpyHostToDevice, pl->stream));
    const double d = (e1 + e0) / 2.0, c = (e1 - e0) / 2.0;
    double w = 0.0;
    for (int k = 1; !s.done && k <= maxiter;) {
      const int per_check = batch_len(k - 1);
      for (int j = 0; j < per_check && k <= maxiter; ++j, ++k) {
        double pk;
        if (k == 1) {
          pk = 0.0;
          w = 1.0 / d;
        } else if (k == 2) {
          pk = -(1.0 / 2.0) * (c / d) * (c / d);
          w = 1.0 / (d - c * c / 2.0 / d);
        } else {
          pk = -(c * c / 4.0) * w * w;
          w = 1.0 / (d - c * c * w / 4.0);
        }
        hipLaunchKernelGGL(fh::it::k_it_chebyshev, dim3(gb), dim3(256), 0, pl->stream, nv, x, v, r, pk, w, halt);
        RETCHK(run_operator<double>(pl, x, Ax, pl->coef_kind, pl->coef64, nullptr, halt, nullptr));
        hipLaunchKernelGGL(fh::it::k_it_residual, dim3(gb), dim3(256), 0, pl->stream, nv, b, Ax, r, pl->partial,
                           halt);
        hipLaunchKernelGGL(fh::it::k_it_finish, dim3(1), dim3(256), 0, pl->stream, st, pl->partial, gb);
      }
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(&s, st, sizeof(s), hipMemcpyDeviceToHost, pl->stream));
      HIPCHK(hipStreamSynchronize(pl->stream));
      RETCHK(call_back(maxiter));
    }
    return FH_OK;
  };
  rc = run();
  if (rc == FH_OK) {
    hipError_t e = hipStreamSynchronize(pl->stream);
    if (e == hipSuccess && res_hist && hist_len > 0) {
      const int cnt = std::min(hist_len, s.kit + 1);
      if (cnt > 1) e = hipMemcpy(res_hist + 1, pl->hist + 1, (size_t)(cnt - 1) * sizeof(double), hipMemcpyDeviceToHost);
      res_hist[0] = norm_init;
    }
    if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("iterative solve: ") + hipGetErrorString(e));
  }
  (void)hipFree(st);
  if (rc == FH_OK) {
    if (kit) *kit = s.kit;
    if (norm_res) *norm_res = s.norm;
  }
  return rc;
}


// ---- Coefficients and random fields generated on the device -------------------------------

int fh_set_coeff_inclusions(fh_plan* pl, int nincl, const int* kinds, const double* params, const double* positions,
                            int kind, const double* vals, int invert) {
  RETCHK(check_plan(pl));
  drop_graphs(pl);
  if (!kinds || !params || !positions || !vals) return fail(FH_ERR_ARG, "null pointer argument");
  if (nincl < 1 || nincl > 4096) return fail(FH_ERR_ARG, "number of inclusions out of range");
  int ncoef, layout;
  if (kind == FH_COEF_ISO) {
    ncoef = 1;
    layout = 1;
  } else if (kind == FH_COEF_LAME) {
    if (pl->physics != FH_ELASTICITY) return fail(FH_ERR_ARG, "Lame coefficients need elasticity");
    ncoef = 2;
    layout = 2;
  } else if (kind == FH_COEF_FULL) {
    ncoef = pl->ncomp * pl->ncomp;
    layout = 3;
  } else {
    return fail(FH_ERR_ARG, "bad coefficient kind");
  }
  const int dim = pl->dim;
  std::vector<fh::ga::Incl> h(nincl);
  for (int ii = 0; ii < nincl; ++ii) {
    fh::ga::Incl& in = h[ii];
    std::memset(&in, 0, sizeof(in));
    in.kind = kinds[ii];
    if (in.kind < fh::ga::INCL_CUBE || in.kind > fh::ga::INCL_OTHERWISE) return fail(FH_ERR_ARG, "unknown inclusion kind");
    if (in.kind == fh::ga::INCL_OTHERWISE && ii != nincl - 1)
      return fail(FH_ERR_ARG, "'otherwise' must be the last inclusion");
    for (int a = 0; a < dim; ++a) {
      in.p[a] = params[3 * ii + a];
      in.pos[a] = positions[3 * ii + a];
      if (!std::isfinite(in.p[a]) || !std::isfinite(in.pos[a])) return fail(FH_ERR_ARG, "non-finite inclusion");
    }
    for (int c = 0; c < ncoef; ++c) {
      in.v[c] = vals[(size_t)ii * ncoef + c];
      if (!std::isfinite(in.v[c])) return fail(FH_ERR_ARG, "coefficient has a non-finite value");
    }
  }
  fh::micro::NodalGrid g{};
  g.dim = dim;
  for (int a = 0; a < dim; ++a) {
    g.N[a] = pl->Nu[a];
    g.Y[a] = pl->Y[a];
  }
  g.n = pl->n;
  g.ncoef = ncoef;
  g.layout = layout;
  g.k = pl->ncomp;
  g.invert = invert ? 1 : 0;
  const size_t cnt = (size_t)ncoef * pl->n;
  fh::ga::Incl* d = nullptr;
  int* derr = nullptr;
  double* c64 = nullptr;
  float* c32 = nullptr;
  int rc = FH_OK;
  hipError_t e = hipMalloc(&d, nincl * sizeof(fh::ga::Incl));
  if (e == hipSuccess) e = hipMalloc(&derr, 2 * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&c64, cnt * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c32, cnt * sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(d, h.data(), nincl * sizeof(fh::ga::Incl), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(derr, 0, 2 * sizeof(int));
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fh::micro::k_nodal_coef, dim3(grid_blocks(pl->n, 256, 16384)), dim3(256), 0, pl->stream, d,
                       nincl, g, c64, derr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL((fh::k_convert<double, float>), dim3(grid_blocks((long long)cnt)), dim3(256), 0, pl->stream,
                       (long long)cnt, c64, c32);
    e = hipGetLastError();
  }
  int herr[2] = {0, 0};
  if (e == hipSuccess) e = hipStreamSynchronize(pl->stream);
  if (e == hipSuccess) e = hipMemcpy(herr, derr, 2 * sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("nodal coefficient: ") + hipGetErrorString(e));
  else if (herr[0]) rc = fail(FH_ERR_ARG, "Overlapping inclusions!");
  else if (herr[1]) rc = fail(FH_ERR_ARG, "singular coefficient matrix at a node");
  (void)hipFree(d);
  (void)hipFree(derr);
  if (rc != FH_OK) {
    (void)hipFree(c64);
    (void)hipFree(c32);
    return rc;
  }
  if (pl->coef64) (void)hipFree(pl->coef64);
  if (pl->coef32) (void)hipFree(pl->coef32);
  pl->coef64 = c64;
  pl->coef32 = c32;
  pl->coef_kind = kind;
  return FH_OK;
}

int fh_set_coeff_device(fh_plan* pl, int kind, const void* dev_vals) {
  RETCHK(check_plan(pl));
  drop_graphs(pl);
  if (!dev_vals) return fail(FH_ERR_ARG, "null coefficient array");
  int per;
  if (kind == FH_COEF_ISO) per = 1;
  else if (kind == FH_COEF_LAME && pl->physics == FH_ELASTICITY) per = 2;
  else if (kind == FH_COEF_FULL) per = pl->ncomp * pl->ncomp;
  else return fail(FH_ERR_ARG, "bad coefficient kind");
  const size_t cnt = (size_t)per * pl->n;
  double* c64 = nullptr;
  float* c32 = nullptr;
  int* dbad = nullptr;
  int bad = 0;
  hipError_t e = hipMalloc(&c64, cnt * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&c32, cnt * sizeof(float));
  if (e == hipSuccess) e = hipMalloc(&dbad, sizeof(int));
  if (e == hipSuccess) e = hipMemsetAsync(dbad, 0, sizeof(int), pl->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(c64, dev_vals, cnt * sizeof(double), hipMemcpyDeviceToDevice, pl->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(fh::micro::k_count_nonfinite, dim3(grid_blocks((long long)cnt)), dim3(256), 0, pl->stream,
                       (long long)cnt, c64, dbad);
    hipLaunchKernelGGL((fh::k_convert<double, float>), dim3(grid_blocks((long long)cnt)), dim3(256), 0, pl->stream,
                       (long long)cnt, c64, c32);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(pl->stream);
  if (e == hipSuccess) e = hipMemcpy(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost);
  if (dbad) (void)hipFree(dbad);
  if (e != hipSuccess || bad) {
    if (c64) (void)hipFree(c64);
    if (c32) (void)hipFree(c32);
    if (e != hipSuccess) return fail(FH_ERR_HIP, std::string("fh_set_coeff_device: ") + hipGetErrorString(e));
    return fail(FH_ERR_ARG, "coefficient has a non-finite value");
  }
  if (pl->coef64) (void)hipFree(pl->coef64);
  if (pl->coef32) (void)hipFree(pl->coef32);
  pl->coef64 = c64;
  pl->coef32 = c32;
  pl->coef_kind = kind;
  return FH_OK;
}

int fh_coeff_to_host(const fh_plan* pl, double* host_out) {
  RETCHK(check_plan(pl));
  if (!host_out) return fail(FH_ERR_ARG, "null pointer argument");
  if (!pl->coef_kind) return fail(FH_ERR_STATE, "fh_set_coeff has not been called");
  const int per = pl->coef_kind == FH_COEF_ISO ? 1 : (pl->coef_kind == FH_COEF_LAME ? 2 : pl->ncomp * pl->ncomp);
  HIPCHK(hipMemcpy(host_out, pl->coef64, (size_t)per * pl->n * sizeof(double), hipMemcpyDeviceToHost));
  return FH_OK;
}

int fh_grf(fh_plan* pl, const void* white, double corr_len, double stdev, int exponentiate, void* out_) {
  RETCHK(check_plan(pl));
  if (!white || !out_) return fail(FH_ERR_ARG, "null pointer argument");
  if (!(corr_len >= 0.0) || !std::isfinite(stdev)) return fail(FH_ERR_ARG, "bad correlation length / std");
  // s = std / sqrt(mean g^2), g^2 = prod_a exp(-2 pi^2 l^2 (k_a / Y_a)^2) summed axis by axis
  const double pi = 3.14159265358979323846;
  double m2 = 1.0;
  for (int a = 0; a < pl->dim; ++a) {
    const int n = pl->Nu[a];
    double acc = 0.0;
    for (int i = 0; i < n; ++i) {
      const int k = (2 * i < n) ? i : i - n;
      const double q = (double)k / pl->Y[a];
      acc += std::exp(-2.0 * pi * pi * corr_len * corr_len * q * q);
    }
    m2 *= acc / (double)n;
  }
  const double sc = stdev / std::sqrt(m2);
  double* out = (double*)out_;
  cpx<double>* tmp = nullptr;
  HIPCHK(hipMalloc(&tmp, (size_t)pl->nh * sizeof(cpx<double>)));
  int rc = rfftn_impl<double>(pl, 1, (const double*)white, tmp);
  if (rc == FH_OK) {
    const double iy0 = pl->dim == 3 ? 1.0 / pl->Y[0] : 0.0;
    const double iy1 = 1.0 / pl->Y[pl->dim == 3 ? 1 : 0], iy2 = 1.0 / pl->Y[pl->dim - 1];
    hipLaunchKernelGGL(fh::micro::k_grf_filter, dim3(grid_blocks(pl->nh, 256, 16384)), dim3(256), 0, pl->stream, tmp,
                       pl->n0, pl->n1, pl->n2, iy0, iy1, iy2, corr_len, sc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("grf: ") + hipGetErrorString(e));
  }
  if (rc == FH_OK) rc = irfftn_impl<double>(pl, 1, tmp, out);
  if (rc == FH_OK && exponentiate)
    hipLaunchKernelGGL(fh::micro::k_exp_inplace, dim3(grid_blocks(pl->n, 256, 16384)), dim3(256), 0, pl->stream,
                       (long long)pl->n, out);
  if (rc == FH_OK) {
    hipError_t e = hipStreamSynchronize(pl->stream);
    if (e != hipSuccess) rc = fail(FH_ERR_HIP, std::string("grf: ") + hipGetErrorString(e));
  }
  (void)hipFree(tmp);
  return rc;
}


// ---- potential formulation (ffthompy/tensorsLowRank/homogenisation.py:43-128) ------------------

int fh_potential_solve(fh_plan* pl, int mode, const double* E, double tol, int maxiter, double delta, void* Fu_,
                       void* X_, int* kit, double* norm_res, double* res_hist, int hist_len, int* cg_status) {
  RETCHK(check_plan(pl));
  if (!E || !Fu_) return fail(FH_ERR_ARG, "null pointer argument");
  if (mode != FH_SOLVE_FP64 && mode != FH_SOLVE_MIXED) return fail(FH_ERR_ARG, "bad solve mode");
  if (pl->physics != FH_SCALAR) return fail(FH_ERR_UNSUPPORTED, "the potential formulation is scalar only");
  if (!pl->coef_kind) return fail(FH_ERR_STATE, "fh_set_coeff has not been called");
  if (!(tol >= 0.0) || maxiter < 0) return fail(FH_ERR_ARG, "bad tol/maxiter");
  if (mode == FH_SOLVE_MIXED && !(delta > 0.0 && delta < 1.0)) return fail(FH_ERR_ARG, "delta must be in (0,1)");
  // the logs of a previous fh_cg_solve never describe this solve
  pl->res_kit.clear();
  pl->res_norm.clear();
  if (pl->res_every > 0)
    return fail(FH_ERR_UNSUPPORTED, "the residual log (fh_plan_set_residual_log) is kept by fh_cg_solve only");
  const bool mixed = mode == FH_SOLVE_MIXED;
  pl->last_graph_batches = 0;
  const fh::pot::PotParams g = pot_params(pl);
  const long long nh = pl->nh;
  const long long nhp = (nh + 15) / 16 * 16;  // 256-byte aligned sub-buffers
  const int nc = pl->ncomp;
  // fp64: b, r, p, Ap; mixed: b and A x in fp64, r, p, Ap and the correction xc in fp32; both: the
  // real field of the right-hand side (and of the one-way passes where the fused kernels do not apply)
  const size_t vec_bytes = mixed ? 2 * (size_t)nhp * sizeof(cpx<double>) + 4 * (size_t)nhp * sizeof(cpx<float>)
                                 : 4 * (size_t)nhp * sizeof(cpx<double>);
  RETCHK(ensure_work(pl, vec_bytes + (size_t)nc * pl->n * sizeof(double)));
  cpx<double>* b = reinterpret_cast<cpx<double>*>(pl->work);
  cpx<double>* v1 = b + nhp;  // fp64: r; mixed: A x
  double* re = reinterpret_cast<double*>(reinterpret_cast<char*>(pl->work) + vec_bytes);
  cpx<double>* x = reinterpret_cast<cpx<double>*>(Fu_);  // the CG iterate; P x at the end
  const int hl = std::max(hist_len, 1);
  RETCHK(ensure_hist(pl, hl));
  CGState s;
  std::memset(&s, 0, sizeof(s));
  s.tol = tol;
  s.maxiter = maxiter;
  s.mixed = mixed;
  s.delta = delta;
  s.inv_nvox = 1.0 / ((double)pl->n * (double)pl->n);  // 'r' spectra: / n^2
  s.hist_len = hist_len > 0 ? hist_len : 0;
  s.hist = pl->hist;
  HIPCHK(hipMemcpyAsync(pl->st, &s, sizeof(s), hipMemcpyHostToDevice, pl->stream));
  HIPCHK(hipEventRecord(pl->ev0, pl->stream));
  const int gb = grid_blocks(nh);
  const dim3 grid(gb), block(256);
  // right-hand side b = P div F(A E) (homogenisation.py:106-110), x0 = 0
  hipLaunchKernelGGL(fh::pot::k_pot_fill, dim3(grid_blocks((long long)nc * pl->n)), block, 0, pl->stream, re,
                     pl->n, nc, E[0], E[1], nc > 2 ? E[2] : 0.0);
  HIPCHK(hipGetLastError());
  cpx<double>* spec = reinterpret_cast<cpx<double>*>(pl->spec);
  const GParams gp = make_gparams(pl, 1.0);
  RETCHK(launch_rows_fwd<double>(pl, nc, re, pl->coef64, pl->coef_kind, spec, nullptr));
  RETCHK(launch_cols<double>(pl, 0, 1, nc, spec, gp, nullptr));
  if (pl->n0 > 1) RETCHK(launch_cols<double>(pl, 0, 0, nc, spec, gp, nullptr));
  hipLaunchKernelGGL((fh::pot::k_pot_div<double>), grid, block, 0, pl->stream, spec, b, g, 1.0, nullptr, nullptr,
                     nullptr);
  HIPCHK(hipMemsetAsync(x, 0, (size_t)nh * sizeof(cpx<double>), pl->stream));
  const int* halt = &pl->st->halt;
  const int per_check = 8;
  if (!mixed) {
    cpx<double>* r = v1;
    cpx<double>* p = r + nhp;
    cpx<double>* Ap = p + nhp;
    hipLaunchKernelGGL((fh::pot::k_pot_residual<double>), grid, block, 0, pl->stream, b, nullptr, r, g, pl->partial);
    hipLaunchKernelGGL(fh::k_cg_restart, dim3(1), dim3(256), 0, pl->stream, pl->st, pl->partial, gb, pl->hist, false);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(p, r, (size_t)nh * sizeof(cpx<double>), hipMemcpyDeviceToDevice, pl->stream));
    auto batch64 = [&]() -> int {
      for (int k = 0; k < per_check; ++k) {
        int nb = 0;
        RETCHK(pot_operator<double>(pl, p, r, Ap, re, p, halt, &nb));
        hipLaunchKernelGGL(fh::k_cg_alpha, dim3(1), dim3(kFinThreads), 0, pl->stream, pl->st, pl->partial, nb);
        hipLaunchKernelGGL((fh::pot::k_pot_update<double>), grid, block, 0, pl->stream, x, r, p, Ap, g, pl->st,
                           pl->partial);
        hipLaunchKernelGGL(fh::k_cg_beta, dim3(1), dim3(kFinThreads), 0, pl->stream, pl->st, pl->partial, gb);
        HIPCHK(hipGetLastError());
      }
      return FH_OK;
    };
    for (;;) {
      RETCHK(batch64());
      HIPCHK(hipMemcpyAsync(&s, pl->st, sizeof(s), hipMemcpyDeviceToHost, pl->stream));
      HIPCHK(hipStreamSynchronize(pl->stream));
      if (s.done) break;
    }
  } else {
    // fp32 iterations on the correction xc; a reliable update (x += xc, r = b - A x in fp64)
    // whenever the recursive residual has dropped by delta or meets the tolerance
    cpx<double>* Ax = v1;
    cpx<float>* r = reinterpret_cast<cpx<float>*>(Ax + nhp);
    cpx<float>* p = r + nhp;
    cpx<float>* Ap = p + nhp;
    cpx<float>* xc = Ap + nhp;
    float* re32 = reinterpret_cast<float*>(re);
    HIPCHK(hipMemsetAsync(xc, 0, (size_t)nh * sizeof(cpx<float>), pl->stream));
    hipLaunchKernelGGL((fh::pot::k_pot_residual<float>), grid, block, 0, pl->stream, b, nullptr, r, g, pl->partial);
    hipLaunchKernelGGL(fh::k_cg_restart, dim3(1), dim3(256), 0, pl->stream, pl->st, pl->partial, gb, pl->hist, false);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(p, r, (size_t)nh * sizeof(cpx<float>), hipMemcpyDeviceToDevice, pl->stream));
    const int fb = grid_blocks(2 * nh);
    auto batch32 = [&]() -> int {
      for (int k = 0; k < per_check; ++k) {
        int nb = 0;
        RETCHK(pot_operator<float>(pl, p, r, Ap, re32, p, halt, &nb));
        hipLaunchKernelGGL(fh::k_cg_alpha, dim3(1), dim3(kFinThreads), 0, pl->stream, pl->st, pl->partial, nb);
        hipLaunchKernelGGL((fh::pot::k_pot_update<float>), grid, block, 0, pl->stream, xc, r, p, Ap, g, pl->st,
                           pl->partial);
        hipLaunchKernelGGL(fh::k_cg_beta, dim3(1), dim3(kFinThreads), 0, pl->stream, pl->st, pl->partial, gb);
        HIPCHK(hipGetLastError());
      }
      return FH_OK;
    };
    for (;;) {
      RETCHK(batch32());
      HIPCHK(hipMemcpyAsync(&s, pl->st, sizeof(s), hipMemcpyDeviceToHost, pl->stream));
      HIPCHK(hipStreamSynchronize(pl->stream));
      if (s.done) break;
      if (s.pause) {
        hipLaunchKernelGGL(fh::k_fold_correction, dim3(fb), block, 0, pl->stream, 2 * nh, (double*)x, (float*)xc, false);
        RETCHK(pot_operator<double>(pl, x, nullptr, Ax, re, nullptr, nullptr, nullptr));
        hipLaunchKernelGGL((fh::pot::k_pot_residual<float>), grid, block, 0, pl->stream, b, Ax, r, g, pl->partial);
        hipLaunchKernelGGL(fh::k_cg_restart, dim3(1), dim3(256), 0, pl->stream, pl->st, pl->partial, gb, pl->hist, false);
        HIPCHK(hipGetLastError());
      }
    }
    hipLaunchKernelGGL(fh::k_fold_correction, dim3(fb), block, 0, pl->stream, 2 * nh, (double*)x, (float*)xc, false);
    HIPCHK(hipGetLastError());
  }
  // Fu = P x (homogenisation.py:117); e = F^-1 grad Fu
  cpx<double>* tmp = v1;
  hipLaunchKernelGGL(fh::pot::k_pot_precond, grid, block, 0, pl->stream, x, tmp, g);
  HIPCHK(hipMemcpyAsync(x, tmp, (size_t)nh * sizeof(cpx<double>), hipMemcpyDeviceToDevice, pl->stream));
  if (X_) {
    hipLaunchKernelGGL((fh::pot::k_pot_grad<double>), grid, block, 0, pl->stream, x, spec, g, 0, nullptr);
    HIPCHK(hipGetLastError());
    if (pl->n0 > 1) RETCHK(launch_cols<double>(pl, 1, 0, nc, spec, gp, nullptr));
    RETCHK(launch_cols<double>(pl, 1, 1, nc, spec, gp, nullptr));
    RETCHK(launch_rows_inv<double>(pl, nc, spec, (double*)X_, 1.0 / (double)pl->n, nullptr, nullptr, nullptr,
                                   nullptr));
  }
  HIPCHK(hipEventRecord(pl->ev1, pl->stream));
  HIPCHK(hipMemcpyAsync(&s, pl->st, sizeof(s), hipMemcpyDeviceToHost, pl->stream));
  HIPCHK(hipStreamSynchronize(pl->stream));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, pl->ev0, pl->ev1));
  pl->last_solve_ms = ms;
  pl->last_updates = s.nupdates;
  if (kit) *kit = s.kit;
  if (norm_res) *norm_res = (s.kit == 0) ? 0.0 : std::sqrt(s.rr);
  if (cg_status) *cg_status = s.status;
  if (res_hist && hist_len > 0) {
    const int cnt = std::min(hist_len, s.kit + 1);
    HIPCHK(hipMemcpy(res_hist, pl->hist, (size_t)cnt * sizeof(double), hipMemcpyDeviceToHost));
  }
  return FH_OK;
}

}  // extern "C"
