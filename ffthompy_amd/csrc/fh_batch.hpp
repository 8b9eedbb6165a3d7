// Batched 2-D scalar GaNi solves on gfx950: one persistent kernel, one workgroup per
// (sample, primal|dual) task, per-task CG state kept by the workgroup.
//
// Each task is the reference's scalar GaNi problem on its own grid (n0, n1)
// (ffthompy/applications.py:24-90): right-hand sides B = G(-A E) for the unit loads E,
// CG to an absolute tolerance on sqrt(<r,r>) (ffthompy/general/solver.py:80-139), and
// the energy matrix <A e_i, e_j> (ffthompy/postprocess.py:53-70). A workgroup of NT
// threads owns one task at a time and takes the next from a device counter when done, so
// uneven grids balance over the chip without host round trips.
//
// Matrix-vector product y = F^-1 G F (c x) inside the workgroup, both components of the
// field at once:
//   rows_fwd : per grid row i0, the complex row z = c x0 + i c x1 is transformed along
//              axis 1 (Stockham stages in LDS), then split into the two half spectra
//              U = (Z_k + conj Z_{n-k})/2, V = (Z_k - conj Z_{n-k})/(2i).
//   cols_G   : per half-spectrum column k1, forward transform along axis 0 for both
//              components, G(xi) applied in registers between the last forward and the
//              first inverse stage, inverse transform.
//   rows_inv : per row, Z = U + i V rebuilt with Hermitian symmetry, inverse transform;
//              real and imaginary parts are the two output components.
// The half spectrum (n0 rows of 2*(n1/2+1) complex values, layout [k0][component][k1])
// sits in LDS in single precision (at most 135 x 136 complex64 = 146880 bytes) and in a
// per-workgroup global scratch buffer in double precision.
//
// Mixed precision: the inner CG runs in fp32; when the recursive residual has dropped by
// `delta` against the last true residual, or the stop test fires, the fp32 correction is
// projected with the fp64 operator and added to the fp64 solution, and the true residual
// b - A x is recomputed in fp64 (reliable update); the search direction restarts from it. The
// solution stays in the range of G; a breakdown ends with the fp64 projection x <- G x.
#pragma once
#include "fh_fixed.hpp"
#include "fh_kernels.hpp"

namespace fh {
namespace batch {

constexpr int NT = 512;      // threads per workgroup
constexpr int MAXL = 135;    // longest supported axis
constexpr int TWS = 136;     // twiddle table slot (entries)
constexpr int NLEN = 6;      // supported axis lengths
constexpr int kLens[NLEN] = {27, 45, 63, 81, 105, 135};

__host__ __device__ constexpr int len_id(int L) {
  return L == 27 ? 0 : L == 45 ? 1 : L == 63 ? 2 : L == 81 ? 3 : L == 105 ? 4 : L == 135 ? 5 : -1;
}

// LDS layout (bytes): fp64 twiddles of axes 0/1, fp32 twiddles of axes 0/1, reduction
// slots, then the fp32 half spectrum.
constexpr int OFF_TW64 = 0;
constexpr int OFF_TW32 = OFF_TW64 + 2 * TWS * 16;
constexpr int OFF_RED = OFF_TW32 + 2 * TWS * 8;
constexpr int OFF_SPEC = OFF_RED + 64 * 8;

__host__ __device__ constexpr long long spec_elems(int n0, int n1) { return (long long)n0 * 2 * (n1 / 2 + 1); }

struct Sample {
  int n0, n1;
  long long off;  // first voxel of this sample in the concatenated coefficient arrays
  double iY0, iY1;
};

struct Params {
  double tol, delta;
  int maxiter, mixed;
};

struct Args {
  const Sample* smp;
  const int* order;  // task t -> sample * 2 + pd (pd 0 primal, 1 dual), largest grids first
  int ntasks;
  int* counter;      // next task; cleared by k_batch_reset before every launch
  const double* a64; // conductivity a(x), concatenated over samples, origin-0 C order
  const float* a32;
  const cpx<double>* tw64;  // NLEN tables of TWS entries: exp(-2 pi i m / L)
  const cpx<float>* tw32;
  double* ws64;      // per-workgroup workspace, stride ws64_stride doubles
  float* ws32;       // per-workgroup workspace, stride ws32_stride floats
  long long ws64_stride, ws32_stride, nvec;  // nvec: one field (2 components), padded
  double* energy;    // [sample][pd][4]  <c e_i, e_j>
  int* kit;          // [sample][pd][2]
  int* nupd;         // [sample][pd][2]  true-residual evaluations
  double* normres;   // [sample][pd][2]
  int* status;       // [sample][pd]     0 ok, 1 breakdown
  const Params* par;
};

struct Task {
  int n0, n1, h, nv;  // h = n1/2 + 1, nv = n0*n1 voxels
  int pd;
  double iY0, iY1, scale;
};

// ---- LDS access ---------------------------------------------------------------------------
// opaque(): a value the compiler cannot see through. Each pass launders the thread index,
// its LDS offsets and base pointers, so index and address arithmetic is rebuilt inside the
// pass instead of being hoisted to the kernel entry and kept live across the whole solve.
__device__ __forceinline__ int opaque(int v) {
  asm volatile("" : "+s"(v));
  return v;
}
__device__ __forceinline__ int opaque_tid() {
  int v = threadIdx.x;
  asm volatile("" : "+v"(v));
  return v;
}
template <typename P>
__device__ __forceinline__ P* opaque_ptr(P* p) {
  asm volatile("" : "+s"(p));
  return p;
}

template <typename T>
__device__ __forceinline__ cpx<T>* lds_tw(int axis) {
  extern __shared__ __align__(16) unsigned char fh_batch_smem[];
  const int off = opaque(sizeof(T) == 8 ? OFF_TW64 : OFF_TW32);
  return reinterpret_cast<cpx<T>*>(fh_batch_smem + off) + axis * TWS;
}
__device__ __forceinline__ double* lds_red() {
  extern __shared__ __align__(16) unsigned char fh_batch_smem[];
  return reinterpret_cast<double*>(fh_batch_smem + OFF_RED);
}
__device__ __forceinline__ cpx<float>* lds_spec() {
  extern __shared__ __align__(16) unsigned char fh_batch_smem[];
  return reinterpret_cast<cpx<float>*>(fh_batch_smem + opaque(OFF_SPEC));
}

// ---- workgroup reductions (fixed order: every thread gets the same bits) --------------------
__device__ __forceinline__ double wg_sum(double v) {
  double* red = lds_red();
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

__device__ __forceinline__ dd wg_sum_dd(dd v) {
  double* red = lds_red();
  v = wave_sum_dd(v);
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = v.hi;
    red[2 * (threadIdx.x >> 6) + 1] = v.lo;
  }
  __syncthreads();
  dd s;
  s.hi = 0.0;
  s.lo = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) {
    dd x;
    x.hi = red[2 * w];
    x.lo = red[2 * w + 1];
    s = dd_add(s, x);
  }
  __syncthreads();
  return s;
}

// ---- operands of the matrix-vector product --------------------------------------------------
// Input at voxel vox (both components), times the coefficient c (none, a, or 1/a):
//   mode 0: the constant field (e0, e1); mode 1: the vector v; mode 2: v <- r + beta v
//   (the CG direction update, written back) and then v.
template <typename T>
struct InOp {
  int mode, coef, nv;
  T e0, e1, beta;
  T* v;
  const T* r;
  const T* a;
  __device__ __forceinline__ cpx<T> operator()(int vox) const {
    T v0, v1;
    if (mode == 0) {
      v0 = e0;
      v1 = e1;
    } else {
      v0 = v[vox];
      v1 = v[nv + vox];
      if (mode == 2) {
        v0 = r[vox] + beta * v0;
        v1 = r[nv + vox] + beta * v1;
        v[vox] = v0;
        v[nv + vox] = v1;
      }
    }
    if (coef) {
      T c = a[vox];
      if (coef == 2) c = T(1) / c;
      v0 *= c;
      v1 *= c;
    }
    return mk(v0, v1);
  }
};

// Output y at voxel vox:
//   mode 0: y stored; mode 1: y stored and acc += <p, y>;
//   mode 2 (T = double): t = b - y stored as the residual (fp32 r32 or fp64 r64), acc += t.t
template <typename T>
struct OutOp {
  int mode, nv;
  T* y;
  const T* p;
  const double* b;
  float* r32;
  double* r64;
  double acc;
  __device__ __forceinline__ void operator()(int vox, T y0, T y1) {
    if (mode == 2) {
      const double t0 = b[vox] - (double)y0, t1 = b[nv + vox] - (double)y1;
      if (r32) {
        r32[vox] = (float)t0;
        r32[nv + vox] = (float)t1;
      } else {
        r64[vox] = t0;
        r64[nv + vox] = t1;
      }
      acc += t0 * t0 + t1 * t1;
    } else {
      y[vox] = y0;
      y[nv + vox] = y1;
      if (mode == 1) acc += (double)p[vox] * (double)y0 + (double)p[nv + vox] * (double)y1;
    }
  }
};

// Stockham stages S..SEND-1 in place (load, barrier, compute + store, barrier); NC pencils
// per thread; inactive threads take part in the barriers only.
template <typename T, class FP, int DIR, int S, int SEND, int NC, class Acc>
__device__ __forceinline__ void stages_ip(int t, bool act, Acc&& acc, const cpx<T>* twl) {
  if constexpr (S < SEND) {
    FxStage<T, FP, S, DIR> st[NC];
    if (act) {
#pragma unroll
      for (int c = 0; c < NC; ++c) st[c].load(t, [&](int j) { return acc(j, c); });
    }
    __syncthreads();
    if (act) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        st[c].compute(t, twl);
        st[c].store(t, [&](int j, cpx<T> v) { acc(j, c) = v; });
      }
    }
    __syncthreads();
    stages_ip<T, FP, DIR, S + 1, SEND, NC>(t, act, acc, twl);
  }
}

template <typename T, bool LDS>
__device__ __forceinline__ cpx<T>* spec_ptr(cpx<T>* gspec) {
  if constexpr (LDS) {
    static_assert(sizeof(T) == 4, "the LDS spectrum is single precision");
    return lds_spec();
  } else {
    return opaque_ptr(gspec);
  }
}

// ---- pass 1: coefficient multiply, forward transform along axis 1, split -------------------
template <typename T, class FP, bool LDS>
__device__ __forceinline__ void rows_fwd(const Task& tk, cpx<T>* gspec, const InOp<T>& in) {
  constexpr int L = FP::L, TPP = FP::TPP, H = L / 2 + 1, G = NT / TPP, NST = FP::NST;
  static_assert(NST >= 2, "at least two stages");
  cpx<T>* spec = spec_ptr<T, LDS>(gspec);
  const cpx<T>* twl = lds_tw<T>(1);
  const int tid = opaque_tid(), g = tid / TPP, t = tid - g * TPP;
  for (int base = 0; base < tk.n0; base += G) {
    const int i0 = base + g;
    const bool act = g < G && i0 < tk.n0;
    cpx<T>* slot = spec + (long long)(act ? i0 : 0) * (2 * H);
    {
      FxStage<T, FP, 0, -1> s0;
      if (act) {
        s0.load(t, [&](int j) { return in(i0 * L + j); });
        s0.compute(t, twl);
        s0.store(t, [&](int j, cpx<T> v) { slot[j] = v; });
      }
    }
    __syncthreads();
    stages_ip<T, FP, -1, 1, NST, 1>(t, act, [&](int j, int) -> cpx<T>& { return slot[j]; }, twl);
    constexpr int KM = (H + TPP - 1) / TPP;
    cpx<T> U[KM], V[KM];
#pragma unroll
    for (int m = 0; m < KM; ++m) {
      const int k = t + m * TPP;
      if (act && k < H) {
        const cpx<T> a = slot[k];
        const cpx<T> b = conj(slot[k == 0 ? 0 : L - k]);
        U[m] = T(0.5) * (a + b);
        const cpx<T> d = a - b;
        V[m] = mk(T(0.5) * d.y, T(-0.5) * d.x);
      }
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < KM; ++m) {
      const int k = t + m * TPP;
      if (act && k < H) {
        slot[k] = U[m];
        slot[H + k] = V[m];
      }
    }
  }
  __syncthreads();
}

// G(xi) of the 2-D scalar problem times 1/(n0 n1): G1 = xi xi^T/|xi|^2 (primal),
// G2 = I - G1 off xi = 0 (dual); zero at xi = 0 and at Nyquist frequencies.
template <typename T>
__device__ __forceinline__ void apply_G2d(cpx<T>& v0, cpx<T>& v1, int k0, int k1, const Task& tk) {
  bool nyq = false;
  const int s0 = signed_freq(k0, tk.n0, nyq);
  const int s1 = signed_freq(k1, tk.n1, nyq);
  const T x0 = T(s0 * tk.iY0), x1 = T(s1 * tk.iY1);
  const T q = x0 * x0 + x1 * x1;
  if (nyq || q == T(0)) {
    v0 = mk(T(0), T(0));
    v1 = mk(T(0), T(0));
    return;
  }
  const T sc = T(tk.scale);
  const cpx<T> d = x0 * v0 + x1 * v1;
  const T f = sc / q;
  const cpx<T> w0 = (x0 * f) * d, w1 = (x1 * f) * d;
  if (tk.pd) {
    v0 = sc * v0 - w0;
    v1 = sc * v1 - w1;
  } else {
    v0 = w0;
    v1 = w1;
  }
}

// ---- pass 2: per half-spectrum column: forward, G, inverse along axis 0 ---------------------
// FUSE: G applied in registers between the last forward and the first inverse stage (both
// components held at once); otherwise the forward transforms are stored, G is applied in a
// sweep over the column and the inverse transforms follow (half the registers, used for
// double precision).
template <typename T, class FP, class FPR, bool LDS, bool FUSE>
__device__ __forceinline__ void cols_G(const Task& tk, cpx<T>* gspec) {
  static_assert(FP::TPP == FPR::TPP && FP::L == FPR::L, "schedules must match");
  static_assert(FP::R[FP::NST - 1] == FPR::R[0], "inverse must start with the last forward radix");
  constexpr int L = FP::L, TPP = FP::TPP, G = NT / TPP, NF = FP::NST, NI = FPR::NST;
  static_assert(NF >= 2 && NI >= 2, "at least two stages");
  cpx<T>* spec = spec_ptr<T, LDS>(gspec);
  const cpx<T>* twl = lds_tw<T>(0);
  const int H = tk.h, rs = 2 * H;
  const int tid = opaque_tid();
  const int g = tid % G, t = tid / G;  // consecutive threads: consecutive columns
  for (int base = 0; base < H; base += G) {
    const int k1 = base + g;
    const bool act = t < TPP && k1 < H;
    const int col = act ? k1 : 0;
    auto acc = [&](int j, int c) -> cpx<T>& { return spec[j * rs + c * H + col]; };
    if constexpr (FUSE) {
      stages_ip<T, FP, -1, 0, NF - 1, 2>(t, act, acc, twl);
      using SL = FxStage<T, FP, NF - 1, -1>;
      SL last[2];
      if (act) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          last[c].load(t, [&](int j) { return acc(j, c); });
          last[c].compute(t, twl);
        }
        // last forward stage outputs sit at k0 = i + j*p
#pragma unroll
        for (int m = 0; m < SL::M; ++m) {
          const int i = t + m * TPP;
          if (SL::valid(i)) {
#pragma unroll
            for (int j = 0; j < SL::R; ++j) apply_G2d<T>(last[0].v[m][j], last[1].v[m][j], i + j * SL::p, k1, tk);
          }
        }
      }
      FxStage<T, FPR, 0, +1> r0[2];
      if (act) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
#pragma unroll
          for (int m = 0; m < SL::M; ++m)
#pragma unroll
            for (int j = 0; j < SL::R; ++j) r0[c].v[m][j] = last[c].v[m][j];
          r0[c].compute(t, twl);
        }
      }
      __syncthreads();  // every load of the last forward stage is done
      if (act) {
#pragma unroll
        for (int c = 0; c < 2; ++c) r0[c].store(t, [&](int j, cpx<T> v) { acc(j, c) = v; });
      }
      __syncthreads();
      stages_ip<T, FPR, +1, 1, NI, 2>(t, act, acc, twl);
    } else {
      auto acc0 = [&](int j, int) -> cpx<T>& { return acc(j, 0); };
      auto acc1 = [&](int j, int) -> cpx<T>& { return acc(j, 1); };
      stages_ip<T, FP, -1, 0, NF, 1>(t, act, acc0, twl);
      stages_ip<T, FP, -1, 0, NF, 1>(t, act, acc1, twl);
      if (act) {
        for (int k0 = t; k0 < L; k0 += TPP) apply_G2d<T>(acc(k0, 0), acc(k0, 1), k0, k1, tk);
      }
      __syncthreads();
      stages_ip<T, FPR, +1, 0, NI, 1>(t, act, acc0, twl);
      stages_ip<T, FPR, +1, 0, NI, 1>(t, act, acc1, twl);
    }
  }
  __syncthreads();
}

// ---- pass 3: Hermitian rebuild, inverse transform along axis 1, output ---------------------
template <typename T, class FP, bool LDS>
__device__ __forceinline__ void rows_inv(const Task& tk, cpx<T>* gspec, OutOp<T>& out) {
  constexpr int L = FP::L, TPP = FP::TPP, H = L / 2 + 1, G = NT / TPP, NST = FP::NST;
  cpx<T>* spec = spec_ptr<T, LDS>(gspec);
  const cpx<T>* twl = lds_tw<T>(1);
  const int tid = opaque_tid(), g = tid / TPP, t = tid - g * TPP;
  for (int base = 0; base < tk.n0; base += G) {
    const int i0 = base + g;
    const bool act = g < G && i0 < tk.n0;
    cpx<T>* slot = spec + (long long)(act ? i0 : 0) * (2 * H);
    FxStage<T, FP, 0, +1> s0;
    if (act) {
      s0.load(t, [&](int j) {
        cpx<T> u, v;
        if (j < H) {
          u = slot[j];
          v = slot[H + j];
        } else {
          u = conj(slot[L - j]);
          v = conj(slot[H + L - j]);
        }
        return mk(u.x - v.y, u.y + v.x);
      });
    }
    __syncthreads();
    if (act) {
      s0.compute(t, twl);
      s0.store(t, [&](int j, cpx<T> v) { slot[j] = v; });
    }
    __syncthreads();
    stages_ip<T, FP, +1, 1, NST - 1, 1>(t, act, [&](int j, int) -> cpx<T>& { return slot[j]; }, twl);
    if (act) {
      FxStage<T, FP, NST - 1, +1> sl;
      sl.load(t, [&](int j) { return slot[j]; });
      sl.compute(t, twl);
      sl.store(t, [&](int j, cpx<T> v) { out(i0 * L + j, v.x, v.y); });
    }
  }
  __syncthreads();
}

// ---- the matrix-vector product on one task -------------------------------------------------
#define FH_BATCH_LENGTHS(X) \
  X(27, FxP27, FxP27r)      \
  X(45, FxP45, FxP45r)      \
  X(63, FxP63, FxP63r)      \
  X(81, FxP81, FxP81)       \
  X(105, FxP105, FxP105r)   \
  X(135, FxP135, FxP135r)

// Radix schedules: VAR 0 transforms rows with the forward schedule of each length and the
// columns forward / reversed; VAR 1 swaps the roles (rows reversed, columns reversed /
// forward). Both are valid for every length; which is faster depends on the mix of grid
// sizes, so the host picks one by timing (fh_batch_autotune).
constexpr int NVAR = 2;

template <typename T, bool LDS, int VAR>
__device__ __forceinline__ void matvec(const Task& tk, cpx<T>* gspec, const InOp<T>& in, OutOp<T>& out) {
  __syncthreads();  // operands written by other threads of the workgroup are complete
  switch (tk.n1) {
#define X(LEN, FP, FPR)                                        \
  case LEN:                                                    \
    if constexpr (VAR == 0)                                    \
      rows_fwd<T, FP, LDS>(tk, gspec, in);                     \
    else                                                       \
      rows_fwd<T, FPR, LDS>(tk, gspec, in);                    \
    break;
    FH_BATCH_LENGTHS(X)
#undef X
  }
  switch (tk.n0) {
#define X(LEN, FP, FPR)                                        \
  case LEN:                                                    \
    if constexpr (VAR == 0)                                    \
      cols_G<T, FP, FPR, LDS, (sizeof(T) == 4)>(tk, gspec);    \
    else                                                       \
      cols_G<T, FPR, FP, LDS, (sizeof(T) == 4)>(tk, gspec);    \
    break;
    FH_BATCH_LENGTHS(X)
#undef X
  }
  switch (tk.n1) {
#define X(LEN, FP, FPR)                                        \
  case LEN:                                                    \
    if constexpr (VAR == 0)                                    \
      rows_inv<T, FP, LDS>(tk, gspec, out);                    \
    else                                                       \
      rows_inv<T, FPR, LDS>(tk, gspec, out);                   \
    break;
    FH_BATCH_LENGTHS(X)
#undef X
  }
}

__global__ void k_batch_reset(int* counter);
template <int VAR>
__global__ void __launch_bounds__(NT) k_batch_solve(Args A);

// Host side (fh_batch.hip): allow `lds` bytes of dynamic LDS, and enqueue counter reset +
// persistent solve on `stream` (graph-capturable: no memset nodes).
hipError_t allow_lds(size_t lds);
hipError_t enqueue(const Args& A, int nslots, size_t lds, hipStream_t stream, int var);

}  // namespace batch
}  // namespace fh
