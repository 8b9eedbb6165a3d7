// Compile-time specialised FFT passes for the grid lengths the method uses.
//
// For a length L with a fixed radix schedule (e.g. 243 = 9*9*3) every index is a
// compile-time constant, loops unroll and butterflies live in registers:
//  * TPP threads own one pencil; in stage s a thread computes ceil((L/R_s)/TPP)
//    radix-R_s butterflies (Stockham, natural order output).
//  * The first stage reads HBM straight into registers and the last stage writes
//    registers straight to HBM; LDS is only used to exchange data between stages.
//  * In the projection pass the forward FFT along the last axis ends, the projection
//    G(xi) is applied in registers, and the inverse FFT starts with the same radix, so
//    the thread already holds its inputs (no LDS round trip around G).
#pragma once
#include "fh_kernels.hpp"

namespace fh {

// 9-point DFT as 3x3 with internal twiddles W9^{n2 k1}.
template <typename T, int DIR>
__device__ __forceinline__ void dft9(cpx<T> (&v)[9]) {
  cpx<T> a[3][3];
#pragma unroll
  for (int n2 = 0; n2 < 3; ++n2) {
    cpx<T> t3[3] = {v[n2], v[n2 + 3], v[n2 + 6]};
    dft<T, 3, DIR>(t3);
#pragma unroll
    for (int k1 = 0; k1 < 3; ++k1) a[n2][k1] = t3[k1];
  }
  // W9^m, m = 1, 2, 4: cos/sin(2 pi m / 9)
  const T c1 = T(0.76604444311897803520), s1 = T(0.64278760968653932632);
  const T c2 = T(0.17364817766693034885), s2 = T(0.98480775301220805936);
  const T c4 = T(-0.93969262078590838405), s4 = T(0.34202014332566873304);
  const T sg = DIR < 0 ? T(-1) : T(1);
  a[1][1] = cmul(a[1][1], mk(c1, sg * s1));
  a[1][2] = cmul(a[1][2], mk(c2, sg * s2));
  a[2][1] = cmul(a[2][1], mk(c2, sg * s2));
  a[2][2] = cmul(a[2][2], mk(c4, sg * s4));
#pragma unroll
  for (int k1 = 0; k1 < 3; ++k1) {
    cpx<T> t3[3] = {a[0][k1], a[1][k1], a[2][k1]};
    dft<T, 3, DIR>(t3);
#pragma unroll
    for (int k2 = 0; k2 < 3; ++k2) v[k1 + 3 * k2] = t3[k2];
  }
}

template <typename T, int R, int DIR>
__device__ __forceinline__ void dft_any(cpx<T> (&v)[R]) {
  if constexpr (R == 9)
    dft9<T, DIR>(v);
  else
    dft<T, R, DIR>(v);
}


// ---- buffer-resource addressing ------------------------------------------------------
// A buffer resource holds the 64-bit base in scalar registers; every access then needs
// only a 32-bit per-lane byte offset, and reads past num_records return 0 (so a ragged
// tile can load unconditionally). Word 3 = 0x00020000 is the raw-buffer format of the
// gfx9 family (gfx950 included).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, unsigned long long bytes) {
  const unsigned nb = bytes > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (unsigned)bytes;
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)nb, 0x00020000);
}

template <typename T>
__device__ __forceinline__ cpx<T> bload_c(__amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(cpx<T>, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  } else {
    return __builtin_bit_cast(cpx<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  }
}

template <typename T>
__device__ __forceinline__ void bstore_c(__amdgpu_buffer_rsrc_t r, unsigned off, cpx<T> v) {
  if constexpr (sizeof(T) == 4) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, off, 0, 0);
  } else {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, v), r, off, 0, 0);
  }
}

template <typename T>
__device__ __forceinline__ T bload_r(__amdgpu_buffer_rsrc_t r, unsigned off) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  }
}

template <typename T>
__device__ __forceinline__ void bstore_r(__amdgpu_buffer_rsrc_t r, unsigned off, T v) {
  if constexpr (sizeof(T) == 4) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b32(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(V, v), r, off, 0, 0);
  } else {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), r, off, 0, 0);
  }
}

template <int L_, int TPP_, int... Rs>
struct FxPlan {
  static constexpr int L = L_;
  static constexpr int TPP = TPP_;
  static constexpr int NST = sizeof...(Rs);
  static constexpr int R[NST] = {Rs...};
};

template <class FP>
constexpr int fx_pre(int s) {
  int p = 1;
  for (int i = 0; i < s; ++i) p *= FP::R[i];
  return p;
}

template <typename T, class FP, int S, int DIR>
struct FxStage {
  static constexpr int R = FP::R[S];
  static constexpr int p = fx_pre<FP>(S);
  static constexpr int B = FP::L / R;
  static constexpr int TPP = FP::TPP;
  static constexpr int M = (B + TPP - 1) / TPP;
  static constexpr bool EXACT = (B % TPP) == 0;
  cpx<T> v[M][R];

  __device__ __forceinline__ static bool valid(int i) { return EXACT || i < B; }

  template <class Ld>
  __device__ __forceinline__ void load(int t, Ld&& ld) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int i = t + m * TPP;
      if (valid(i)) {
#pragma unroll
        for (int j = 0; j < R; ++j) v[m][j] = ld(i + j * B);
      }
    }
  }

  __device__ __forceinline__ void compute(int t, const cpx<T>* __restrict__ twl) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int i = t + m * TPP;
      if (valid(i)) {
        if constexpr (p > 1) {
          const int k = i % p;
          constexpr int st = FP::L / (p * R);
#pragma unroll
          for (int j = 1; j < R; ++j) {
            cpx<T> w = twl[j * k * st];
            if (DIR > 0) w.y = -w.y;
            v[m][j] = cmul(v[m][j], w);
          }
        }
        dft_any<T, R, DIR>(v[m]);
      }
    }
  }

  template <class St>
  __device__ __forceinline__ void store(int t, St&& st) {
#pragma unroll
    for (int m = 0; m < M; ++m) {
      const int i = t + m * TPP;
      if (valid(i)) {
        const int k = i % p;
        const int base = (i - k) * R + k;
#pragma unroll
        for (int j = 0; j < R; ++j) st(base + j * p, v[m][j]);
      }
    }
  }
};

// Stages S..SEND-1, LDS to LDS, for NC pencils per thread (accessor lds(j, c)).
// Barriers are uniform: every thread runs every stage.
template <typename T, class FP, int DIR, int S, int SEND, int NC, class Lds>
__device__ __forceinline__ void fx_lds_stages(int t, Lds&& lds, const cpx<T>* __restrict__ twl) {
  if constexpr (S < SEND) {
    FxStage<T, FP, S, DIR> st[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) st[c].load(t, [&](int j) { return lds(j, c); });
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      st[c].compute(t, twl);
      st[c].store(t, [&](int j, cpx<T> v) { lds(j, c) = v; });
    }
    __syncthreads();
    fx_lds_stages<T, FP, DIR, S + 1, SEND, NC>(t, lds, twl);
  }
}

// Stage sizing helpers for launch shapes (host and device).
template <class FP, typename T>
struct FxShape {
  // columns per strided tile and frequency points per projection tile (fp32 with 16 or more
  // threads per pencil, i.e. lengths 105 and 243: 4 points x 3 components keep a 243-point tile
  // at 23 KB of LDS, six workgroups per CU, so one workgroup's loads overlap the others'
  // transforms; 8 points measured 12 % slower at 243^3)
  static constexpr int P = (sizeof(T) == 4) ? (FP::TPP >= 16 ? 16 : (FP::TPP >= 8 ? 32 : 64))
                                            : (FP::TPP >= 16 ? 8 : (FP::TPP >= 8 ? 16 : 32));
  static constexpr int F = (sizeof(T) == 4) ? (FP::TPP >= 16 ? 4 : (FP::TPP >= 8 ? 16 : 32))
                                            : (FP::TPP >= 16 ? 4 : (FP::TPP >= 8 ? 8 : 16));
  // rows per workgroup in the contiguous-axis passes (even)
  static constexpr int RPB = (FP::TPP >= 16) ? 8 : (FP::TPP >= 8 ? 16 : 32);
  // Schedule variants a plan can select (fh_plan_set_schedule / fh_plan_autotune); variant 0 is
  // the default. Projection pass: frequency points per tile (TPP*F*NC threads, so F also sets how
  // full the last wavefront of a workgroup is); strided passes: columns per tile.
  static constexpr int NPROJ = 4, NCOLS = 2;
  static constexpr int proj_f(int v) {
    constexpr int c32a[NPROJ] = {4, 6, 5, 7}, c32b[NPROJ] = {16, 8, 12, 24}, c32c[NPROJ] = {32, 16, 24, 48};
    constexpr int c64a[NPROJ] = {4, 2, 3, 6}, c64b[NPROJ] = {8, 4, 6, 12}, c64c[NPROJ] = {16, 8, 12, 24};
    if (sizeof(T) == 4) return FP::TPP >= 16 ? c32a[v] : (FP::TPP >= 8 ? c32b[v] : c32c[v]);
    return FP::TPP >= 16 ? c64a[v] : (FP::TPP >= 8 ? c64b[v] : c64c[v]);
  }
  static constexpr int cols_p(int v) { return v == 0 ? P : (P >= 2 ? P / 2 : P); }
  // contiguous-axis passes: rows per workgroup (even); variant 1 takes 3/2 as many rows in single
  // precision and 1/2 as many in double, variant 2 3/4 as many, which also changes how full the
  // last wavefront is (TPP * (rpb/2) * NC threads: 243 points, 3 components: 8 rows 324 threads
  // in 6 wavefronts, 12 rows 486 in 8, 6 rows 243 in 4, 4 rows 162 in 3) and how many workgroups
  // share a CU's LDS
  static constexpr int NROWS = 3;
  static constexpr int rows_rpb(int v) {
    return v == 0 ? RPB : (v == 1 ? (sizeof(T) == 4 ? RPB / 2 * 3 : RPB / 2) : RPB / 4 * 3);
  }
};

// ---- strided pass: forward (DIR -1) or inverse (DIR +1) along a non-contiguous axis ----
// total: number of complex values in the array (for the buffer range); all byte
// offsets of one launch fit in 32 bits (checked by the launcher).
template <typename T, class FP, int P, int DIR>
__global__ void __launch_bounds__(FP::TPP * P)
k_cols_fx(cpx<T>* __restrict__ data, long long ostride, long long inner, int ntiles, long long total,
          const cpx<T>* __restrict__ twg, const int* __restrict__ halt, int rev, CgFin fin) {
  if (halt && *halt) return;
  if (fin.st && blockIdx.x == 0) cg_fin_block(fin);
  constexpr int L = FP::L;
  constexpr unsigned CS = sizeof(cpx<T>);
  __shared__ cpx<T> s[L * P];
  __shared__ cpx<T> twl[L];
  const int tid = threadIdx.x;
  const int q = tid % P, t = tid / P;
  const int lin0 = xcd_remap(blockIdx.x, gridDim.x);
  const int lin = rev ? (int)gridDim.x - 1 - lin0 : lin0;  // rev: newest data of the previous pass first
  const int o = lin / ntiles, tile = lin - o * ntiles;
  const long long col0 = (long long)tile * P;
  const long long start = (long long)o * ostride + col0;
  const int ncol = (int)min((long long)P, inner - col0);
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(data + start, (unsigned long long)(total - start) * CS);
  const unsigned rowb = (unsigned)inner * CS;
  const unsigned qb = (unsigned)q * CS;
  for (int e = tid; e < L; e += FP::TPP * P) twl[e] = twg[e];
  const bool act = q < ncol;
  FxStage<T, FP, 0, DIR> s0;
  s0.load(t, [=](int j) { return bload_c<T>(rs, (unsigned)j * rowb + qb); });
  s0.compute(t, twl);
  if constexpr (FP::NST == 1) {
    if (act) s0.store(t, [=](int j, cpx<T> v) { bstore_c<T>(rs, (unsigned)j * rowb + qb, v); });
  } else {
    s0.store(t, [&](int j, cpx<T> v) { s[j * P + q] = v; });
    __syncthreads();
    auto at = [&](int j, int) -> cpx<T>& { return s[j * P + q]; };
    fx_lds_stages<T, FP, DIR, 1, FP::NST - 1, 1>(t, at, twl);
    FxStage<T, FP, FP::NST - 1, DIR> sl;
    sl.load(t, [&](int j) { return s[j * P + q]; });
    sl.compute(t, twl);
    if (act) sl.store(t, [=](int j, cpx<T> v) { bstore_c<T>(rs, (unsigned)j * rowb + qb, v); });
  }
}

// 1/x: hardware reciprocal (1 ulp) in single precision, IEEE division in double.
template <typename T>
__device__ __forceinline__ T recip(T x) {
  if constexpr (sizeof(T) == 4)
    return __builtin_amdgcn_rcpf(x);
  else
    return T(1) / x;
}

// G(xi) on NC registers (compile-time component count; same formulas as apply_G).
template <typename T, int NC>
__device__ __forceinline__ void apply_G_fx(cpx<T> (&v)[NC], int k0, int k1, int k2, const GParams& g) {
  bool nyq = false;
  const int s0 = signed_freq(k0, g.n0, nyq);
  const int s1 = signed_freq(k1, g.n1, nyq);
  if ((g.n2 & 1) == 0 && 2 * k2 == g.n2) nyq = true;
  if (g.banded && out_of_band(s0, s1, k2, g)) nyq = true;
  const T x0 = T(s0 * g.iY[0]), x1 = T(s1 * g.iY[1]), x2 = T(k2 * g.iY[2]);
  const T sc = T(g.scale);
  if constexpr (NC == 2 || NC == 3) {
    if (g.physics == 0) {
      // scalar problem: NC = dim; 2-D uses the axes (1, 2) of the 3-D view
      T xi[NC];
      if constexpr (NC == 3) {
        xi[0] = x0; xi[1] = x1; xi[2] = x2;
      } else {
        xi[0] = x1; xi[1] = x2;
      }
      T q = T(0);
#pragma unroll
      for (int c = 0; c < NC; ++c) q += xi[c] * xi[c];
      if (nyq || q == T(0)) {
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = mk(T(0), T(0));
        return;
      }
      cpx<T> dot = mk(T(0), T(0));
#pragma unroll
      for (int c = 0; c < NC; ++c) dot = dot + xi[c] * v[c];
      const T f = sc * recip(q);
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const cpx<T> w = (xi[c] * f) * dot;
        v[c] = g.dual ? (sc * v[c] - w) : w;
      }
      return;
    }
  }
  // elasticity (NC = 3: 2-D Mandel, NC = 6: 3-D Mandel)
  if constexpr (NC == 3 || NC == 6) {
    constexpr int d = NC == 6 ? 3 : 2;
    T xi[3];
    if constexpr (d == 3) {
      xi[0] = x0; xi[1] = x1; xi[2] = x2;
    } else {
      xi[0] = x1; xi[1] = x2; xi[2] = T(0);
    }
    T q = T(0);
#pragma unroll
    for (int a = 0; a < d; ++a) q += xi[a] * xi[a];
    if (nyq || q == T(0)) {
#pragma unroll
      for (int c = 0; c < NC; ++c) v[c] = mk(T(0), T(0));
      return;
    }
    const T r2 = T(1.41421356237309504880), ir2 = T(0.70710678118654752440);
    cpx<T> S[3][3];
    if constexpr (d == 3) {
      S[0][0] = v[0]; S[1][1] = v[1]; S[2][2] = v[2];
      S[1][2] = S[2][1] = ir2 * v[3];
      S[0][2] = S[2][0] = ir2 * v[4];
      S[0][1] = S[1][0] = ir2 * v[5];
    } else {
      S[0][0] = v[0]; S[1][1] = v[1];
      S[0][1] = S[1][0] = ir2 * v[2];
    }
    cpx<T> tv[3];
    cpx<T> qq = mk(T(0), T(0));
#pragma unroll
    for (int i = 0; i < d; ++i) {
      tv[i] = mk(T(0), T(0));
#pragma unroll
      for (int j = 0; j < d; ++j) tv[i] = tv[i] + xi[j] * S[i][j];
      qq = qq + xi[i] * tv[i];
    }
    const T iq = recip(q), iq2 = iq * iq;
    auto gam = [&](int i, int j) -> cpx<T> {
      return iq * (xi[i] * tv[j] + xi[j] * tv[i]) - (xi[i] * xi[j] * iq2) * qq;
    };
    cpx<T> o[NC];
    if constexpr (d == 3) {
      o[0] = gam(0, 0); o[1] = gam(1, 1); o[2] = gam(2, 2);
      o[3] = r2 * gam(1, 2); o[4] = r2 * gam(0, 2); o[5] = r2 * gam(0, 1);
    } else {
      o[0] = gam(0, 0); o[1] = gam(1, 1); o[2] = r2 * gam(0, 1);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = g.dual ? sc * (v[c] - o[c]) : sc * o[c];
  }
}

// ---- projection pass: forward FFT, G(xi) sweep, inverse FFT along the last strided axis ----
// FP: forward schedule; FPR: the same radices reversed (inverse schedule). A tile is F
// consecutive frequency points of the (k1, k2) plane with all NC components (the AoS
// spectrum's contiguous run); a thread owns one (frequency point, component) pencil, so a
// tile uses TPP*F*NC threads (three times the waves of one thread per frequency point at the
// same LDS); G is applied by a sweep over the tile in LDS between the transforms.
template <typename T, class FP, class FPR, int F, int NC>
__global__ void __launch_bounds__(FP::TPP * F * NC)
k_cols_g_sw(cpx<T>* __restrict__ data, long long inner, int ntiles, long long total,
            const cpx<T>* __restrict__ twg, GParams g, int n2h, const int* __restrict__ halt, CgFin fin) {
  if (halt && *halt) return;
  if (fin.st && blockIdx.x == 0) cg_fin_block(fin);
  constexpr int L = FP::L, P = F * NC, TPP = FP::TPP, NT = TPP * P;
  constexpr int NI = FPR::NST;
  constexpr unsigned CS = sizeof(cpx<T>);
  __shared__ cpx<T> s[L * P];
  __shared__ cpx<T> twl[L];
  const int tid = threadIdx.x;
  const int q = tid % P, t = tid / P;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const long long col0 = (long long)tile * P;
  const int ncol = (int)min((long long)P, inner - col0);
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(data + col0, (unsigned long long)(total - col0) * CS);
  const unsigned rowb = (unsigned)inner * CS;
  const unsigned qb = (unsigned)q * CS;
  FxStage<T, FP, 0, -1> s0;
  s0.load(t, [=](int j) { return bload_c<T>(rs, (unsigned)j * rowb + qb); });
  for (int e = tid; e < L; e += NT) twl[e] = twg[e];
  __syncthreads();
  s0.compute(t, twl);
  auto at = [&](int j, int) -> cpx<T>& { return s[j * P + q]; };
  s0.store(t, [&](int j, cpx<T> v) { s[j * P + q] = v; });
  __syncthreads();
  fx_lds_stages<T, FP, -1, 1, FP::NST, 1>(t, at, twl);
  // G(xi) at every (k0, frequency point) of the tile; NT is a multiple of F, so a thread
  // keeps one frequency point f and walks k0
  static_assert(NT % F == 0, "sweep layout");
  const int f = tid % F;
  const int fp = (int)(col0 / NC) + f;
  int ka, kb;
  if (g.n0 > 1) {
    ka = fp / n2h;
    kb = fp - ka * n2h;
  } else {
    ka = 0;
    kb = fp;
  }
  for (int k0 = tid / F; k0 < L; k0 += NT / F) {
    if (f * NC < ncol) {
      cpx<T> v[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) v[c] = s[k0 * P + f * NC + c];
      if (g.n0 > 1)
        apply_G_fx<T, NC>(v, k0, ka, kb, g);
      else
        apply_G_fx<T, NC>(v, 0, k0, kb, g);
#pragma unroll
      for (int c = 0; c < NC; ++c) s[k0 * P + f * NC + c] = v[c];
    }
  }
  __syncthreads();
  fx_lds_stages<T, FPR, +1, 0, NI - 1, 1>(t, at, twl);
  FxStage<T, FPR, NI - 1, +1> sl;
  sl.load(t, [&](int j) { return s[j * P + q]; });
  sl.compute(t, twl);
  if (q < ncol) sl.store(t, [=](int j, cpx<T> v) { bstore_c<T>(rs, (unsigned)j * rowb + qb, v); });
}

// value of component c of A(x) v at voxel vox, for the contiguous-axis pass; byte
// offsets: vb = vox*sizeof(T), nb = n*sizeof(T) (component / coefficient stride)
// pv(d, vb) gives component d of the operand at that voxel.
template <typename T, int COEF, int NC, class PV>
__device__ __forceinline__ T coef_apply(PV&& pv, __amdgpu_buffer_rsrc_t rco, unsigned vb, int c, unsigned nb) {
  if constexpr (COEF == COEF_NONE) {
    return pv(c, vb);
  } else if constexpr (COEF == COEF_ISO) {
    return bload_r<T>(rco, vb) * pv(c, vb);
  } else if constexpr (COEF == COEF_LAME) {
    constexpr int dn = (NC == 6) ? 3 : 2;
    T tr = T(0);
#pragma unroll
    for (int d = 0; d < dn; ++d) tr += pv(d, vb);
    return T(2) * bload_r<T>(rco, nb + vb) * pv(c, vb) + (c < dn ? bload_r<T>(rco, vb) * tr : T(0));
  } else {
    T acc = T(0);
#pragma unroll
    for (int d = 0; d < NC; ++d) acc += bload_r<T>(rco, (unsigned)(c * NC + d) * nb + vb) * pv(d, vb);
    return acc;
  }
}

// ---- pass 1 (fixed length): A(x) v, then R2C along the contiguous axis -------------
// Workgroup: rpb rows (all components), (rpb/2)*nc complex pencils of paired real rows,
// TPP threads per pencil (blockDim = (rpb/2)*nc*TPP). Output: AoS half spectrum.
// UPD = 1: `in` is the CG direction p, updated in place first as p = r + beta p
// (beta from the CG state); the operator is then applied to the new p.
// UPD = 2 (fused iteration, COEF_ISO/NONE): as UPD = 1, and before p is overwritten the
// pending step of the previous iteration is added to the correction, xc += alpha p_old
// (when st->pend); per-workgroup partial sums of <p, a p> go to partial[blockIdx.x]
// (for p in the range of G this is <p, G(a p)>, the CG denominator).
template <typename T, class FP, int COEF, int NC, int UPD>
__global__ void __launch_bounds__(1024)
k_rows_fwd_fx(const T* __restrict__ in, const T* __restrict__ coef, cpx<T>* __restrict__ out,
              long long nrows, long long n, int rpb, const cpx<T>* __restrict__ twg,
              const int* __restrict__ halt, const T* __restrict__ rvec, const CGState* __restrict__ st,
              T* __restrict__ xc, double* __restrict__ partial) {
  constexpr int nc = NC;
  if (halt && *halt) return;
  constexpr int L = FP::L, n2h = L / 2 + 1, TPP = FP::TPP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* s = reinterpret_cast<cpx<T>*>(smem_raw);
  __shared__ cpx<T> twl[L];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int t = tid % TPP, pen = tid / TPP;
  const long long row0 = (long long)xcd_remap(blockIdx.x, gridDim.x) * rpb;
  const int nr = (int)min((long long)rpb, nrows - row0);
  for (int e = tid; e < L; e += nthr) twl[e] = twg[e];
  const int rp = pen / nc, m = pen - rp * nc;
  const int rra = 2 * m, rrb = 2 * m + 1;
  const int rowa = 2 * rp + rra / nc, ca = rra % nc;
  const int rowb = 2 * rp + rrb / nc, cb = rrb % nc;
  const bool va = rowa < nr, vb = rowb < nr;
  constexpr unsigned SR = sizeof(T);
  const unsigned nbytes = (unsigned)n * SR;
  constexpr int kvals = COEF == COEF_NONE ? 0 : (COEF == COEF_ISO ? 1 : (COEF == COEF_LAME ? 2 : NC * NC));
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(in, (unsigned long long)NC * n * SR);
  const __amdgpu_buffer_rsrc_t rco = make_rsrc(coef ? coef : in, (unsigned long long)kvals * n * SR);
  const unsigned ba = (unsigned)((row0 + rowa) * L) * SR, bb = (unsigned)((row0 + rowb) * L) * SR;
  FxStage<T, FP, 0, -1> s0;
  double acc = 0.0;
  if constexpr (UPD == 2) {
    static_assert(COEF == COEF_ISO || COEF == COEF_NONE, "fused iteration: one component per thread");
    const T bta = T(st->beta);
    const bool fold = st->pend != 0;
    const T alp = T(st->alpha);
    const __amdgpu_buffer_rsrc_t rr = make_rsrc(rvec, (unsigned long long)NC * n * SR);
    const __amdgpu_buffer_rsrc_t rx = make_rsrc(xc, (unsigned long long)NC * n * SR);
    auto one = [&](int c, unsigned o) -> T {
      const unsigned oc = (unsigned)c * nbytes + o;
      const T pold = bload_r<T>(rin, oc);
      const T pn = bload_r<T>(rr, oc) + bta * pold;
      if (fold) bstore_r<T>(rx, oc, bload_r<T>(rx, oc) + alp * pold);
      bstore_r<T>(rin, oc, pn);
      const T y = (COEF == COEF_ISO) ? bload_r<T>(rco, o) * pn : pn;
      acc += (double)y * (double)pn;
      return y;
    };
    s0.load(t, [&](int x) {
      const T ya = va ? one(ca, ba + x * SR) : T(0);
      const T yb = vb ? one(cb, bb + x * SR) : T(0);
      return mk(ya, yb);
    });
  } else if constexpr (UPD == 0) {
    auto pv = [=](int d, unsigned vb_) { return bload_r<T>(rin, (unsigned)d * nbytes + vb_); };
    s0.load(t, [=](int x) {
      const T ya = va ? coef_apply<T, COEF, NC>(pv, rco, ba + x * SR, ca, nbytes) : T(0);
      const T yb = vb ? coef_apply<T, COEF, NC>(pv, rco, bb + x * SR, cb, nbytes) : T(0);
      return mk(ya, yb);
    });
  } else {
    const T bta = T(st->beta);
    const __amdgpu_buffer_rsrc_t rr = make_rsrc(rvec, (unsigned long long)NC * n * SR);
    auto pv = [=](int d, unsigned vb_) {
      const unsigned o = (unsigned)d * nbytes + vb_;
      return bload_r<T>(rr, o) + bta * bload_r<T>(rin, o);
    };
    s0.load(t, [=](int x) {
      T ya = T(0), yb = T(0);
      if (va) {
        const unsigned o = ba + x * SR;
        ya = coef_apply<T, COEF, NC>(pv, rco, o, ca, nbytes);
        bstore_r<T>(rin, (unsigned)ca * nbytes + o, pv(ca, o));
      }
      if (vb) {
        const unsigned o = bb + x * SR;
        yb = coef_apply<T, COEF, NC>(pv, rco, o, cb, nbytes);
        bstore_r<T>(rin, (unsigned)cb * nbytes + o, pv(cb, o));
      }
      return mk(ya, yb);
    });
  }
  s0.compute(t, twl);
  cpx<T>* sp = s + pen * L;
  s0.store(t, [&](int j, cpx<T> v) { sp[j] = v; });
  __syncthreads();
  auto at = [&](int j, int) -> cpx<T>& { return sp[j]; };
  fx_lds_stages<T, FP, -1, 1, FP::NST, 1>(t, at, twl);
  // untangle the paired rows and store the AoS half spectrum (contiguous per block)
  const int per = n2h * nc;
  cpx<T>* ob = out + row0 * per;
  for (int e = tid; e < nr * per; e += nthr) {
    const int r = e / per;
    const int rem = e - r * per;
    const int k = rem / nc, c = rem - k * nc;
    const int rr = (r & 1) * nc + c;
    const cpx<T>* z = s + ((r >> 1) * nc + (rr >> 1)) * L;
    const cpx<T> Z = z[k];
    const cpx<T> Zc = conj(z[k ? L - k : 0]);
    cpx<T> Y;
    if (rr & 1) {
      const cpx<T> d = Z - Zc;
      Y = mk(T(0.5) * d.y, T(-0.5) * d.x);
    } else {
      Y = T(0.5) * (Z + Zc);
    }
    ob[e] = Y;
  }
  if constexpr (UPD == 2) {
    const double bsum = block_sum(acc);
    if (tid == 0) partial[blockIdx.x] = bsum;
  }
}

// Radix schedules of the specialised lengths (forward / reversed for the inverse).
using FxP5 = FxPlan<5, 1, 5>;
using FxP9 = FxPlan<9, 1, 9>;
using FxP15 = FxPlan<15, 5, 5, 3>;
using FxP15r = FxPlan<15, 5, 3, 5>;
using FxP25 = FxPlan<25, 5, 5, 5>;
using FxP27 = FxPlan<27, 3, 9, 3>;
using FxP27r = FxPlan<27, 3, 3, 9>;
using FxP45 = FxPlan<45, 9, 9, 5>;
using FxP45r = FxPlan<45, 9, 5, 9>;
using FxP63 = FxPlan<63, 9, 9, 7>;
using FxP63r = FxPlan<63, 9, 7, 9>;
using FxP75 = FxPlan<75, 15, 5, 5, 3>;
using FxP75r = FxPlan<75, 15, 3, 5, 5>;
using FxP81 = FxPlan<81, 9, 9, 9>;
using FxP99 = FxPlan<99, 11, 9, 11>;
using FxP99r = FxPlan<99, 11, 11, 9>;
using FxP105 = FxPlan<105, 21, 7, 5, 3>;
using FxP105r = FxPlan<105, 21, 3, 5, 7>;
using FxP125 = FxPlan<125, 25, 5, 5, 5>;
using FxP135 = FxPlan<135, 15, 9, 3, 5>;
using FxP135r = FxPlan<135, 15, 5, 3, 9>;
using FxP225 = FxPlan<225, 25, 9, 5, 5>;
using FxP225r = FxPlan<225, 25, 5, 5, 9>;
using FxP243 = FxPlan<243, 27, 9, 9, 3>;
using FxP243r = FxPlan<243, 27, 3, 9, 9>;

}  // namespace fh
