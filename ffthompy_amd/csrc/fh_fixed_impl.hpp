// Definitions of the fixed-length launchers for one precision FH_T (included by
// fh_fixed_f32.hip / fh_fixed_f64.hip).
#pragma once
#include "fh_fixed.hpp"
#include "fh_launch.hpp"
#include "fh_rows_st.hpp"
#include "fh_pot_fx.hpp"

namespace fh {

// length -> (forward schedule, reversed schedule)
#define FH_FIXED_LENGTHS(X)   \
  X(5, FxP5, FxP5)            \
  X(9, FxP9, FxP9)            \
  X(15, FxP15, FxP15r)        \
  X(25, FxP25, FxP25)         \
  X(27, FxP27, FxP27r)        \
  X(45, FxP45, FxP45r)        \
  X(63, FxP63, FxP63r)        \
  X(75, FxP75, FxP75r)        \
  X(81, FxP81, FxP81)         \
  X(99, FxP99, FxP99r)        \
  X(105, FxP105, FxP105r)     \
  X(125, FxP125, FxP125)      \
  X(135, FxP135, FxP135r)     \
  X(225, FxP225, FxP225r)     \
  X(243, FxP243, FxP243r)

// the buffer addressing of the fixed kernels uses 32-bit byte offsets
inline bool offsets_fit(long long elems, size_t esize) { return (unsigned long long)elems * esize < 0xFFFFFF00ull; }

// Traversal order of a pass: 1 = workgroups walk the tiles from the end, so the data the
// previous pass wrote last (most likely still in L2 / the Infinity Cache) is read first.
// Forward strided pass (after the forward row pass) and the inverse row pass run reversed.
inline int order_rev(int which) { return which == 0 || which == 1; }

template <class K>
inline void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024) (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)bytes);
}

// RPB_: rows per workgroup. ALL: every (coefficient, component) combination is instantiated;
// otherwise only the scalar ones (no / isotropic coefficient, 2 or 3 components) and the others
// return 0 (the caller then launches the default row count).
template <typename T, class FP, int RPB_, bool ALL>
int rows_fwd_fx_launch_r(const LaunchCtx& c, int nc, const T* in, const T* coef, int kind, cpx<T>* out,
                         const int* halt, const T* rvec, const CGState* st, T* xc, double* partial,
                         size_t partial_cap, int* nblocks) {
  constexpr int rpb = RPB_;
  if (!ALL && !((nc == 2 || nc == 3) && (kind == COEF_NONE || kind == COEF_ISO))) return 0;
  const int npen = (rpb / 2) * nc;
  const int nthr = npen * FP::TPP;
  if (nthr > 1024) return 0;
  const size_t lds = (size_t)npen * FP::L * sizeof(cpx<T>);
  if (lds + FP::L * sizeof(cpx<T>) > 160 * 1024) return 0;
  const int kv = kind == COEF_NONE ? 0 : (kind == COEF_ISO ? 1 : (kind == COEF_LAME ? 2 : nc * nc));
  if (!offsets_fit(c.n * (nc > kv ? nc : kv), sizeof(T))) return 0;
  const long long nrows = (long long)c.n0 * c.n1;
  const long long nb = (nrows + rpb - 1) / rpb;
  const dim3 grid((unsigned)nb), block(nthr);
  const cpx<T>* tw = reinterpret_cast<const cpx<T>*>(c.tw[2]);
#define FH_LAUNCH_RF(KIND, NCC)                                                                       \
  do {                                                                                                \
    if (rvec) {                                                                                       \
      auto kf = k_rows_fwd_fx<T, FP, KIND, NCC, 1>;                                                   \
      allow_lds(kf, lds);                                                                             \
      hipExtLaunchKernelGGL(kf, grid, block, lds, c.stream, c.t0, c.t1, 0, in, coef, out, nrows, c.n, rpb, tw, halt, rvec, \
                         st, xc, partial);                                                            \
    } else {                                                                                          \
      auto kf = k_rows_fwd_fx<T, FP, KIND, NCC, 0>;                                                   \
      allow_lds(kf, lds);                                                                             \
      hipExtLaunchKernelGGL(kf, grid, block, lds, c.stream, c.t0, c.t1, 0, in, coef, out, nrows, c.n, rpb, tw, halt, rvec, \
                         st, xc, partial);                                                            \
    }                                                                                                 \
  } while (0)
  // staged kernels (16-byte HBM chunks through LDS) in single precision, and in double
  // precision for the smaller row counts of variants 1 and 2 (at the default row count their
  // doubled LDS footprint halves the occupancy and the direct kernel, which also serves every
  // coefficient kind, is as fast). The staged kernels finish every direction update before any
  // coefficient product, so the fused update also serves the Lame coefficient there.
  const bool lame_ok = kind == COEF_LAME && (nc == 3 || nc == 6);
  if ((kind == COEF_ISO || kind == COEF_NONE || lame_ok) && (sizeof(T) == 4 || !ALL)) {
    const int upd = xc ? 2 : (rvec ? 1 : 0);
    if (upd == 2 && (!rvec || kind == COEF_NONE || (size_t)nb > partial_cap)) return 0;
#define FH_LAUNCH_ST(KIND, NCC, UPDV)                                                                         \
  do {                                                                                                        \
    const size_t slds = RowStage<FP, T, NCC>::lds_bytes(rpb, KIND == COEF_ISO ? 1 : (KIND == COEF_LAME ? 2 : 0)); \
    if (slds + FP::L * sizeof(cpx<T>) > 160 * 1024) return 0;                                                 \
    if (!offsets_fit(nrows * (FP::L / 2 + 1) * NCC * 2, sizeof(T))) return 0;                                 \
    auto kf = k_rows_fwd_st<T, FP, KIND, NCC, UPDV, RPB_>;                                                    \
    allow_lds(kf, slds);                                                                                      \
    hipExtLaunchKernelGGL(kf, grid, block, slds, c.stream, c.t0, c.t1, 0, in, coef, out, nrows, c.n, rpb, tw, halt, rvec, st, xc, \
                       partial, c.bpart, c.bnp);                                                              \
  } while (0)
#define FH_ST_NC(KIND, UPDV)                 \
  do {                                       \
    switch (nc) {                            \
      case 2: FH_LAUNCH_ST(KIND, 2, UPDV); break; \
      case 3: FH_LAUNCH_ST(KIND, 3, UPDV); break; \
      case 1: if constexpr (ALL) { FH_LAUNCH_ST(KIND, 1, UPDV); break; } else return 0; \
      case 6: if constexpr (ALL) { FH_LAUNCH_ST(KIND, 6, UPDV); break; } else return 0; \
      default: return 0;                     \
    }                                        \
  } while (0)
    if (kind == COEF_NONE) {
      if (upd == 0) FH_ST_NC(COEF_NONE, 0);
      else if (upd == 1) FH_ST_NC(COEF_NONE, 1);
      else return 0;
    } else if (kind == COEF_LAME) {
      if constexpr (!ALL) return 0;
#define FH_ST_LAME(UPDV)                                \
  do {                                                  \
    if (nc == 3) FH_LAUNCH_ST(COEF_LAME, 3, UPDV);      \
    else FH_LAUNCH_ST(COEF_LAME, 6, UPDV);              \
  } while (0)
      if (upd == 0) FH_ST_LAME(0);
      else if (upd == 1) FH_ST_LAME(1);
      else {
        if constexpr (sizeof(T) == 4) FH_ST_LAME(2);
        else return 0;
      }
#undef FH_ST_LAME
    } else {
      if (upd == 0) FH_ST_NC(COEF_ISO, 0);
      else if (upd == 1) FH_ST_NC(COEF_ISO, 1);
      else {
        if constexpr (sizeof(T) == 4) FH_ST_NC(COEF_ISO, 2);
        else return 0;
      }
    }
#undef FH_ST_NC
#undef FH_LAUNCH_ST
    if (nblocks) *nblocks = (int)nb;
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? 1 : -(int)e;
  }
  // the direct kernel's fused direction update is only race-free when each thread reads its
  // own component
  if (rvec && kind != COEF_ISO && kind != COEF_NONE) return 0;
  if (xc) {
    // fused iteration (k_rows_fwd_fx UPD 2), single precision only; not with a deferred beta
    if constexpr (sizeof(T) == 4) {
      if (!rvec || kind != COEF_ISO || (size_t)nb > partial_cap || c.bpart) return 0;
#define FH_LAUNCH_RF2(NCC)                                                                                \
  do {                                                                                                    \
    auto kf = k_rows_fwd_fx<T, FP, COEF_ISO, NCC, 2>;                                                     \
    allow_lds(kf, lds);                                                                                   \
    hipExtLaunchKernelGGL(kf, grid, block, lds, c.stream, c.t0, c.t1, 0, in, coef, out, nrows, c.n, rpb, tw, halt, rvec, st,   \
                       xc, partial);                                                                      \
  } while (0)
      switch (nc) {
        case 2: FH_LAUNCH_RF2(2); break;
        case 3: FH_LAUNCH_RF2(3); break;
        default: return 0;
      }
#undef FH_LAUNCH_RF2
      if (nblocks) *nblocks = (int)nb;
      const hipError_t e = hipGetLastError();
      return e == hipSuccess ? 1 : -(int)e;
    } else {
      return 0;
    }
  }
  const int key = kind * 8 + nc;
  switch (key) {
    case COEF_NONE * 8 + 2: FH_LAUNCH_RF(COEF_NONE, 2); break;
    case COEF_NONE * 8 + 3: FH_LAUNCH_RF(COEF_NONE, 3); break;
    case COEF_ISO * 8 + 2: FH_LAUNCH_RF(COEF_ISO, 2); break;
    case COEF_ISO * 8 + 3: FH_LAUNCH_RF(COEF_ISO, 3); break;
    default:
      if constexpr (ALL) {
        switch (key) {
          case COEF_NONE * 8 + 1: FH_LAUNCH_RF(COEF_NONE, 1); break;
          case COEF_NONE * 8 + 6: FH_LAUNCH_RF(COEF_NONE, 6); break;
          case COEF_LAME * 8 + 3: FH_LAUNCH_RF(COEF_LAME, 3); break;
          case COEF_LAME * 8 + 6: FH_LAUNCH_RF(COEF_LAME, 6); break;
          case COEF_FULL * 8 + 2: FH_LAUNCH_RF(COEF_FULL, 2); break;
          case COEF_FULL * 8 + 3: FH_LAUNCH_RF(COEF_FULL, 3); break;
          case COEF_FULL * 8 + 6: FH_LAUNCH_RF(COEF_FULL, 6); break;
          default: return 0;
        }
      } else {
        return 0;
      }
  }
#undef FH_LAUNCH_RF
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 1 : -(int)e;
}

template <typename T, class FP>
int rows_fwd_fx_launch(const LaunchCtx& c, int nc, const T* in, const T* coef, int kind, cpx<T>* out,
                       const int* halt, const T* rvec, const CGState* st, T* xc, double* partial,
                       size_t partial_cap, int* nblocks) {
  constexpr int R0 = FxShape<FP, T>::rows_rpb(0), R1 = FxShape<FP, T>::rows_rpb(1);
  constexpr int R2 = FxShape<FP, T>::rows_rpb(2);
  if (c.rows_var == 1) {
    const int r = rows_fwd_fx_launch_r<T, FP, R1, false>(c, nc, in, coef, kind, out, halt, rvec, st, xc, partial,
                                                         partial_cap, nblocks);
    if (r != 0) return r;  // 0: the variant does not fit this case; run the default
  } else if (c.rows_var == 2) {
    const int r = rows_fwd_fx_launch_r<T, FP, R2, false>(c, nc, in, coef, kind, out, halt, rvec, st, xc, partial,
                                                         partial_cap, nblocks);
    if (r != 0) return r;
  }
  return rows_fwd_fx_launch_r<T, FP, R0, true>(c, nc, in, coef, kind, out, halt, rvec, st, xc, partial, partial_cap,
                                               nblocks);
}

template <typename T, class FP, int RPB_, bool ALL>
int rows_inv_fx_launch_r(const LaunchCtx& c, int nc, const cpx<T>* in, T* out, T scale, int epi, const T* w,
                         double* partial, size_t partial_cap, const int* halt, int* nblocks, CGState* st) {
  constexpr int rpb = RPB_;
  if (!ALL && nc != 2 && nc != 3) return 0;
  const int npen = (rpb / 2) * nc;
  const int nthr = npen * FP::TPP;
  if (nthr > 1024) return 0;
  const size_t lds = (size_t)npen * FP::L * sizeof(cpx<T>);
  if (lds + FP::L * sizeof(cpx<T>) > 160 * 1024) return 0;
  if (!offsets_fit(c.n * nc, sizeof(T)) || !offsets_fit(c.nh * nc, sizeof(cpx<T>))) return 0;
  const long long nrows = (long long)c.n0 * c.n1;
  const long long nb = (nrows + rpb - 1) / rpb;
  if (epi && (size_t)nb > partial_cap) return 0;
  if (epi == 1 && !w) return 0;
  if (epi == 3 && (!w || !c.r32 || sizeof(T) != 8)) return 0;
  const dim3 grid((unsigned)nb), block(nthr);
  const cpx<T>* tw = reinterpret_cast<const cpx<T>*>(c.tw[2]);
#define FH_LAUNCH_RI(EPI, NCC)                                                                                \
  do {                                                                                                        \
    const size_t slds = RowStage<FP, T, NCC>::lds_bytes(rpb, false);                                          \
    if (slds + FP::L * sizeof(cpx<T>) > 160 * 1024) return 0;                                                 \
    auto kf = k_rows_inv_st<T, FP, EPI, NCC, RPB_>;                                                           \
    allow_lds(kf, slds);                                                                                      \
    hipExtLaunchKernelGGL(kf, grid, block, slds, c.stream, c.t0, c.t1, 0, in, out, nrows, c.n, rpb, tw, scale, w, partial, halt, \
                       st, order_rev(1), c.defer_beta, c.r32);                                            \
  } while (0)
  const int key = epi * 8 + nc;
  switch (key) {
    case 2: FH_LAUNCH_RI(0, 2); break;
    case 3: FH_LAUNCH_RI(0, 3); break;
    case 8 + 2: FH_LAUNCH_RI(1, 2); break;
    case 8 + 3: FH_LAUNCH_RI(1, 3); break;
    case 16 + 2: FH_LAUNCH_RI(2, 2); break;
    case 16 + 3: FH_LAUNCH_RI(2, 3); break;
    case 24 + 2:
      if constexpr (sizeof(T) == 8) { FH_LAUNCH_RI(3, 2); break; } else { return 0; }
    case 24 + 3:
      if constexpr (sizeof(T) == 8) { FH_LAUNCH_RI(3, 3); break; } else { return 0; }
    default:
      if constexpr (ALL) {
        switch (key) {
          case 1: FH_LAUNCH_RI(0, 1); break;
          case 6: FH_LAUNCH_RI(0, 6); break;
          case 8 + 1: FH_LAUNCH_RI(1, 1); break;
          case 8 + 6: FH_LAUNCH_RI(1, 6); break;
          case 16 + 1: FH_LAUNCH_RI(2, 1); break;
          case 16 + 6: FH_LAUNCH_RI(2, 6); break;
          case 24 + 1:
            if constexpr (sizeof(T) == 8) { FH_LAUNCH_RI(3, 1); break; } else { return 0; }
          case 24 + 6:
            if constexpr (sizeof(T) == 8) { FH_LAUNCH_RI(3, 6); break; } else { return 0; }
          default: return 0;
        }
      } else {
        return 0;
      }
  }
#undef FH_LAUNCH_RI
  if (nblocks) *nblocks = (int)nb;
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 1 : -(int)e;
}

template <typename T, class FP>
int rows_inv_fx_launch(const LaunchCtx& c, int nc, const cpx<T>* in, T* out, T scale, int epi, const T* w,
                       double* partial, size_t partial_cap, const int* halt, int* nblocks, CGState* st) {
  constexpr int R0 = FxShape<FP, T>::rows_rpb(0), R1 = FxShape<FP, T>::rows_rpb(1);
  constexpr int R2 = FxShape<FP, T>::rows_rpb(2);
  if (c.rows_var == 1) {
    const int r = rows_inv_fx_launch_r<T, FP, R1, false>(c, nc, in, out, scale, epi, w, partial, partial_cap, halt,
                                                         nblocks, st);
    if (r != 0) return r;
  } else if (c.rows_var == 2) {
    const int r = rows_inv_fx_launch_r<T, FP, R2, false>(c, nc, in, out, scale, epi, w, partial, partial_cap, halt,
                                                         nblocks, st);
    if (r != 0) return r;
  }
  return rows_inv_fx_launch_r<T, FP, R0, true>(c, nc, in, out, scale, epi, w, partial, partial_cap, halt, nblocks, st);
}

inline void strided_geometry(const LaunchCtx& c, int axis, int nc, long long* outer, long long* ostride,
                             long long* inner) {
  if (axis == 1) {
    *outer = c.n0;
    *inner = (long long)c.n2h * nc;
    *ostride = (long long)c.n1 * (*inner);
  } else {
    *outer = 1;
    *inner = (long long)c.n1 * c.n2h * nc;
    *ostride = 0;
  }
}

template <typename T, class FP, int P>
int cols_fx_launch_p(const LaunchCtx& c, int dir, int axis, int nc, cpx<T>* data, const int* halt) {
  long long outer, ostride, inner;
  strided_geometry(c, axis, nc, &outer, &ostride, &inner);
  const long long total = c.nh * nc;
  if (!offsets_fit(total, sizeof(cpx<T>))) return 0;
  const long long ntiles = (inner + P - 1) / P;
  const long long nb = outer * ntiles;
  if (nb > 0x7fffffffLL) return 0;
  const dim3 grid((unsigned)nb), block(FP::TPP * P);
  const cpx<T>* tw = reinterpret_cast<const cpx<T>*>(c.tw[axis]);
  if (dir < 0)
    hipExtLaunchKernelGGL((k_cols_fx<T, FP, P, -1>), grid, block, 0, c.stream, c.t0, c.t1, 0, data, ostride, inner, (int)ntiles,
                       total, tw, halt, order_rev(0), c.fin);
  else
    hipExtLaunchKernelGGL((k_cols_fx<T, FP, P, +1>), grid, block, 0, c.stream, c.t0, c.t1, 0, data, ostride, inner, (int)ntiles,
                       total, tw, halt, order_rev(2), c.fin);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 1 : -(int)e;
}

template <typename T, class FP>
int cols_fx_launch(const LaunchCtx& c, int dir, int axis, int nc, cpx<T>* data, const int* halt) {
  if (c.cols_var == 1) return cols_fx_launch_p<T, FP, FxShape<FP, T>::cols_p(1)>(c, dir, axis, nc, data, halt);
  return cols_fx_launch_p<T, FP, FxShape<FP, T>::cols_p(0)>(c, dir, axis, nc, data, halt);
}

// frequency points per projection tile of schedule variant V, halved until TPP*F*NC threads fit
// a workgroup
template <typename T, class FP, int NC, int V>
constexpr int proj_points() {
  int f = FxShape<FP, T>::proj_f(V);
  while (f > 1 && FP::TPP * f * NC > 1024) f /= 2;
  return f;
}

template <typename T, class FP, class FPR, int NC, int V>
int cols_g_fx_launch_v(const LaunchCtx& c, int axis, cpx<T>* data, const GParams& g, const int* halt) {
  constexpr int F = proj_points<T, FP, NC, V>();
  if (FP::TPP * F * NC > 1024) return 0;
  long long outer, ostride, inner;
  strided_geometry(c, axis, NC, &outer, &ostride, &inner);
  if (outer != 1) return 0;
  const long long total = c.nh * NC;
  if (!offsets_fit(total, sizeof(cpx<T>))) return 0;
  const long long ntiles = (inner + F * NC - 1) / (F * NC);
  if (ntiles > 0x7fffffffLL) return 0;
  const dim3 grid((unsigned)ntiles), block(FP::TPP * F * NC);
  const cpx<T>* tw = reinterpret_cast<const cpx<T>*>(c.tw[axis]);
  hipExtLaunchKernelGGL((k_cols_g_sw<T, FP, FPR, F, NC>), grid, block, 0, c.stream, c.t0, c.t1, 0, data, inner, (int)ntiles, total,
                     tw, g, c.n2h, halt, c.fin);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 1 : -(int)e;
}

template <typename T, class FP, class FPR, int NC>
int cols_g_fx_launch_nc(const LaunchCtx& c, int axis, cpx<T>* data, const GParams& g, const int* halt) {
  // every schedule variant for 2, 3 and 6 components (the tile's points are halved until the
  // workgroup fits; at small grids fewer points per tile give more workgroups than CUs)
  switch (c.proj_var) {
    case 1: return cols_g_fx_launch_v<T, FP, FPR, NC, 1>(c, axis, data, g, halt);
    case 2: return cols_g_fx_launch_v<T, FP, FPR, NC, 2>(c, axis, data, g, halt);
    case 3: return cols_g_fx_launch_v<T, FP, FPR, NC, 3>(c, axis, data, g, halt);
    default: break;
  }
  return cols_g_fx_launch_v<T, FP, FPR, NC, 0>(c, axis, data, g, halt);
}

template <typename T, class FP, class FPR>
int cols_g_fx_launch(const LaunchCtx& c, int axis, int nc, cpx<T>* data, const GParams& g, const int* halt) {
  switch (nc) {
    case 2: return cols_g_fx_launch_nc<T, FP, FPR, 2>(c, axis, data, g, halt);
    case 3: return cols_g_fx_launch_nc<T, FP, FPR, 3>(c, axis, data, g, halt);
    case 6: return cols_g_fx_launch_nc<T, FP, FPR, 6>(c, axis, data, g, halt);
    default: return 0;
  }
}

template <typename T, class FP>
int rows_rt_fx_launch(const LaunchCtx& c, int nc, cpx<T>* spec, const T* coef, T scale, const int* halt) {
  // fewer rows than the one-way passes: the round trip holds its pencils twice as long, and the
  // smaller LDS tile lets more workgroups share a CU; 6 rows for the long pencils (TPP >= 16:
  // 3 * NC * TPP threads fill the last wavefront better than 4 rows do), half the rows otherwise
  constexpr int rpb = FP::TPP >= 16 ? 6 : (FxShape<FP, T>::RPB >= 4 ? FxShape<FP, T>::RPB / 2 : 2);
  const int npen = (rpb / 2) * nc;
  const int nthr = npen * FP::TPP;
  if (nthr > 1024) return 0;
  if (!offsets_fit(c.nh * nc, sizeof(cpx<T>))) return 0;
  const long long nrows = (long long)c.n0 * c.n1;
  const long long nb = (nrows + rpb - 1) / rpb;
  const dim3 grid((unsigned)nb), block(nthr);
  const cpx<T>* tw = reinterpret_cast<const cpx<T>*>(c.tw[2]);
#define FH_LAUNCH_RT(NCC)                                                                                     \
  do {                                                                                                        \
    const size_t slds = (size_t)RowStage<FP, T, NCC>::pen_stride(rpb) * FP::L * sizeof(cpx<T>); /* pencils */ \
    if (slds + FP::L * sizeof(cpx<T>) > 160 * 1024) return 0;                                                 \
    auto kf = k_rows_rt_st<T, FP, NCC>;                                                                       \
    allow_lds(kf, slds);                                                                                      \
    hipExtLaunchKernelGGL(kf, grid, block, slds, c.stream, c.t0, c.t1, 0, spec, coef, nrows, c.n, rpb, tw, scale, halt); \
  } while (0)
  switch (nc) {
    case 2: FH_LAUNCH_RT(2); break;
    case 3: FH_LAUNCH_RT(3); break;
    default: return 0;
  }
#undef FH_LAUNCH_RT
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 1 : -(int)e;
}

template <typename T>
int fixed_rows_rt(const LaunchCtx& c, int nc, cpx<T>* spec, const T* coef, T scale, const int* halt) {
  switch (c.n2) {
#define FH_CASE(L, FP, FPR) \
  case L: return rows_rt_fx_launch<T, FP>(c, nc, spec, coef, scale, halt);
    FH_FIXED_LENGTHS(FH_CASE)
#undef FH_CASE
    default: return 0;
  }
}

// frequency points per tile of the potential's strided passes: 6 for the long pencils
// (TPP >= 16: TPP * 6 * NC threads fill the last wavefront of a workgroup better than 4 do),
// otherwise the default projection tile; halved until the workgroup fits
template <typename T, class FP, int NC>
constexpr int pot_points() {
  if (FP::TPP < 16) return proj_points<T, FP, NC, 0>();
  int f = 6;
  while (f > 1 && FP::TPP * f * NC > 1024) f /= 2;
  return f;
}

template <typename T, class FP, int NC>
int pot_cols_launch_nc(const LaunchCtx& c, int axis, int dir, cpx<T>* data, cpx<T>* p, const cpx<T>* r,
                       const CGState* st, T sgn, cpx<T>* out, const cpx<T>* w, double* partial, size_t partial_cap,
                       const pot::PotParams& g, const int* halt, int* nblocks) {
  constexpr int F = pot_points<T, FP, NC>();
  if (FP::TPP * F * NC > 1024) return 0;
  long long outer, ostride, inner;
  strided_geometry(c, axis, NC, &outer, &ostride, &inner);
  if (outer != 1) return 0;
  const long long total = c.nh * NC;
  if (!offsets_fit(total, sizeof(cpx<T>))) return 0;
  const long long ntiles = (inner + F * NC - 1) / (F * NC);
  if (ntiles > 0x7fffffffLL) return 0;
  if (dir < 0 && w && (size_t)ntiles > partial_cap) return 0;
  const dim3 grid((unsigned)ntiles), block(FP::TPP * F * NC);
  const cpx<T>* tw = reinterpret_cast<const cpx<T>*>(c.tw[axis]);
  if (dir > 0) {
    if (r)
      hipExtLaunchKernelGGL((pot::k_pot_cols_in<T, FP, F, NC, 1>), grid, block, 0, c.stream, c.t0, c.t1, 0, p, r, st,
                            data, inner, total, tw, g, halt);
    else
      hipExtLaunchKernelGGL((pot::k_pot_cols_in<T, FP, F, NC, 0>), grid, block, 0, c.stream, c.t0, c.t1, 0, p, r, st,
                            data, inner, total, tw, g, halt);
  } else {
    hipExtLaunchKernelGGL((pot::k_pot_cols_out<T, FP, F, NC>), grid, block, 0, c.stream, c.t0, c.t1, 0, data, inner,
                          total, tw, g, sgn, out, w, partial, halt);
  }
  if (nblocks) *nblocks = (int)ntiles;
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? 1 : -(int)e;
}

template <typename T>
int fixed_pot_cols(const LaunchCtx& c, int axis, int dir, int nc, cpx<T>* data, cpx<T>* p, const cpx<T>* r,
                   const CGState* st, T sgn, cpx<T>* out, const cpx<T>* w, double* partial, size_t partial_cap,
                   const pot::PotParams& g, const int* halt, int* nblocks) {
  const int L = axis == 0 ? c.n0 : c.n1;
  switch (L) {
#define FH_CASE(LL, FP, FPR)                                                                                      \
  case LL:                                                                                                        \
    if (nc == 2)                                                                                                  \
      return pot_cols_launch_nc<T, FP, 2>(c, axis, dir, data, p, r, st, sgn, out, w, partial, partial_cap, g, halt, \
                                          nblocks);                                                               \
    if (nc == 3)                                                                                                  \
      return pot_cols_launch_nc<T, FP, 3>(c, axis, dir, data, p, r, st, sgn, out, w, partial, partial_cap, g, halt, \
                                          nblocks);                                                               \
    return 0;
    FH_FIXED_LENGTHS(FH_CASE)
#undef FH_CASE
    default: return 0;
  }
}

template <typename T>
int fixed_rows_fwd(const LaunchCtx& c, int nc, const T* in, const T* coef, int kind, cpx<T>* out,
                   const int* halt, const T* rvec, const CGState* st, T* xc, double* partial,
                   size_t partial_cap, int* nblocks) {
  switch (c.n2) {
#define FH_CASE(L, FP, FPR) \
  case L: return rows_fwd_fx_launch<T, FP>(c, nc, in, coef, kind, out, halt, rvec, st, xc, partial, partial_cap, nblocks);
    FH_FIXED_LENGTHS(FH_CASE)
#undef FH_CASE
    default: return 0;
  }
}

template <typename T>
int fixed_rows_inv(const LaunchCtx& c, int nc, const cpx<T>* in, T* out, T scale, int epi, const T* w,
                   double* partial, size_t partial_cap, const int* halt, int* nblocks, CGState* st) {
  switch (c.n2) {
#define FH_CASE(L, FP, FPR) \
  case L: return rows_inv_fx_launch<T, FP>(c, nc, in, out, scale, epi, w, partial, partial_cap, halt, nblocks, st);
    FH_FIXED_LENGTHS(FH_CASE)
#undef FH_CASE
    default: return 0;
  }
}

template <typename T>
int fixed_cols(const LaunchCtx& c, int dir, int axis, int nc, cpx<T>* data, const int* halt) {
  const int L = axis == 0 ? c.n0 : c.n1;
  switch (L) {
#define FH_CASE(LL, FP, FPR) \
  case LL: return cols_fx_launch<T, FP>(c, dir, axis, nc, data, halt);
    FH_FIXED_LENGTHS(FH_CASE)
#undef FH_CASE
    default: return 0;
  }
}

template <typename T>
int fixed_cols_g(const LaunchCtx& c, int axis, int nc, cpx<T>* data, const GParams& g, const int* halt) {
  const int L = axis == 0 ? c.n0 : c.n1;
  switch (L) {
#define FH_CASE(LL, FP, FPR) \
  case LL: return cols_g_fx_launch<T, FP, FPR>(c, axis, nc, data, g, halt);
    FH_FIXED_LENGTHS(FH_CASE)
#undef FH_CASE
    default: return 0;
  }
}

}  // namespace fh
