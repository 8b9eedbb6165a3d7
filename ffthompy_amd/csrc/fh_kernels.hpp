// Device kernels of the Fourier-Galerkin (GaNi) operator y = F^-1 G(xi) F (A(x) v(x)),
// the conjugate-gradient updates and the energy forms that give the bound matrices.
//
// Layouts (HBM):
//  * real fields: component-major, (ncomp, n0, n1, n2) C-order, exactly the
//    layout of the reference's Tensor.val (ffthompy/tensors/objects.py:92-133);
//    2-D grids are viewed as n0 = 1.
//  * internal half spectrum: (n0, n1, n2/2+1, ncomp) complex, the components of one
//    frequency adjacent, so the projection reads them from one LDS tile.
#pragma once
#include "fh_fft.hpp"

namespace fh {

// A small matrix (up to 6 x 6 doubles, row-major) passed by value as a kernel argument.
struct Mat6 {
  double w[36];
};

constexpr int kMaxComp = 6;

enum CoefKind { COEF_NONE = 0, COEF_ISO = 1, COEF_LAME = 2, COEF_FULL = 3 };

// Conjugate-gradient scalars on the device (see the CG kernels below).
struct CGState {
  double rr, rr_old, pAp, alpha, beta;
  double tol, delta, rmax, inv_nvox;
  int kit, maxiter, done, pause, halt, status, mixed, hist_len, nupdates;
  int bpend;     // fused iteration, small grids: the beta of the last inverse pass is not finalised yet
  int pend;      // fused iteration: alpha * p of the last iteration not yet added to x
  int xfresh;    // fused iteration: the single-precision correction holds no value since the last
                 // (re)start, so its stale contents are read as zero (no clearing pass)
  double* hist;  // device residual log (nullable)
};

// Launch-invariant description of the projection G(xi) of the operator.
struct GParams {
  int dim;      // spatial dimension of the problem (2 or 3)
  int physics;  // 0 scalar (ncomp = dim), 1 small-strain elasticity (Mandel, ncomp = 3 or 6)
  int dual;     // 0: G1 (compatible fields), 1: G2 = I[xi != 0] - G1 (equilibrated fields)
  int ncomp;
  int n0, n1, n2;
  double iY[3];  // 1/Y per axis of the 3-D view
  double scale;  // folded normalisation (1/n for the matvec, 1 for a bare projection)
  int banded;    // 1: G is zero outside |k_a| <= band[a] (the enlarged projection of a Ga solve)
  int band[3];   // per view axis
};

// Map a linear workgroup id so that consecutive tiles run on the same XCD
// (workgroups are dispatched round-robin over the 8 XCDs, each with its own L2).
__device__ __forceinline__ int xcd_remap(int b, int G) {
  const int q = G >> 3, r = G & 7;
  const int x = b & 7, i = b >> 3;
  return x * q + (x < r ? x : r) + i;
}

// Signed frequency of index k on an axis of n points; Nyquist modes of even n
// are flagged (the reference nulls them by building G on the odd grid n-1 and
// zero-padding, ffthompy/projections.py:36-40, 101-105).
__device__ __forceinline__ int signed_freq(int k, int n, bool& nyq) {
  if ((n & 1) == 0 && 2 * k == n) nyq = true;
  return (2 * k > n) ? k - n : k;
}

// Frequency outside the band of an enlarged projection: G built on the solution grid N and
// zero-padded to the Ga grid (hG.enlarge(Nbar), ffthompy/applications.py:28-31).
__device__ __forceinline__ bool out_of_band(int s0, int s1, int k2, const GParams& g) {
  return abs(s0) > g.band[0] || abs(s1) > g.band[1] || k2 > g.band[2];
}

// Apply G(xi) at frequency (k0, k1, k2) to the ncomp values v, times g.scale.
//  scalar : G1 = xi xi^T / |xi|^2 (ffthompy/projections.py:77-83; closed form
//           ffthompy/tensors/projection.py:6-31); G2 = I - G1 (projections.py:84-85).
//  elastic: G1 = Mandel of Gamma_ijkl = sym(delta_ik xi_j xi_l)/|xi|^2 - xi_i xi_j xi_k xi_l/|xi|^4
//           (ffthompy/tensors/projection.py:46-48, equal to G1h+G1s of
//           ffthompy/projections.py:114-267); G2 = I[xi != 0] - G1 (projections.py:239-240).
//  Both vanish at xi = 0 and where `nyq` is set. apply_G_at takes the signed integer frequencies
//  (s0, s1, s2) of the 3-D view; apply_G derives them from the indices of the 'r' layout and sets
//  `nyq` at the Nyquist modes of even axes (the NyqNul projections) and outside a band.
template <typename T>
__device__ __forceinline__ void apply_G_at(cpx<T>* v, int s0, int s1, int s2, bool nyq, const GParams& g) {
  T xi3[3] = {T(s0 * g.iY[0]), T(s1 * g.iY[1]), T(s2 * g.iY[2])};
  const T* xi = (g.dim == 3) ? xi3 : xi3 + 1;
  const int d = g.dim;
  T q = T(0);
  for (int a = 0; a < d; ++a) q += xi[a] * xi[a];
  const int nc = g.ncomp;
  if (nyq || q == T(0)) {
    for (int c = 0; c < nc; ++c) v[c] = mk(T(0), T(0));
    return;
  }
  const T sc = T(g.scale);
  if (g.physics == 0) {
    cpx<T> dot = mk(T(0), T(0));
    for (int c = 0; c < d; ++c) dot = dot + xi[c] * v[c];
    const T f = sc / q;
    for (int c = 0; c < d; ++c) {
      const cpx<T> w = (xi[c] * f) * dot;
      v[c] = g.dual ? (sc * v[c] - w) : w;
    }
    return;
  }
  // Elasticity, Mandel notation: 3-D (11,22,33,23,13,12), 2-D (11,22,12); shears carry sqrt(2).
  const T r2 = T(1.41421356237309504880);
  const T ir2 = T(0.70710678118654752440);
  cpx<T> S[3][3];
  if (d == 3) {
    S[0][0] = v[0]; S[1][1] = v[1]; S[2][2] = v[2];
    S[1][2] = S[2][1] = ir2 * v[3];
    S[0][2] = S[2][0] = ir2 * v[4];
    S[0][1] = S[1][0] = ir2 * v[5];
  } else {
    S[0][0] = v[0]; S[1][1] = v[1];
    S[0][1] = S[1][0] = ir2 * v[2];
  }
  cpx<T> t[3];
  cpx<T> qq = mk(T(0), T(0));
  for (int i = 0; i < d; ++i) {
    t[i] = mk(T(0), T(0));
    for (int j = 0; j < d; ++j) t[i] = t[i] + xi[j] * S[i][j];
    qq = qq + xi[i] * t[i];
  }
  const T iq = T(1) / q;
  const T iq2 = iq * iq;
  auto gam = [&](int i, int j) -> cpx<T> {
    return iq * (xi[i] * t[j] + xi[j] * t[i]) - (xi[i] * xi[j] * iq2) * qq;
  };
  cpx<T> o[kMaxComp];
  if (d == 3) {
    o[0] = gam(0, 0); o[1] = gam(1, 1); o[2] = gam(2, 2);
    o[3] = r2 * gam(1, 2); o[4] = r2 * gam(0, 2); o[5] = r2 * gam(0, 1);
  } else {
    o[0] = gam(0, 0); o[1] = gam(1, 1); o[2] = r2 * gam(0, 1);
  }
  for (int c = 0; c < nc; ++c) v[c] = g.dual ? sc * (v[c] - o[c]) : sc * o[c];
}

template <typename T>
__device__ __forceinline__ void apply_G(cpx<T>* v, int k0, int k1, int k2, const GParams& g) {
  bool nyq = false;
  const int s0 = signed_freq(k0, g.n0, nyq);
  const int s1 = signed_freq(k1, g.n1, nyq);
  if ((g.n2 & 1) == 0 && 2 * k2 == g.n2) nyq = true;
  if (g.banded && out_of_band(s0, s1, k2, g)) nyq = true;
  apply_G_at<T>(v, s0, s1, k2, nyq, g);
}

// y = A(x) v at one voxel (ffthompy/tensors/objects.py:223-232, multype 21).
//  ISO : A = a I          (1 value per voxel)
//  LAME: C = lambda m m^T + 2 mu I in Mandel form, m = (1,..,1,0,..) (2 values per voxel)
//  FULL: any ncomp x ncomp matrix per voxel, entry (i,j) at coef[(i*ncomp+j)*n + vox]
template <typename T, int COEF>
__device__ __forceinline__ void apply_coef(T (&y)[kMaxComp], const T (&v)[kMaxComp], int nc,
                                           const T* __restrict__ coef, long long vox, long long n) {
  if constexpr (COEF == COEF_NONE) {
#pragma unroll
    for (int c = 0; c < kMaxComp; ++c) y[c] = v[c];
  } else if constexpr (COEF == COEF_ISO) {
    const T a = coef[vox];
#pragma unroll
    for (int c = 0; c < kMaxComp; ++c) y[c] = a * v[c];
  } else if constexpr (COEF == COEF_LAME) {
    const T lam = coef[vox], mu2 = T(2) * coef[n + vox];
    const int dn = (nc == 6) ? 3 : 2;
    T tr = T(0);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (c < dn) tr += v[c];
#pragma unroll
    for (int c = 0; c < kMaxComp; ++c) y[c] = mu2 * v[c] + (c < dn ? lam * tr : T(0));
  } else {
#pragma unroll
    for (int i = 0; i < kMaxComp; ++i) {
      T acc = T(0);
      if (i < nc) {
        for (int j = 0; j < nc; ++j) acc += coef[(long long)(i * nc + j) * n + vox] * v[j];
      }
      y[i] = acc;
    }
  }
}

// ---- reductions (wave64 shuffles, fixed tree: bitwise reproducible) -----------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Block sum; result valid in thread 0. Ends with a barrier so it can be reused.
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[16];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (wid == 0) {
    r = (lane < (int)((blockDim.x + 63) >> 6)) ? sh[lane] : 0.0;
    r = wave_sum(r);
  }
  __syncthreads();
  return r;
}

// double-double accumulation for the energy forms
struct dd {
  double hi, lo;
};
__device__ __forceinline__ dd dd_add(dd a, dd b) {
  const double s = a.hi + b.hi;
  const double bb = s - a.hi;
  double e = (a.hi - (s - bb)) + (b.hi - bb);
  e += a.lo + b.lo;
  dd r;
  r.hi = s + e;
  r.lo = e - (r.hi - s);
  return r;
}
__device__ __forceinline__ dd wave_sum_dd(dd v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    dd w;
    w.hi = __shfl_down(v.hi, o, 64);
    w.lo = __shfl_down(v.lo, o, 64);
    v = dd_add(v, w);
  }
  return v;
}
__device__ __forceinline__ dd block_sum_dd(dd v) {
  __shared__ double shh[16], shl[16];
  v = wave_sum_dd(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    shh[wid] = v.hi;
    shl[wid] = v.lo;
  }
  __syncthreads();
  dd r;
  r.hi = 0.0;
  r.lo = 0.0;
  if (wid == 0) {
    if (lane < (int)((blockDim.x + 63) >> 6)) {
      r.hi = shh[lane];
      r.lo = shl[lane];
    }
    r = wave_sum_dd(r);
  }
  __syncthreads();
  return r;
}

// ---- CG scalar updates (run by one workgroup after the reduction kernels) -------------
// alpha = rr / <p, A p>; breakdown (<p,Ap> <= 0 or not finite) stops the solve.
__device__ __forceinline__ void finalize_alpha(CGState* st, double psum) {
  const double pAp = psum * st->inv_nvox;
  st->pAp = pAp;
  if (!(pAp > 0.0) || !isfinite(pAp)) {
    st->status = 1;
    st->done = 1;
    st->halt = 1;
  } else {
    st->alpha = st->rr / pAp;
  }
}

// The step that <r,r> = rsum (unscaled) makes: rr_new, beta = rr_new / rr, the iteration count,
// the norm and the stop decisions. One function, so that every kernel that evaluates it from the
// same sum takes the same decision.
struct BetaStep {
  double rr, beta, nr;
  int kit;
  bool conv, maxed, pause_mixed;
};
__device__ __forceinline__ BetaStep beta_step(const CGState* st, double rsum) {
  BetaStep b;
  b.rr = rsum * st->inv_nvox;
  b.beta = b.rr / st->rr;
  b.kit = st->kit + 1;
  b.nr = sqrt(b.rr);
  b.conv = !(b.nr > st->tol);
  b.maxed = b.kit >= st->maxiter;
  b.pause_mixed = b.conv || b.nr < st->delta * st->rmax || b.maxed;
  return b;
}

// rr <- <r,r>; beta = rr_new / rr_old; iteration count, residual log and stop test.
__device__ __forceinline__ void finalize_beta(CGState* st, double rsum) {
  const BetaStep b = beta_step(st, rsum);
  st->rr_old = st->rr;
  st->rr = b.rr;
  st->beta = b.beta;
  st->kit = b.kit;
  if (st->hist && st->kit < st->hist_len) st->hist[st->kit] = b.nr;
  if (st->mixed) {
    if (b.pause_mixed) st->pause = 1;
  } else {
    st->done = b.conv || b.maxed;
  }
  st->halt = st->done | st->pause;
}

// Sum of np per-workgroup partials in an order that does not depend on the block size: the first
// 32 lanes add lane-strided runs and combine them by an xor butterfly (every lane ends with the same
// bits). Every kernel that calls it has at least 32 threads per block. Returned to every thread of
// the block; call from uniform control flow.
__device__ __forceinline__ double fixed_sum_partials(const double* __restrict__ p, int np) {
  __shared__ double res;
  if (threadIdx.x < 32) {
    // 16 loads in flight per lane before the additions (which keep their order): the partials
    // were just written by workgroups on every XCD, so each load is a trip to memory
    double a = 0.0;
    for (int i0 = 0; i0 < np; i0 += 16 * 32) {
      double v[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int i = i0 + k * 32 + (int)threadIdx.x;
        v[k] = i < np ? p[i] : 0.0;
      }
#pragma unroll
      for (int k = 0; k < 16; ++k) a += v[k];
    }
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) a += __shfl_xor(a, o, 32);
    if (threadIdx.x == 0) res = a;
  }
  __syncthreads();
  const double r = res;
  __syncthreads();
  return r;
}

// Finalisation carried by block 0 of the first strided pass of a fused iteration (in place of the
// one-workgroup alpha kernel, and on a small grid of the beta kernel too): the deferred beta of the
// previous iteration from the inverse pass's partials p2 (p2 set and st->bpend), then alpha of
// this iteration from the forward pass's partials p1. Null st: nothing to do.
struct CgFin {
  CGState* st;
  const double* p1;
  const double* p2;
  int n1, n2;
};
// The caller has already returned on a halted state (read before this kernel wrote anything), so
// every thread of the block enters; the decisions are shared through LDS, never re-read from the
// state this block writes, so the block's control flow stays uniform.
__device__ __forceinline__ void cg_fin_block(const CgFin& f) {
  CGState* st = f.st;
  __shared__ int stop;
  if (f.p2 && st->bpend) {
    const double rs = fixed_sum_partials(f.p2, f.n2);
    if (threadIdx.x == 0) {
      finalize_beta(st, rs);
      st->bpend = 0;
      stop = st->halt;
    }
    __syncthreads();
    if (stop) return;
  }
  const double ps = fixed_sum_partials(f.p1, f.n1);
  if (threadIdx.x == 0) {
    finalize_alpha(st, ps);
    if (st->pend) st->xfresh = 0;  // this iteration's forward pass wrote the correction
    st->pend = st->status == 0 ? 1 : 0;
  }
}

// ---- pass 1: coefficient multiply + real-to-complex FFT along the contiguous axis ----
// A workgroup owns rows_pb consecutive grid rows (all components). Real rows are
// paired into one complex transform (odd lengths have no half-length packing):
// within a row pair, the 2*ncomp real rows (row, component) are paired in order.
// UPD = 1: `in` is the CG direction p, replaced first by r + beta p (beta from *st).
template <typename T, int COEF, int UPD>
__global__ void __launch_bounds__(1024)
k_rows_fwd(T* __restrict__ in, const T* __restrict__ coef, cpx<T>* __restrict__ out,
           long long nrows, int nc, long long n, int rows_pb, AxisPlan ap,
           const cpx<T>* __restrict__ tw, const int* __restrict__ halt, const T* __restrict__ rvec,
           const CGState* __restrict__ st) {
  if (halt && *halt) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* s = reinterpret_cast<cpx<T>*>(smem_raw);
  T* sr = reinterpret_cast<T*>(smem_raw);
  const int L = ap.L, n2h = L / 2 + 1, M = ap.M;  // M: pencil stride in LDS (>= L)
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long row0 = (long long)xcd_remap(blockIdx.x, gridDim.x) * rows_pb;
  const int nr = (int)min((long long)rows_pb, nrows - row0);
  const int npen = (rows_pb >> 1) * nc;
  for (int e = tid; e < rows_pb * L; e += nthr) {
    const int r = e / L, x = e - r * L;
    T y[kMaxComp];
    if (r < nr) {
      const long long vox = (row0 + r) * L + x;
      T v[kMaxComp];
#pragma unroll
      for (int c = 0; c < kMaxComp; ++c) v[c] = (c < nc) ? in[c * n + vox] : T(0);
      if constexpr (UPD == 1) {
        const T bta = T(st->beta);
#pragma unroll
        for (int c = 0; c < kMaxComp; ++c) {
          if (c < nc) {
            v[c] = rvec[c * n + vox] + bta * v[c];
            in[c * n + vox] = v[c];
          }
        }
      }
      apply_coef<T, COEF>(y, v, nc, coef, vox, n);
    } else {
#pragma unroll
      for (int c = 0; c < kMaxComp; ++c) y[c] = T(0);
    }
    const int rp = r >> 1, half = r & 1;
#pragma unroll
    for (int c = 0; c < kMaxComp; ++c) {
      if (c < nc) {
        const int rr = half * nc + c;
        const int pen = rp * nc + (rr >> 1);
        sr[2 * (pen * M + x) + (rr & 1)] = y[c];
      }
    }
  }
  __syncthreads();
  fft_lds<T, -1>(s, M, 1, npen, ap, tw, tid, nthr, true);
  const int nrp = rows_pb >> 1;
  for (int e = tid; e < nrp * n2h; e += nthr) {
    const int rp = e / n2h, k = e - rp * n2h;
    const int kn = (k == 0) ? 0 : L - k;
    for (int m = 0; m < nc; ++m) {
      const cpx<T> Z = s[(rp * nc + m) * M + k];
      const cpx<T> Zc = conj(s[(rp * nc + m) * M + kn]);
      const cpx<T> Ya = T(0.5) * (Z + Zc);
      const cpx<T> dz = Z - Zc;
      const cpx<T> Yb = mk(T(0.5) * dz.y, T(-0.5) * dz.x);
      int rr = 2 * m;
      int row = 2 * rp + rr / nc;
      if (row < nr) out[((row0 + row) * n2h + k) * nc + rr % nc] = Ya;
      rr = 2 * m + 1;
      row = 2 * rp + rr / nc;
      if (row < nr) out[((row0 + row) * n2h + k) * nc + rr % nc] = Yb;
    }
  }
}

// ---- last pass: complex-to-real inverse FFT along the contiguous axis --------------
// Imaginary parts of the k=0 (and even-length Nyquist) coefficients are dropped,
// as numpy.fft.irfftn does. EPI = 1 also accumulates sum(out * w) per workgroup.
template <typename T, int EPI>
__global__ void __launch_bounds__(1024)
k_rows_inv(const cpx<T>* __restrict__ in, T* __restrict__ out, long long nrows, int nc,
           long long n, int rows_pb, AxisPlan ap, const cpx<T>* __restrict__ tw, T scale,
           const T* __restrict__ w, double* __restrict__ partial, const int* __restrict__ halt,
           CGState* __restrict__ st) {
  if (halt && *halt) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* s = reinterpret_cast<cpx<T>*>(smem_raw);
  const int L = ap.L, n2h = L / 2 + 1, M = ap.M;  // M: pencil stride in LDS (>= L)
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long row0 = (long long)xcd_remap(blockIdx.x, gridDim.x) * rows_pb;
  const int nr = (int)min((long long)rows_pb, nrows - row0);
  const int npen = (rows_pb >> 1) * nc;
  const int nrp = rows_pb >> 1;
  for (int e = tid; e < nrp * n2h; e += nthr) {
    const int rp = e / n2h, k = e - rp * n2h;
    const bool self = (k == 0) || (2 * k == L);
    for (int m = 0; m < nc; ++m) {
      int rr = 2 * m;
      int row = 2 * rp + rr / nc;
      cpx<T> Ya = mk(T(0), T(0)), Yb = mk(T(0), T(0));
      if (row < nr) Ya = in[((row0 + row) * n2h + k) * nc + rr % nc];
      rr = 2 * m + 1;
      row = 2 * rp + rr / nc;
      if (row < nr) Yb = in[((row0 + row) * n2h + k) * nc + rr % nc];
      if (self) {
        Ya.y = T(0);
        Yb.y = T(0);
      }
      cpx<T>* z = s + (rp * nc + m) * M;
      z[k] = mk(Ya.x - Yb.y, Ya.y + Yb.x);
      if (!self) z[L - k] = mk(Ya.x + Yb.y, Yb.x - Ya.y);
    }
  }
  __syncthreads();
  fft_lds<T, +1>(s, M, 1, npen, ap, tw, tid, nthr, true);
  const T* sr = reinterpret_cast<const T*>(smem_raw);
  double acc = 0.0;
  for (int e = tid; e < rows_pb * L; e += nthr) {
    const int r = e / L, x = e - r * L;
    if (r < nr) {
      const long long vox = (row0 + r) * L + x;
      const int rp = r >> 1, half = r & 1;
      for (int c = 0; c < nc; ++c) {
        const int rr = half * nc + c;
        const int pen = rp * nc + (rr >> 1);
        const T val = scale * sr[2 * (pen * M + x) + (rr & 1)];
        out[c * n + vox] = val;
        if constexpr (EPI == 1) acc += (double)val * (double)w[c * n + vox];
      }
    }
  }
  if constexpr (EPI == 1) {
    const double b = block_sum(acc);
    if (tid == 0) partial[blockIdx.x] = b;
  }
}

// ---- strided passes: FFT along a non-contiguous axis ---------------------------------
// The half spectrum is viewed as (outer, L, inner) complex. A workgroup loads an
// L x P tile (P consecutive inner columns, coalesced row by row) into LDS.
// MODE 0 forward, 1 inverse, 2 forward + G(xi) + inverse (the last forward axis:
// all components of a frequency are inside the tile because P is a multiple of ncomp).
template <typename T, int MODE>
__global__ void __launch_bounds__(1024)
k_cols(cpx<T>* __restrict__ data, long long ostride, long long inner, int P, int ntiles,
       AxisPlan ap, const cpx<T>* __restrict__ tw, GParams g, int n2h,
       const int* __restrict__ halt) {
  if (halt && *halt) return;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* s = reinterpret_cast<cpx<T>*>(smem_raw);
  const int L = ap.L;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int lin = xcd_remap(blockIdx.x, gridDim.x);
  const int o = lin / ntiles, tile = lin - o * ntiles;
  const long long col0 = (long long)tile * P;
  const int ncol = (int)min((long long)P, inner - col0);
  cpx<T>* base = data + (long long)o * ostride + col0;
  for (int e = tid; e < L * P; e += nthr) {
    const int j = e / P, q = e - j * P;
    s[e] = (q < ncol) ? base[(long long)j * inner + q] : mk(T(0), T(0));
  }
  __syncthreads();
  if constexpr (MODE == 1)
    fft_lds<T, +1>(s, 1, P, P, ap, tw, tid, nthr, false);
  else
    fft_lds<T, -1>(s, 1, P, P, ap, tw, tid, nthr, false);
  if constexpr (MODE == 2) {
    const int nc = g.ncomp, nf = P / nc;
    const long long f0 = col0 / nc;
    for (int e = tid; e < L * nf; e += nthr) {
      const int j = e / nf, f = e - j * nf;
      if (f * nc < ncol) {
        cpx<T> v[kMaxComp];
        cpx<T>* sv = s + j * P + f * nc;
        for (int c = 0; c < nc; ++c) v[c] = sv[c];
        const long long fp = f0 + f;
        int k0, k1, k2;
        if (g.n0 > 1) {
          k0 = j;
          k1 = (int)(fp / n2h);
          k2 = (int)(fp - (long long)k1 * n2h);
        } else {
          k0 = 0;
          k1 = j;
          k2 = (int)fp;
        }
        apply_G<T>(v, k0, k1, k2, g);
        for (int c = 0; c < nc; ++c) sv[c] = v[c];
      }
    }
    __syncthreads();
    fft_lds<T, +1>(s, 1, P, P, ap, tw, tid, nthr, false);
  }
  for (int e = tid; e < L * P; e += nthr) {
    const int j = e / P, q = e - j * P;
    if (q < ncol) base[(long long)j * inner + q] = s[e];
  }
}

// G(xi) on a half spectrum in the reference's component-major 'r' layout
// (ncomp, n0, n1, n2/2+1): the einsum of ffthompy/tensors/objects.py:599-604.
template <typename T>
__global__ void k_apply_G_soa(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out,
                              long long nh, int n2h, GParams g) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nh) return;
  const long long plane = (long long)g.n1 * n2h;
  const int k0 = (int)(i / plane);
  const long long rem = i - (long long)k0 * plane;
  const int k1 = (int)(rem / n2h);
  const int k2 = (int)(rem - (long long)k1 * n2h);
  cpx<T> v[kMaxComp];
  for (int c = 0; c < g.ncomp; ++c) v[c] = in[c * nh + i];
  apply_G<T>(v, k0, k1, k2, g);
  for (int c = 0; c < g.ncomp; ++c) out[c * nh + i] = v[c];
}

// ---- conjugate gradients (ffthompy/general/solver.py:80-139) ------------------------
// Scalars live on the device; every kernel returns at once when the solve has
// finished (done) or waits for a host-side reliable update (pause), so the host
// can enqueue several iterations ahead without reading anything back.


enum CGStatus { CG_OK = 0, CG_BREAKDOWN = 1 };

template <typename T>
__global__ void k_axpby(long long n, double a, const T* __restrict__ x, double b, T* __restrict__ y) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = T(a * (double)x[i] + b * (double)y[i]);
}

template <typename T>
__global__ void k_add_const(long long nvox, int nc, Mat6 E, T* __restrict__ y) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nvox; i += stride)
    for (int c = 0; c < nc; ++c) y[c * nvox + i] += T(E.w[c]);
}

template <typename TS, typename TD>
__global__ void k_convert(long long n, const TS* __restrict__ x, TD* __restrict__ y) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    y[i] = TD(x[i]);
}

template <typename T>
__global__ void k_dot_partial(long long n, const T* __restrict__ x, const T* __restrict__ y,
                              double* __restrict__ partial) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += (double)x[i] * (double)y[i];
  const double b = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

// partial sums of |b - Ax|^2 (nothing written: the true residual of a residual log)
static __global__ void __launch_bounds__(256)
k_resid_norm_partial(long long n, const double* __restrict__ b, const double* __restrict__ Ax,
                     double* __restrict__ partial) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = b[i] - Ax[i];
    acc += v * v;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

static __global__ void k_sum_partials(const double* __restrict__ partial, int np, double scale,
                               double* __restrict__ out) {
  double acc = 0.0;
#pragma unroll 8  // loads issued together; the additions keep their order
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partial[i];
  const double b = block_sum(acc);
  if (threadIdx.x == 0) *out = b * scale;
}

// r = b - Ax (Ax may be null: r = b), in the residual precision TR; partial sums of r.r
template <typename TR>
__global__ void k_residual(long long n, const double* __restrict__ b, const double* __restrict__ Ax,
                           TR* __restrict__ r, double* __restrict__ partial) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = Ax ? b[i] - Ax[i] : b[i];
    r[i] = TR(v);
    acc += v * v;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// (Re)start: rr = <r,r> of the true residual; beta for continuing the direction, or 0 (p = r) with
// newdir, when the true residual differs from the recurrence by more than rounding.
static __global__ void k_cg_restart(CGState* st, const double* __restrict__ partial, int np,
                             double* __restrict__ hist, bool newdir) {
  double acc = 0.0;
#pragma unroll 8  // loads issued together; the additions keep their order
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partial[i];
  const double b = block_sum(acc);
  if (threadIdx.x == 0) {
    const double rr = b * st->inv_nvox;
    st->beta = (st->kit > 0 && st->rr_old > 0.0 && !newdir) ? rr / st->rr_old : 0.0;
    st->rr = rr;
    const double nr = sqrt(rr);
    st->rmax = nr;
    st->pause = 0;
    st->pend = 0;
    st->bpend = 0;
    st->xfresh = 1;
    if (st->kit == 0 && hist && st->hist_len > 0) hist[0] = nr;
    st->done = (!(nr > st->tol)) || (st->kit >= st->maxiter);
    st->halt = st->done;
    st->nupdates += 1;
  }
}

static __global__ void k_cg_alpha(CGState* st, const double* __restrict__ partial, int np) {
  if (st->halt) return;
  double acc = 0.0;
#pragma unroll 8  // loads issued together; the additions keep their order
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partial[i];
  const double b = block_sum(acc);
  if (threadIdx.x == 0) {
    finalize_alpha(st, b);
    if (st->pend) st->xfresh = 0;  // this iteration's forward pass wrote the correction
    st->pend = st->status == 0 ? 1 : 0;
  }
}

static __global__ void k_cg_beta(CGState* st, const double* __restrict__ partial, int np) {
  if (st->halt) return;
  double acc = 0.0;
#pragma unroll 8  // loads issued together; the additions keep their order
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partial[i];
  const double b = block_sum(acc);
  if (threadIdx.x == 0) finalize_beta(st, b);
}

// The deferred beta of a fused iteration on a small grid, at the end of a batch of iterations, so
// that the host reads a finalised state (the same sum and decision as cg_fin_block).
static __global__ void k_cg_flush(CGState* st, const double* __restrict__ partial, int np) {
  if (st->halt || !st->bpend) return;
  const double rs = fixed_sum_partials(partial, np);
  if (threadIdx.x == 0) {
    finalize_beta(st, rs);
    st->bpend = 0;
  }
}

// 16-byte vector of T (4 floats or 2 doubles) for the streaming CG kernels.
template <typename T>
struct alignas(16) V16 {
  static constexpr int W = 16 / sizeof(T);
  T e[W];
};

template <typename T>
__global__ void k_cg_update(long long n, T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
                            const T* __restrict__ Ap, CGState* __restrict__ st,
                            double* __restrict__ partial) {
  if (st->halt) return;
  constexpr int W = V16<T>::W;
  const T a = T(st->alpha);
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nv = n / W;
  double acc = 0.0;
  V16<T>* xv = reinterpret_cast<V16<T>*>(x);
  V16<T>* rv = reinterpret_cast<V16<T>*>(r);
  const V16<T>* pv = reinterpret_cast<const V16<T>*>(p);
  const V16<T>* av = reinterpret_cast<const V16<T>*>(Ap);
  for (long long i = gid; i < nv; i += stride) {
    V16<T> X = xv[i], R = rv[i];
    const V16<T> Pp = pv[i], AP = av[i];
#pragma unroll
    for (int k = 0; k < W; ++k) {
      X.e[k] = X.e[k] + a * Pp.e[k];
      R.e[k] = R.e[k] - a * AP.e[k];
      acc += (double)R.e[k] * (double)R.e[k];
    }
    xv[i] = X;
    rv[i] = R;
  }
  for (long long i = nv * W + gid; i < n; i += stride) {
    x[i] = x[i] + a * p[i];
    const T rn = r[i] - a * Ap[i];
    r[i] = rn;
    acc += (double)rn * (double)rn;
  }
  const double sm = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = sm;
}

template <typename T>
__global__ void k_cg_pupdate(long long n, T* __restrict__ p, const T* __restrict__ r,
                             const CGState* __restrict__ st) {
  if (st->halt) return;
  constexpr int W = V16<T>::W;
  const T bta = T(st->beta);
  const long long stride = (long long)gridDim.x * blockDim.x;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long nv = n / W;
  V16<T>* pv = reinterpret_cast<V16<T>*>(p);
  const V16<T>* rv = reinterpret_cast<const V16<T>*>(r);
  for (long long i = gid; i < nv; i += stride) {
    V16<T> Pp = pv[i];
    const V16<T> R = rv[i];
#pragma unroll
    for (int k = 0; k < W; ++k) Pp.e[k] = R.e[k] + bta * Pp.e[k];
    pv[i] = Pp;
  }
  for (long long i = nv * W + gid; i < n; i += stride) p[i] = r[i] + bta * p[i];
}

// Fused iteration: x64 += x32 + alpha p when a step is pending (st->pend); x32 is read as zero
// when it holds no value (st->xfresh) and left as it is: the restart that follows marks it fresh.
// With set, x64 = x32 + alpha p (the correction alone, for a projection before it is added).
static __global__ void k_fold_pending(long long n, double* __restrict__ x64, const float* __restrict__ x32,
                                      const float* __restrict__ p, const CGState* __restrict__ st, bool set) {
  const float a = st->pend ? (float)st->alpha : 0.0f;
  const bool fresh = st->xfresh != 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    x64[i] = (set ? 0.0 : x64[i]) + (double)((fresh ? 0.0f : x32[i]) + a * p[i]);
}

// x64 += x32 (x64 = x32 with set), x32 = 0 (mixed precision: fold the single-precision correction)
static __global__ void k_fold_correction(long long n, double* __restrict__ x64, float* __restrict__ x32, bool set) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    x64[i] = (set ? 0.0 : x64[i]) + (double)x32[i];
    x32[i] = 0.0f;
  }
}

// ---- energy forms: AH[i][j] = <A e_i, e_j> (ffthompy/postprocess.py:53-70) ------------
// All NC fields are read once; per-workgroup double-double partial sums.
// Up to six device field pointers, passed by value (kernel arguments need no device copy).
struct FieldPtrs {
  const double* p[6];
};

template <int NC, int COEF>
__global__ void __launch_bounds__(256)
k_energy(FieldPtrs F, const double* __restrict__ coef, long long n, double* __restrict__ partial) {
  double acc[NC * NC];
#pragma unroll
  for (int k = 0; k < NC * NC; ++k) acc[k] = 0.0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    double e[NC][NC];
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int c = 0; c < NC; ++c) e[i][c] = F.p[i][c * n + v];
#pragma unroll
    for (int i = 0; i < NC; ++i) {
      double Ae[NC];
      if constexpr (COEF == COEF_ISO) {
        const double a = coef[v];
#pragma unroll
        for (int c = 0; c < NC; ++c) Ae[c] = a * e[i][c];
      } else if constexpr (COEF == COEF_LAME) {
        const int dn = (NC == 6) ? 3 : 2;
        const double lam = coef[v], mu2 = 2.0 * coef[n + v];
        double tr = 0.0;
#pragma unroll
        for (int c = 0; c < dn; ++c) tr += e[i][c];
#pragma unroll
        for (int c = 0; c < NC; ++c) Ae[c] = mu2 * e[i][c] + (c < dn ? lam * tr : 0.0);
      } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          double s = 0.0;
#pragma unroll
          for (int d = 0; d < NC; ++d) s += coef[(long long)(c * NC + d) * n + v] * e[i][d];
          Ae[c] = s;
        }
      }
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) s += Ae[c] * e[j][c];
        acc[i * NC + j] += s;
      }
    }
  }
  for (int k = 0; k < NC * NC; ++k) {
    dd t;
    t.hi = acc[k];
    t.lo = 0.0;
    t = block_sum_dd(t);
    if (threadIdx.x == 0) {
      partial[(2 * (long long)blockIdx.x) * NC * NC + k] = t.hi;
      partial[(2 * (long long)blockIdx.x + 1) * NC * NC + k] = t.lo;
    }
  }
}

static __global__ void k_energy_final(const double* __restrict__ partial, int np, int nn, double inv_n,
                               double* __restrict__ AH) {
  for (int k = 0; k < nn; ++k) {
    dd t;
    t.hi = 0.0;
    t.lo = 0.0;
    for (int b = threadIdx.x; b < np; b += blockDim.x) {
      dd u;
      u.hi = partial[(2 * (long long)b) * nn + k];
      u.lo = partial[(2 * (long long)b + 1) * nn + k];
      t = dd_add(t, u);
    }
    t = block_sum_dd(t);
    if (threadIdx.x == 0) AH[k] = (t.hi + t.lo) * inv_n;
  }
}

// ---- gradient of the energy forms w.r.t. the coefficient (envelope theorem) -----------
// For the minimisers e_i of the primal energy, dAH_ij/da(x) = (A'(x) e_i . e_j)/n with the
// fields held fixed. For a loss L(AH) with W = dL/dAH:
//   ISO : grad[x]            = chain[x] * sum_ij W_ij (e_i(x).e_j(x)) / n
//   LAME: grad[x], grad[n+x] = sum_ij W_ij tr e_i tr e_j / n,  sum_ij W_ij 2 e_i.e_j / n
// chain (nullable) multiplies the ISO gradient per voxel (e.g. a = exp(theta) -> da/dtheta = a).
template <int NC, int COEF>
__global__ void __launch_bounds__(256)
k_energy_grad(FieldPtrs F, long long n, Mat6 W, double inv_n, const double* __restrict__ chain,
              double* __restrict__ grad) {
  double w[NC * NC];
#pragma unroll
  for (int k = 0; k < NC * NC; ++k) w[k] = W.w[k];
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    double e[NC][NC];
#pragma unroll
    for (int i = 0; i < NC; ++i)
#pragma unroll
      for (int c = 0; c < NC; ++c) e[i][c] = F.p[i][c * n + v];
    double gdot = 0.0, gtr = 0.0;
#pragma unroll
    for (int i = 0; i < NC; ++i) {
#pragma unroll
      for (int j = 0; j < NC; ++j) {
        double d = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) d += e[i][c] * e[j][c];
        gdot += w[i * NC + j] * d;
        if constexpr (COEF == COEF_LAME) {
          constexpr int dn = NC == 6 ? 3 : 2;
          double ti = 0.0, tj = 0.0;
#pragma unroll
          for (int c = 0; c < dn; ++c) {
            ti += e[i][c];
            tj += e[j][c];
          }
          gtr += w[i * NC + j] * ti * tj;
        }
      }
    }
    if constexpr (COEF == COEF_ISO) {
      grad[v] = gdot * inv_n * (chain ? chain[v] : 1.0);
    } else {
      grad[v] = gtr * inv_n;
      grad[n + v] = 2.0 * gdot * inv_n;
    }
  }
}

// Adjoint gradient of a loss on the solution fields (isotropic coefficient):
// grad[x] = -chain[x] * sum_i lam_i(x) . e_i(x), with e_i the fields of the unit loads and
// lam_i the adjoint fields (solutions of G(a lam_i) = G dL/de_i). All pointers fp64.
template <int NC>
__global__ void __launch_bounds__(256)
k_field_grad(FieldPtrs F, FieldPtrs Lm, int nf, long long n, const double* __restrict__ chain,
             double* __restrict__ grad) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    double acc = 0.0;
    for (int i = 0; i < nf; ++i) {
#pragma unroll
      for (int c = 0; c < NC; ++c) acc += Lm.p[i][c * n + v] * F.p[i][c * n + v];
    }
    grad[v] = -acc * (chain ? chain[v] : 1.0);
  }
}

}  // namespace fh
