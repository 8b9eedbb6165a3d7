// Staged contiguous-axis passes (fixed lengths, scalar or isotropic coefficient).
//
// A workgroup owns rpb consecutive grid rows of every component. Per component those rows
// are one contiguous run of rpb*L values in HBM, so all real-field traffic of the pass is
// moved in aligned 16-byte chunks (4 floats / 2 doubles per lane, a wave covers 1 KiB),
// and LDS absorbs the reordering into FFT pencils:
//  * forward: chunks -> CG vector updates in registers -> real-row staging in LDS
//    (16-byte LDS writes); the coefficient rows are staged the same way; the first FFT
//    stage reads a(x) v(x) from the staging, pairs two real rows into one complex pencil,
//    and overwrites the staging with pencils (after a barrier); stages run in LDS; the
//    half spectra are untangled into the AoS spectrum.
//  * inverse: AoS spectrum -> Hermitian-completed complex pencils, stages in LDS, the
//    last stage writes its two real rows into the real-row staging; the epilogue (store,
//    dot product or residual update) reads the staging and HBM in 16-byte chunks.
// A staging row run of component c starts at an LDS float index congruent to its HBM
// element index modulo the chunk width, so chunk k of HBM lands on a 16-byte boundary in
// LDS; chunks straddling the run's ends are loaded whole (the extra values are never used)
// and stored element by element.
#pragma once
#include "fh_fixed.hpp"

namespace fh {

template <typename T>
struct alignas(16) Chunk {
  static constexpr int W = 16 / sizeof(T);
  T e[W];
};

template <typename T>
__device__ __forceinline__ Chunk<T> bload_chunk(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(Chunk<T>, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
template <typename T>
__device__ __forceinline__ void bstore_chunk(__amdgpu_buffer_rsrc_t r, unsigned off, const Chunk<T>& v) {
  using V = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, v), r, off, 0, 0);
}

// LDS geometry of the staged passes (host and device).
template <class FP, typename T, int NC>
struct RowStage {
  static constexpr int W = 16 / sizeof(T);
  // real-row staging of one component: rpb*L values plus 3 chunks of alignment slack (a run
  // starts W + (s0 mod W) values into its region and may be read up to W - 1 values past
  // its end), a whole number of chunks
  __host__ __device__ static constexpr int comp_stride(int rpb) { return (rpb * FP::L + 3 * W + W - 1) / W * W; }
  // bytes: max(pencils, staging) + coefficient staging (forward pass only)
  __host__ __device__ static constexpr size_t lds_bytes(int rpb, bool coef) {
    const size_t pen = (size_t)(rpb / 2) * NC * FP::L * sizeof(cpx<T>);
    const size_t stg = (size_t)NC * comp_stride(rpb) * sizeof(T);
    return (pen > stg ? pen : stg) + (coef ? (size_t)comp_stride(rpb) * sizeof(T) : 0);
  }
};

// staging base (in floats/doubles) of a run that starts at HBM element s0
template <typename T>
__device__ __forceinline__ int stage_base(int region, unsigned s0) {
  constexpr int W = 16 / sizeof(T);
  return region + W + (int)(s0 % W);
}

// ---- forward: A(x) v, R2C along the contiguous axis, AoS half spectrum -------------------
// UPD 0: v = in. UPD 1: p = r + beta p (in = p, written back), v = p.
// UPD 2: as UPD 1 and, first, xc += alpha p_old when st->pend; partial sums of <p, a p>.
template <typename T, class FP, int COEF, int NC, int UPD>
__global__ void __launch_bounds__(1024)
k_rows_fwd_st(const T* __restrict__ in, const T* __restrict__ coef, cpx<T>* __restrict__ out,
              long long nrows, long long n, int rpb, const cpx<T>* __restrict__ twg,
              const int* __restrict__ halt, const T* __restrict__ rvec, const CGState* __restrict__ st,
              T* __restrict__ xc, double* __restrict__ partial) {
  static_assert(COEF == COEF_NONE || COEF == COEF_ISO, "staged pass: scalar or isotropic coefficient");
  if (halt && *halt) return;
  constexpr int L = FP::L, n2h = L / 2 + 1, TPP = FP::TPP, W = Chunk<T>::W;
  constexpr unsigned SR = sizeof(T);
  using RS = RowStage<FP, T, NC>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* s = reinterpret_cast<cpx<T>*>(smem_raw);
  T* sr = reinterpret_cast<T*>(smem_raw);
  const int cstr = RS::comp_stride(rpb);
  const size_t main_bytes = RS::lds_bytes(rpb, false);
  T* sa = reinterpret_cast<T*>(smem_raw + main_bytes);
  __shared__ cpx<T> twl[L];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long row0 = (long long)xcd_remap(blockIdx.x, gridDim.x) * rpb;
  const int nr = (int)min((long long)rpb, nrows - row0);
  const unsigned len = (unsigned)nr * L;
  for (int e = tid; e < L; e += nthr) twl[e] = twg[e];
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(in, (unsigned long long)NC * n * SR);
  const unsigned rbase = (unsigned)(row0 * L);
  // coefficient rows -> LDS
  if constexpr (COEF == COEF_ISO) {
    const __amdgpu_buffer_rsrc_t rco = make_rsrc(coef, (unsigned long long)n * SR);
    const int ab = stage_base<T>(0, rbase);
    const unsigned q0 = rbase / W, q1 = (rbase + len + W - 1) / W;
    for (unsigned q = q0 + tid; q < q1; q += nthr) {
      const Chunk<T> v = bload_chunk<T>(rco, q * 16u);
      *reinterpret_cast<Chunk<T>*>(sa + ab + (int)(q * W - rbase)) = v;
    }
  }
  // CG vector updates, the operand v -> real-row staging
  const T bta = UPD ? T(st->beta) : T(0);
  const bool fold = UPD == 2 && st->pend != 0;
  const T alp = UPD == 2 ? T(st->alpha) : T(0);
  const __amdgpu_buffer_rsrc_t rr = make_rsrc(UPD ? rvec : in, (unsigned long long)NC * n * SR);
  const __amdgpu_buffer_rsrc_t rx = make_rsrc(UPD == 2 ? xc : in, (unsigned long long)NC * n * SR);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const unsigned s0 = (unsigned)(c * n) + rbase;
    const int sb = stage_base<T>(c * cstr, s0);
    const unsigned q0 = s0 / W, q1 = (s0 + len + W - 1) / W;
    for (unsigned q = q0 + tid; q < q1; q += nthr) {
      const unsigned ob = q * 16u;
      Chunk<T> P = bload_chunk<T>(rin, ob);
      if constexpr (UPD >= 1) {
        const Chunk<T> R = bload_chunk<T>(rr, ob);
        Chunk<T> X;
        if (fold) X = bload_chunk<T>(rx, ob);
        Chunk<T> Pn;
#pragma unroll
        for (int e = 0; e < W; ++e) {
          if (fold) X.e[e] += alp * P.e[e];
          Pn.e[e] = R.e[e] + bta * P.e[e];
        }
        const bool full = q * W >= s0 && q * W + W <= s0 + len;
        if (full) {
          bstore_chunk<T>(rin, ob, Pn);
          if (fold) bstore_chunk<T>(rx, ob, X);
        } else {
#pragma unroll
          for (int e = 0; e < W; ++e) {
            const unsigned f = q * W + e;
            if (f >= s0 && f < s0 + len) {
              bstore_r<T>(rin, ob + e * SR, Pn.e[e]);
              if (fold) bstore_r<T>(rx, ob + e * SR, X.e[e]);
            }
          }
        }
        P = Pn;
      }
      *reinterpret_cast<Chunk<T>*>(sr + sb + (int)(q * W - s0)) = P;
    }
  }
  __syncthreads();
  // first stage: a(x) v(x) of two real rows per complex pencil, from the staging
  const int t = tid % TPP, pen = tid / TPP;
  const int rp = pen / NC, m = pen - rp * NC;
  const int rra = 2 * m, rrb = 2 * m + 1;
  const int rowa = 2 * rp + rra / NC, ca = rra % NC;
  const int rowb = 2 * rp + rrb / NC, cb = rrb % NC;
  const bool va = rowa < nr, vb = rowb < nr;
  const int pa = stage_base<T>(ca * cstr, (unsigned)(ca * n) + rbase) + rowa * L;
  const int pb = stage_base<T>(cb * cstr, (unsigned)(cb * n) + rbase) + rowb * L;
  const int aa = stage_base<T>(0, rbase) + rowa * L, ab2 = stage_base<T>(0, rbase) + rowb * L;
  double acc = 0.0;
  {
    FxStage<T, FP, 0, -1> s0;
    s0.load(t, [&](int x) {
      T ya = T(0), yb = T(0);
      if (va) {
        const T v = sr[pa + x];
        ya = (COEF == COEF_ISO) ? sa[aa + x] * v : v;
        if constexpr (UPD == 2) acc += (double)ya * (double)v;
      }
      if (vb) {
        const T v = sr[pb + x];
        yb = (COEF == COEF_ISO) ? sa[ab2 + x] * v : v;
        if constexpr (UPD == 2) acc += (double)yb * (double)v;
      }
      return mk(ya, yb);
    });
    __syncthreads();  // the pencils overwrite the staging
    s0.compute(t, twl);
    cpx<T>* sp = s + pen * L;
    s0.store(t, [&](int j, cpx<T> v) { sp[j] = v; });
  }
  __syncthreads();
  {
    cpx<T>* sp = s + pen * L;
    auto at = [&](int j, int) -> cpx<T>& { return sp[j]; };
    fx_lds_stages<T, FP, -1, 1, FP::NST, 1>(t, at, twl);
  }
  // untangle the paired rows and store the AoS half spectrum (contiguous per block): a
  // thread takes one (row, k) and writes its NC adjacent components
  cpx<T>* obp = out + row0 * (n2h * NC);
  for (int e = tid; e < nr * n2h; e += nthr) {
    const int r = e / n2h, k = e - r * n2h;
    const int kc = k ? L - k : 0;
    const cpx<T>* zr = s + (r >> 1) * NC * L;
    cpx<T> Y[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int rr2 = (r & 1) * NC + c;
      const cpx<T>* z = zr + (rr2 >> 1) * L;
      const cpx<T> Z = z[k];
      const cpx<T> Zc = conj(z[kc]);
      if (rr2 & 1) {
        const cpx<T> d = Z - Zc;
        Y[c] = mk(T(0.5) * d.y, T(-0.5) * d.x);
      } else {
        Y[c] = T(0.5) * (Z + Zc);
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) obp[e * NC + c] = Y[c];
  }
  if constexpr (UPD == 2) {
    const double bsum = block_sum(acc);
    if (tid == 0) partial[blockIdx.x] = bsum;
  }
}

// ---- inverse: C2R along the contiguous axis, then the epilogue --------------------------
// EPI 0: out = y. EPI 1: out = y, partial sums of <y, w>. EPI 2: out is the CG residual r,
// r -= alpha y (alpha from st), partial sums of <r, r>.
template <typename T, class FP, int EPI, int NC>
__global__ void __launch_bounds__(1024)
k_rows_inv_st(const cpx<T>* __restrict__ in, T* __restrict__ out, long long nrows, long long n, int rpb,
              const cpx<T>* __restrict__ twg, T scale, const T* __restrict__ w, double* __restrict__ partial,
              const int* __restrict__ halt, CGState* __restrict__ st, int rev) {
  if (halt && *halt) return;
  constexpr int L = FP::L, n2h = L / 2 + 1, TPP = FP::TPP, W = Chunk<T>::W;
  constexpr unsigned SR = sizeof(T);
  using RS = RowStage<FP, T, NC>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* s = reinterpret_cast<cpx<T>*>(smem_raw);
  T* sr = reinterpret_cast<T*>(smem_raw);
  const int cstr = RS::comp_stride(rpb);
  __shared__ cpx<T> twl[L];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int t = tid % TPP, pen = tid / TPP;
  const int blk0 = xcd_remap(blockIdx.x, gridDim.x);
  const long long row0 = (long long)(rev ? (int)gridDim.x - 1 - blk0 : blk0) * rpb;
  const int nr = (int)min((long long)rpb, nrows - row0);
  const unsigned len = (unsigned)nr * L;
  const unsigned rbase = (unsigned)(row0 * L);
  for (int e = tid; e < L; e += nthr) twl[e] = twg[e];
  // Hermitian completion: pencil (rp, mm) = spectrum of real row a + i spectrum of row b
  const int nrp = rpb >> 1;
  const cpx<T>* ib = in + row0 * (long long)n2h * NC;
  for (int e = tid; e < nrp * n2h; e += nthr) {
    const int rp = e / n2h, k = e - rp * n2h;
    const bool self = (k == 0) || (2 * k == L);
#pragma unroll
    for (int mm = 0; mm < NC; ++mm) {
      int rr = 2 * mm;
      int row = 2 * rp + rr / NC;
      cpx<T> Ya = mk(T(0), T(0)), Yb = mk(T(0), T(0));
      if (row < nr) Ya = ib[((long long)row * n2h + k) * NC + rr % NC];
      rr = 2 * mm + 1;
      row = 2 * rp + rr / NC;
      if (row < nr) Yb = ib[((long long)row * n2h + k) * NC + rr % NC];
      if (self) {
        Ya.y = T(0);
        Yb.y = T(0);
      }
      cpx<T>* z = s + (rp * NC + mm) * L;
      z[k] = mk(Ya.x - Yb.y, Ya.y + Yb.x);
      if (!self) z[L - k] = mk(Ya.x + Yb.y, Yb.x - Ya.y);
    }
  }
  __syncthreads();
  cpx<T>* sp = s + pen * L;
  auto at = [&](int j, int) -> cpx<T>& { return sp[j]; };
  fx_lds_stages<T, FP, +1, 0, FP::NST - 1, 1>(t, at, twl);
  // last stage: its two real rows go to the real-row staging (which overlays the pencils)
  {
    const int rp_ = pen / NC, m_ = pen - rp_ * NC;
    const int rowa = 2 * rp_ + (2 * m_) / NC, ca = (2 * m_) % NC;
    const int rowb = 2 * rp_ + (2 * m_ + 1) / NC, cb = (2 * m_ + 1) % NC;
    const int pa = stage_base<T>(ca * cstr, (unsigned)(ca * n) + rbase) + rowa * L;
    const int pb = stage_base<T>(cb * cstr, (unsigned)(cb * n) + rbase) + rowb * L;
    FxStage<T, FP, FP::NST - 1, +1> sl;
    sl.load(t, [&](int j) { return sp[j]; });
    __syncthreads();
    sl.compute(t, twl);
    sl.store(t, [&](int j, cpx<T> v) {
      sr[pa + j] = scale * v.x;
      sr[pb + j] = scale * v.y;
    });
  }
  __syncthreads();
  // epilogue over each component's run of rows, in 16-byte chunks
  const __amdgpu_buffer_rsrc_t rout = make_rsrc(out, (unsigned long long)NC * n * SR);
  const __amdgpu_buffer_rsrc_t rw = make_rsrc(EPI == 1 ? w : out, (unsigned long long)NC * n * SR);
  const T alp = EPI == 2 ? T(st->alpha) : T(0);
  double acc = 0.0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const unsigned s0 = (unsigned)(c * n) + rbase;
    const int sb = stage_base<T>(c * cstr, s0);
    const unsigned q0 = s0 / W, q1 = (s0 + len + W - 1) / W;
    for (unsigned q = q0 + tid; q < q1; q += nthr) {
      const unsigned ob = q * 16u;
      const Chunk<T> Y = *reinterpret_cast<const Chunk<T>*>(sr + sb + (int)(q * W - s0));
      Chunk<T> O = Y;
      const bool full = q * W >= s0 && q * W + W <= s0 + len;
      if constexpr (EPI >= 1) {
        const Chunk<T> V = bload_chunk<T>(rw, ob);
#pragma unroll
        for (int e = 0; e < W; ++e) {
          const unsigned f = q * W + e;
          const bool inr = full || (f >= s0 && f < s0 + len);
          if constexpr (EPI == 1) {
            if (inr) acc += (double)Y.e[e] * (double)V.e[e];
          } else {
            O.e[e] = V.e[e] - alp * Y.e[e];
            if (inr) acc += (double)O.e[e] * (double)O.e[e];
          }
        }
      }
      if (full) {
        bstore_chunk<T>(rout, ob, O);
      } else {
#pragma unroll
        for (int e = 0; e < W; ++e) {
          const unsigned f = q * W + e;
          if (f >= s0 && f < s0 + len) bstore_r<T>(rout, ob + e * SR, O.e[e]);
        }
      }
    }
  }
  if constexpr (EPI >= 1) {
    const double bsum = block_sum(acc);
    if (tid == 0) partial[blockIdx.x] = bsum;
  }
}

}  // namespace fh
