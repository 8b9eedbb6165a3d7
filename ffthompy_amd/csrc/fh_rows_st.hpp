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
//    half spectra are untangled in registers, written back over the pencils in HBM order
//    (the block's AoS half spectrum is one contiguous run) and stored in 16-byte chunks.
//  * FFT threads: lane q = tid % npen owns pencil q (consecutive lanes, consecutive
//    pencils at the same element: conflict-free LDS access, shared twiddles).
//  * inverse: AoS spectrum (every load of a thread issued first) -> Hermitian-completed
//    complex pencils, stages in LDS, the last stage writes its two real rows into the
//    real-row staging; the epilogue (store, dot product, residual update, or the residual
//    of a right-hand side in single precision) reads the staging and HBM in 16-byte chunks.
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
  // pencils are interleaved, element j of pencil q at j*S + q, with S the pencil count: thread
  // (t, q) = (tid / S, tid % S) of a stage reads and writes element i*S + q, so the lanes of a
  // wavefront touch consecutive LDS words whenever their indices i advance with t; the
  // Stockham index patterns of every fixed length then issue without bank conflicts (an odd
  // stride doubled the LDS cycles of the 243-point passes through 2-way conflicts)
  __host__ __device__ static constexpr int pen_stride(int rpb) { return (rpb / 2) * NC; }
  // the forward pass's half spectrum of rpb rows in HBM order, plus one 16-byte chunk of
  // alignment slack (the run starts at its HBM complex index modulo the chunk's complex count)
  __host__ __device__ static constexpr size_t spec_bytes(int rpb) {
    return (size_t)rpb * (FP::L / 2 + 1) * NC * sizeof(cpx<T>) + 16;
  }
  // bytes: max(pencils, staging, spectrum run) + staging of ncoef coefficient runs (forward
  // pass only)
  __host__ __device__ static constexpr size_t lds_bytes(int rpb, int ncoef) {
    const size_t pen = (size_t)pen_stride(rpb) * FP::L * sizeof(cpx<T>);
    const size_t stg = (size_t)NC * comp_stride(rpb) * sizeof(T);
    const size_t spc = spec_bytes(rpb);
    const size_t mx = pen > stg ? pen : stg;
    return (mx > spc ? mx : spc) + (size_t)ncoef * comp_stride(rpb) * sizeof(T);
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
// UPD 2: as UPD 1 and, first, xc += alpha p_old when st->pend (xc = alpha p_old, without reading
// it, while it holds no value since the last restart: st->xfresh); partial sums of <p, a p>.
template <typename T, class FP, int COEF, int NC, int UPD, int RPB = FxShape<FP, T>::RPB>
__global__ void __launch_bounds__(1024)
k_rows_fwd_st(const T* __restrict__ in, const T* __restrict__ coef, cpx<T>* __restrict__ out,
              long long nrows, long long n, int rpb, const cpx<T>* __restrict__ twg,
              const int* __restrict__ halt, const T* __restrict__ rvec, const CGState* __restrict__ st,
              T* __restrict__ xc, double* __restrict__ partial, const double* __restrict__ bpart, int bnp) {
  static_assert(COEF == COEF_NONE || COEF == COEF_ISO || (COEF == COEF_LAME && (NC == 3 || NC == 6)),
                "staged pass: no coefficient, isotropic a, or Lame (lambda, mu) in Mandel form");
  constexpr int KV = COEF == COEF_NONE ? 0 : (COEF == COEF_ISO ? 1 : 2);  // coefficient runs
  constexpr int DN = NC == 6 ? 3 : 2;                                       // normal components
  if (halt && *halt) return;
  constexpr int L = FP::L, n2h = L / 2 + 1, TPP = FP::TPP, W = Chunk<T>::W;
  constexpr unsigned SR = sizeof(T);
  using RS = RowStage<FP, T, NC>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* s = reinterpret_cast<cpx<T>*>(smem_raw);
  T* sr = reinterpret_cast<T*>(smem_raw);
  const int cstr = RS::comp_stride(rpb);
  const size_t main_bytes = RS::lds_bytes(rpb, 0);
  T* sa = reinterpret_cast<T*>(smem_raw + main_bytes);
  __shared__ cpx<T> twl[L];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long long row0 = (long long)xcd_remap(blockIdx.x, gridDim.x) * rpb;
  const int nr = (int)min((long long)rpb, nrows - row0);
  const unsigned len = (unsigned)nr * L;
  for (int e = tid; e < L; e += nthr) twl[e] = twg[e];
  const __amdgpu_buffer_rsrc_t rin = make_rsrc(in, (unsigned long long)NC * n * SR);
  const unsigned rbase = (unsigned)(row0 * L);
  // every HBM load of the block is issued before the first use (CPT chunks per thread per run)
  // RPB (template): rows per workgroup; the launch passes rpb == RPB
  constexpr int NT = (RPB / 2) * NC * TPP;
  constexpr int QMAX = (RPB * L + 2 * W - 2) / W + 1;  // chunks touched by one run of RPB rows
  constexpr int CPT = (QMAX + NT - 1) / NT;
  T bta = UPD ? T(st->beta) : T(0);
  const bool fold = UPD == 2 && st->pend != 0;
  const bool fresh = UPD == 2 && st->xfresh != 0;  // xc read as zero
  const T alp = UPD == 2 ? T(st->alpha) : T(0);
  const __amdgpu_buffer_rsrc_t rr = make_rsrc(UPD ? rvec : in, (unsigned long long)NC * n * SR);
  const __amdgpu_buffer_rsrc_t rx = make_rsrc(UPD == 2 ? xc : in, (unsigned long long)NC * n * SR);
  Chunk<T> P[NC][CPT], R[UPD ? NC : 1][CPT], X[UPD == 2 ? NC : 1][CPT], A[KV ? KV : 1][CPT];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const unsigned s0 = (unsigned)(c * n) + rbase;
    const unsigned q0 = s0 / W, q1 = (s0 + len + W - 1) / W;
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const unsigned q = q0 + tid + i * NT;
      if (q < q1) {
        P[c][i] = bload_chunk<T>(rin, q * 16u);
        if constexpr (UPD >= 1) R[c][i] = bload_chunk<T>(rr, q * 16u);
        if constexpr (UPD == 2) {
          if (fold && !fresh) X[c][i] = bload_chunk<T>(rx, q * 16u);
        }
      }
    }
  }
  if constexpr (KV > 0) {
    // coefficient run k (a; or lambda, mu) of the block's rows -> LDS region k of sa
    const __amdgpu_buffer_rsrc_t rco = make_rsrc(coef, (unsigned long long)KV * n * SR);
#pragma unroll
    for (int k = 0; k < KV; ++k) {
      const unsigned s0 = (unsigned)(k * n) + rbase;
      const unsigned q0 = s0 / W, q1 = (s0 + len + W - 1) / W;
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const unsigned q = q0 + tid + i * NT;
        if (q < q1) A[k][i] = bload_chunk<T>(rco, q * 16u);
      }
    }
#pragma unroll
    for (int k = 0; k < KV; ++k) {
      const unsigned s0 = (unsigned)(k * n) + rbase;
      const unsigned q0 = s0 / W, q1 = (s0 + len + W - 1) / W;
      const int ab = stage_base<T>(k * cstr, s0);
#pragma unroll
      for (int i = 0; i < CPT; ++i) {
        const unsigned q = q0 + tid + i * NT;
        if (q < q1) *reinterpret_cast<Chunk<T>*>(sa + ab + (int)(q * W - s0)) = A[k][i];
      }
    }
  }
  if constexpr (UPD == 2) {
    // small grids: the beta of the previous iteration is not finalised yet; every workgroup sums
    // the inverse pass's partials in one fixed order (its loads overlap the block's own, issued
    // above) and takes the decision the finalising block of the next pass takes: on a stop or a
    // reliable update nothing is written, as when the beta kernel halts the iteration
    if (bpart && st->bpend) {
      const BetaStep b = beta_step(st, fixed_sum_partials(bpart, bnp));
      if (b.pause_mixed) return;
      bta = T(b.beta);
    }
  }
  // CG vector updates, the operand v -> real-row staging
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const unsigned s0 = (unsigned)(c * n) + rbase;
    const int sb = stage_base<T>(c * cstr, s0);
    const unsigned q0 = s0 / W, q1 = (s0 + len + W - 1) / W;
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const unsigned q = q0 + tid + i * NT;
      if (q >= q1) continue;
      const unsigned ob = q * 16u;
      Chunk<T> Pv = P[c][i];
      if constexpr (UPD >= 1) {
        Chunk<T> Pn;
#pragma unroll
        for (int e = 0; e < W; ++e) {
          if constexpr (UPD == 2) {
            if (fold) X[c][i].e[e] = (fresh ? T(0) : X[c][i].e[e]) + alp * Pv.e[e];
          }
          Pn.e[e] = R[c][i].e[e] + bta * Pv.e[e];
        }
        const bool full = q * W >= s0 && q * W + W <= s0 + len;
        if (full) {
          bstore_chunk<T>(rin, ob, Pn);
          if constexpr (UPD == 2) {
            if (fold) bstore_chunk<T>(rx, ob, X[c][i]);
          }
        } else {
#pragma unroll
          for (int e = 0; e < W; ++e) {
            const unsigned f = q * W + e;
            if (f >= s0 && f < s0 + len) {
              bstore_r<T>(rin, ob + e * SR, Pn.e[e]);
              if constexpr (UPD == 2) {
                if (fold) bstore_r<T>(rx, ob + e * SR, X[c][i].e[e]);
              }
            }
          }
        }
        Pv = Pn;
      }
      *reinterpret_cast<Chunk<T>*>(sr + sb + (int)(q * W - s0)) = Pv;
    }
  }
  __syncthreads();
  // first stage: a(x) v(x) of two real rows per complex pencil, from the staging
  const int npen = (rpb / 2) * NC, S = RS::pen_stride(rpb);
  const int pen = tid % npen, t = tid / npen;
  const int rp = pen / NC, m = pen - rp * NC;
  const int rra = 2 * m, rrb = 2 * m + 1;
  const int rowa = 2 * rp + rra / NC, ca = rra % NC;
  const int rowb = 2 * rp + rrb / NC, cb = rrb % NC;
  const bool va = rowa < nr, vb = rowb < nr;
  const int pa = stage_base<T>(ca * cstr, (unsigned)(ca * n) + rbase) + rowa * L;
  const int pb = stage_base<T>(cb * cstr, (unsigned)(cb * n) + rbase) + rowb * L;
  const int aa = stage_base<T>(0, rbase) + rowa * L, ab2 = stage_base<T>(0, rbase) + rowb * L;
  // Lame: mu run and the rows' normal components (for the trace)
  const int ma = stage_base<T>(cstr, (unsigned)n + rbase) + rowa * L;
  const int mb = stage_base<T>(cstr, (unsigned)n + rbase) + rowb * L;
  int tra[DN], trb[DN];
#pragma unroll
  for (int d = 0; d < DN; ++d) {
    tra[d] = stage_base<T>(d * cstr, (unsigned)(d * n) + rbase) + rowa * L;
    trb[d] = stage_base<T>(d * cstr, (unsigned)(d * n) + rbase) + rowb * L;
  }
  double acc = 0.0;
  {
    FxStage<T, FP, 0, -1> s0;
    // rows past nr (last block) read in-bounds staging garbage that the selects discard
    s0.load(t, [&](int x) {
      const T v0 = sr[pa + x], v1 = sr[pb + x];
      T y0 = v0, y1 = v1;
      if constexpr (COEF == COEF_ISO) {
        y0 = sa[aa + x] * v0;
        y1 = sa[ab2 + x] * v1;
      } else if constexpr (COEF == COEF_LAME) {
        // C:e = lambda tr(e) m + 2 mu e (Mandel): lambda run 0, mu run 1 of sa
        T t0 = T(0), t1 = T(0);
#pragma unroll
        for (int d = 0; d < DN; ++d) {
          t0 += sr[tra[d] + x];
          t1 += sr[trb[d] + x];
        }
        y0 = T(2) * sa[ma + x] * v0 + (ca < DN ? sa[aa + x] * t0 : T(0));
        y1 = T(2) * sa[mb + x] * v1 + (cb < DN ? sa[ab2 + x] * t1 : T(0));
      }
      const T ya = va ? y0 : T(0), yb = vb ? y1 : T(0);
      if constexpr (UPD == 2) acc += (double)ya * (double)(va ? v0 : T(0)) + (double)yb * (double)(vb ? v1 : T(0));
      return mk(ya, yb);
    });
    __syncthreads();  // the pencils overwrite the staging
    s0.compute(t, twl);
    s0.store(t, [&](int j, cpx<T> v) { s[j * S + pen] = v; });
  }
  __syncthreads();
  {
    auto at = [&](int j, int) -> cpx<T>& { return s[j * S + pen]; };
    fx_lds_stages<T, FP, -1, 1, FP::NST, 1>(t, at, twl);
  }
  // untangle the paired rows into registers: a thread takes one row pair and one k; pencil m
  // holds the real rows 2m and 2m+1 of the pair's 2*NC (row, component) rows, so two LDS
  // reads give two outputs. Row pair fastest: the lanes' reads s[k*S + rp*NC + m] are NC
  // elements apart.
  const int nrp_v = (nr + 1) >> 1;
  constexpr int UIT = ((RPB / 2) * n2h + NT - 1) / NT;
  cpx<T> Y[UIT][2 * NC];  // Y[u][h*NC + c] is (row 2rp+h, component c)
#pragma unroll
  for (int u = 0; u < UIT; ++u) {
    const int e = tid + u * NT;
    if (e >= nrp_v * n2h) continue;
    const int k = e / nrp_v, rp = e - k * nrp_v;
    const int kc = k ? L - k : 0;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const int q = rp * NC + m;
      const cpx<T> Z = s[k * S + q];
      const cpx<T> Zc = conj(s[kc * S + q]);
      const cpx<T> d = Z - Zc;
      Y[u][2 * m] = T(0.5) * (Z + Zc);
      Y[u][2 * m + 1] = mk(T(0.5) * d.y, T(-0.5) * d.x);
    }
  }
  // the block's half spectrum is one contiguous AoS run of nr * n2h * NC values in HBM: it is
  // written over the pencils in HBM order (starting at its HBM index modulo the chunk's count
  // CW, so HBM chunk q sits on a 16-byte LDS boundary) and stored in 16-byte chunks, instead of
  // each lane storing its NC values at a row pair's stride
  constexpr int CW = 16 / (int)sizeof(cpx<T>);
  const long long c0 = row0 * (long long)(n2h * NC);
  const int so = (int)(c0 % CW);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < UIT; ++u) {
    const int e = tid + u * NT;
    if (e >= nrp_v * n2h) continue;
    const int k = e / nrp_v, rp = e - k * nrp_v;
    cpx<T>* o0 = s + so + ((2 * rp) * n2h + k) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) o0[c] = Y[u][c];
    if (2 * rp + 1 < nr) {
#pragma unroll
      for (int c = 0; c < NC; ++c) o0[n2h * NC + c] = Y[u][NC + c];
    }
  }
  __syncthreads();
  {
    const __amdgpu_buffer_rsrc_t rsp = make_rsrc(out, (unsigned long long)nrows * n2h * NC * sizeof(cpx<T>));
    const long long c1 = c0 + (long long)nr * n2h * NC;  // one past the run's last value
    const unsigned q0 = (unsigned)(c0 / CW), q1 = (unsigned)((c1 + CW - 1) / CW);
    for (unsigned q = q0 + tid; q < q1; q += nthr) {
      const long long g = (long long)q * CW;  // the chunk's first complex index
      const int l = so + (int)(g - c0);       // its LDS index (a multiple of CW)
      if (g >= c0 && g + CW <= c1) {
        bstore_chunk<T>(rsp, q * 16u, *reinterpret_cast<const Chunk<T>*>(s + l));
      } else {
#pragma unroll
        for (int e = 0; e < CW; ++e) {
          if (g + e >= c0 && g + e < c1) bstore_c<T>(rsp, (unsigned)((g + e) * sizeof(cpx<T>)), s[l + e]);
        }
      }
    }
  }
  if constexpr (UPD == 2) {
    const double bsum = block_sum(acc);
    if (tid == 0) partial[blockIdx.x] = bsum;
  }
}

// ---- inverse: C2R along the contiguous axis, then the epilogue --------------------------
// EPI 0: out = y. EPI 1: out = y, partial sums of <y, w>. EPI 2: out is the CG residual r,
// r -= alpha y (alpha from st), partial sums of <r, r>. EPI 3 (double precision): the
// residual of the right-hand side b = w, r32 = float(b - y) (out unused), partial sums of
// <b - y, b - y> in double: the operator's output never goes to HBM.
template <typename T, class FP, int EPI, int NC, int RPB = FxShape<FP, T>::RPB>
__global__ void __launch_bounds__(1024)
k_rows_inv_st(const cpx<T>* __restrict__ in, T* __restrict__ out, long long nrows, long long n, int rpb,
              const cpx<T>* __restrict__ twg, T scale, const T* __restrict__ w, double* __restrict__ partial,
              const int* __restrict__ halt, CGState* __restrict__ st, int rev, int defer_beta,
              float* __restrict__ r32) {
  static_assert(EPI != 3 || sizeof(T) == 8, "residual epilogue: double precision operator");
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
  const int npen = (rpb / 2) * NC, S = RS::pen_stride(rpb);
  const int pen = tid % npen, t = tid / npen;
  const int blk0 = xcd_remap(blockIdx.x, gridDim.x);
  const long long row0 = (long long)(rev ? (int)gridDim.x - 1 - blk0 : blk0) * rpb;
  const int nr = (int)min((long long)rpb, nrows - row0);
  const unsigned len = (unsigned)nr * L;
  const unsigned rbase = (unsigned)(row0 * L);
  for (int e = tid; e < L; e += nthr) twl[e] = twg[e];
  // Hermitian completion: pencil q = rp*NC + m holds real rows 2m (real part) and 2m+1
  // (imaginary part) of row pair rp's 2*NC (row, component) rows. Rows past nr read 0
  // (buffer range); at k = 0 (and k = L/2) both writes hit one element with one value.
  const int nrp = rpb >> 1;
  constexpr unsigned CS = sizeof(cpx<T>);
  const __amdgpu_buffer_rsrc_t rsp =
      make_rsrc(in + row0 * (long long)n2h * NC, (unsigned long long)nr * n2h * NC * CS);
  // every spectrum load of the thread is issued before the first LDS write (the launch passes
  // rpb == RPB and NTH threads)
  constexpr int NTH = (RPB / 2) * NC * TPP;
  constexpr int UIT = ((RPB / 2) * n2h + NTH - 1) / NTH;  // (row pair, k) items per thread
  cpx<T> Y[UIT][2 * NC];
#pragma unroll
  for (int u = 0; u < UIT; ++u) {
    const int e = tid + u * NTH;
    if (e >= nrp * n2h) continue;
    const int k = e / nrp, rp = e - k * nrp;
    const unsigned o0 = (unsigned)(((2 * rp) * n2h + k) * NC) * CS;
    const unsigned o1 = o0 + (unsigned)(n2h * NC) * CS;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      Y[u][c] = bload_c<T>(rsp, o0 + c * CS);
      Y[u][NC + c] = bload_c<T>(rsp, o1 + c * CS);
    }
  }
#pragma unroll
  for (int u = 0; u < UIT; ++u) {
    const int e = tid + u * NTH;
    if (e >= nrp * n2h) continue;
    // row pair fastest: the lanes' writes s[k*S + rp*NC + m] are NC elements apart
    const int k = e / nrp, rp = e - k * nrp;
    const bool self = (k == 0) || (2 * k == L);
    const int kc = k ? L - k : 0;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      cpx<T> Ya = Y[u][2 * m], Yb = Y[u][2 * m + 1];
      if (self) {
        Ya.y = T(0);
        Yb.y = T(0);
      }
      const int q = rp * NC + m;
      s[k * S + q] = mk(Ya.x - Yb.y, Ya.y + Yb.x);
      s[kc * S + q] = mk(Ya.x + Yb.y, Yb.x - Ya.y);
    }
  }
  __syncthreads();
  auto at = [&](int j, int) -> cpx<T>& { return s[j * S + pen]; };
  fx_lds_stages<T, FP, +1, 0, FP::NST - 1, 1>(t, at, twl);
  // last stage: its two real rows go to the real-row staging (which overlays the pencils)
  {
    const int rp_ = pen / NC, m_ = pen - rp_ * NC;
    const int rowa = 2 * rp_ + (2 * m_) / NC, ca = (2 * m_) % NC;
    const int rowb = 2 * rp_ + (2 * m_ + 1) / NC, cb = (2 * m_ + 1) % NC;
    const int pa = stage_base<T>(ca * cstr, (unsigned)(ca * n) + rbase) + rowa * L;
    const int pb = stage_base<T>(cb * cstr, (unsigned)(cb * n) + rbase) + rowb * L;
    FxStage<T, FP, FP::NST - 1, +1> sl;
    sl.load(t, [&](int j) { return s[j * S + pen]; });
    __syncthreads();
    sl.compute(t, twl);
    sl.store(t, [&](int j, cpx<T> v) {
      sr[pa + j] = scale * v.x;
      sr[pb + j] = scale * v.y;
    });
  }
  __syncthreads();
  // epilogue over each component's run of rows, in 16-byte chunks
  // RPB (template): rows per workgroup; the launch passes rpb == RPB
  constexpr int NT = (RPB / 2) * NC * TPP;
  constexpr int QMAX = (RPB * L + 2 * W - 2) / W + 1;  // chunks touched by one run of RPB rows
  constexpr int CPT = (QMAX + NT - 1) / NT;
  const __amdgpu_buffer_rsrc_t rout = make_rsrc(EPI == 3 ? (void*)r32 : (void*)out,
                                                (unsigned long long)NC * n * (EPI == 3 ? 4u : SR));
  const T alp = EPI == 2 ? T(st->alpha) : T(0);
  // the epilogue operand (w, b or r) is read here, not prefetched before the transform: holding
  // it in registers through the transform costs more occupancy than the early load gains
  const __amdgpu_buffer_rsrc_t rv = make_rsrc((EPI == 1 || EPI == 3) ? w : out, (unsigned long long)NC * n * SR);
  double acc = 0.0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const unsigned s0 = (unsigned)(c * n) + rbase;
    const int sb = stage_base<T>(c * cstr, s0);
    const unsigned q0 = s0 / W, q1 = (s0 + len + W - 1) / W;
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
      const unsigned q = q0 + tid + i * NT;
      if (q >= q1) continue;
      const unsigned ob = q * 16u;
      const Chunk<T> Y = *reinterpret_cast<const Chunk<T>*>(sr + sb + (int)(q * W - s0));
      Chunk<T> O = Y;
      Chunk<T> V;
      if constexpr (EPI >= 1) V = bload_chunk<T>(rv, ob);
      const bool full = q * W >= s0 && q * W + W <= s0 + len;
      if constexpr (EPI >= 1) {
#pragma unroll
        for (int e = 0; e < W; ++e) {
          const unsigned f = q * W + e;
          const bool inr = full || (f >= s0 && f < s0 + len);
          if constexpr (EPI == 1) {
            if (inr) acc += (double)Y.e[e] * (double)V.e[e];
          } else if constexpr (EPI == 2) {
            O.e[e] = V.e[e] - alp * Y.e[e];
            if (inr) acc += (double)O.e[e] * (double)O.e[e];
          } else {
            O.e[e] = V.e[e] - Y.e[e];
            if (inr) acc += (double)O.e[e] * (double)O.e[e];
          }
        }
      }
      if constexpr (EPI == 3) {
        // W = 2 doubles -> two floats (8 bytes) at the same element offset of r32
        if (full) {
          bstore_c<float>(rout, q * 8u, mk((float)O.e[0], (float)O.e[1]));
        } else {
#pragma unroll
          for (int e = 0; e < W; ++e) {
            const unsigned f = q * W + e;
            if (f >= s0 && f < s0 + len) bstore_r<float>(rout, f * 4u, (float)O.e[e]);
          }
        }
        continue;
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
  // EPI 2 on small grids: beta is finalised by the next forward pass's readers (st->bpend)
  if (EPI == 2 && defer_beta && blockIdx.x == 0 && tid == 0) st->bpend = 1;
}

// ---- round trip: C2R, a(x) v, R2C along the contiguous axis, in place ---------------------
// The middle of F A F^-1 for a field held as a spectrum (the potential formulation): the AoS
// half spectrum of rpb rows x NC components is completed into pencils as in k_rows_inv_st,
// transformed back (times scale), multiplied by the isotropic coefficient and transformed
// forward as in k_rows_fwd_st; the result overwrites the block's own rows. A pencil holds the
// same two real rows (row, component) in both directions, so the product needs no exchange
// between pencils and the real field never leaves LDS. Each thread loads the a(x) values of its
// last inverse stage's outputs from HBM before that stage, so the tile holds only the pencils.
template <typename T, class FP, int NC>
__global__ void __launch_bounds__(1024)
k_rows_rt_st(cpx<T>* __restrict__ spec, const T* __restrict__ coef, long long nrows, long long n, int rpb,
             const cpx<T>* __restrict__ twg, T scale, const int* __restrict__ halt) {
  if (halt && *halt) return;
  constexpr int L = FP::L, n2h = L / 2 + 1;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cpx<T>* s = reinterpret_cast<cpx<T>*>(smem_raw);
  __shared__ cpx<T> twl[L];
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int npen = (rpb / 2) * NC, S = RowStage<FP, T, NC>::pen_stride(rpb);
  const int pen = tid % npen, t = tid / npen;
  const long long row0 = (long long)xcd_remap(blockIdx.x, gridDim.x) * rpb;
  const int nr = (int)min((long long)rpb, nrows - row0);
  for (int e = tid; e < L; e += nthr) twl[e] = twg[e];
  // Hermitian completion (rows past nr read 0 through the buffer range)
  const int nrp = rpb >> 1;
  constexpr unsigned CS = sizeof(cpx<T>);
  cpx<T>* blk = spec + row0 * (long long)n2h * NC;
  const __amdgpu_buffer_rsrc_t rsp = make_rsrc(blk, (unsigned long long)nr * n2h * NC * CS);
  for (int e = tid; e < nrp * n2h; e += nthr) {
    // row pair fastest: the lanes' writes s[k*S + rp*NC + m] are NC elements apart
    const int k = e / nrp, rp = e - k * nrp;
    const bool self = (k == 0) || (2 * k == L);
    const unsigned o0 = (unsigned)(((2 * rp) * n2h + k) * NC) * CS;
    const unsigned o1 = o0 + (unsigned)(n2h * NC) * CS;
    cpx<T> Y[2 * NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      Y[c] = bload_c<T>(rsp, o0 + c * CS);
      Y[NC + c] = bload_c<T>(rsp, o1 + c * CS);
    }
    const int kc = k ? L - k : 0;
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      cpx<T> Ya = Y[2 * m], Yb = Y[2 * m + 1];
      if (self) {
        Ya.y = T(0);
        Yb.y = T(0);
      }
      const int q = rp * NC + m;
      s[k * S + q] = mk(Ya.x - Yb.y, Ya.y + Yb.x);
      s[kc * S + q] = mk(Ya.x + Yb.y, Yb.x - Ya.y);
    }
  }
  __syncthreads();
  auto at = [&](int j, int) -> cpx<T>& { return s[j * S + pen]; };
  fx_lds_stages<T, FP, +1, 0, FP::NST - 1, 1>(t, at, twl);
  {
    // last inverse stage, then a(x) times the two real rows, back into the pencil. The thread's
    // a(x) values (its outputs' indices of the rows rowa, rowb) are loaded from HBM before the
    // stage (the NC pencils of a row pair read the same values, from cache after the first), so
    // the tile holds only the pencils in LDS
    using SL = FxStage<T, FP, FP::NST - 1, +1>;
    const int rp_ = pen / NC, m_ = pen - rp_ * NC;
    const int rowa = 2 * rp_ + (2 * m_) / NC, rowb = 2 * rp_ + (2 * m_ + 1) / NC;
    const bool va = rowa < nr, vb = rowb < nr;
    const __amdgpu_buffer_rsrc_t rco = make_rsrc(coef + row0 * L, (unsigned long long)nr * L * sizeof(T));
    T ka[SL::M][SL::R], kb[SL::M][SL::R];
#pragma unroll
    for (int m = 0; m < SL::M; ++m) {
      const int i = t + m * SL::TPP;
      if (SL::valid(i)) {
        const int k = i % SL::p, base = (i - k) * SL::R + k;
#pragma unroll
        for (int j = 0; j < SL::R; ++j) {
          const int x = base + j * SL::p;
          ka[m][j] = bload_r<T>(rco, (unsigned)(rowa * L + x) * (unsigned)sizeof(T));
          kb[m][j] = bload_r<T>(rco, (unsigned)(rowb * L + x) * (unsigned)sizeof(T));
        }
      }
    }
    SL sl;
    sl.load(t, [&](int j) { return s[j * S + pen]; });
    __syncthreads();
    sl.compute(t, twl);
    // the outputs in the order of FxStage::store
#pragma unroll
    for (int m = 0; m < SL::M; ++m) {
      const int i = t + m * SL::TPP;
      if (SL::valid(i)) {
        const int k = i % SL::p, base = (i - k) * SL::R + k;
#pragma unroll
        for (int j = 0; j < SL::R; ++j) {
          const cpx<T> v = sl.v[m][j];
          s[(base + j * SL::p) * S + pen] =
              mk(va ? scale * v.x * ka[m][j] : T(0), vb ? scale * v.y * kb[m][j] : T(0));
        }
      }
    }
  }
  __syncthreads();
  {
    FxStage<T, FP, 0, -1> s0;
    s0.load(t, [&](int x) { return s[x * S + pen]; });
    __syncthreads();
    s0.compute(t, twl);
    s0.store(t, [&](int j, cpx<T> v) { s[j * S + pen] = v; });
  }
  __syncthreads();
  fx_lds_stages<T, FP, -1, 1, FP::NST, 1>(t, at, twl);
  // untangle the paired rows into the AoS half spectrum (as k_rows_fwd_st)
  const int nrp_v = (nr + 1) >> 1;
  for (int e = tid; e < nrp_v * n2h; e += nthr) {
    // row pair fastest: the lanes' reads s[k*S + rp*NC + m] are NC elements apart
    const int k = e / nrp_v, rp = e - k * nrp_v;
    const int kc = k ? L - k : 0;
    cpx<T> Y[2 * NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
      const int q = rp * NC + m;
      const cpx<T> Z = s[k * S + q];
      const cpx<T> Zc = conj(s[kc * S + q]);
      const cpx<T> d = Z - Zc;
      Y[2 * m] = T(0.5) * (Z + Zc);
      Y[2 * m + 1] = mk(T(0.5) * d.y, T(-0.5) * d.x);
    }
    cpx<T>* o0 = blk + ((long long)(2 * rp) * n2h + k) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) o0[c] = Y[c];
    if (2 * rp + 1 < nr) {
      cpx<T>* o1 = o0 + (long long)n2h * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) o1[c] = Y[NC + c];
    }
  }
}

}  // namespace fh
