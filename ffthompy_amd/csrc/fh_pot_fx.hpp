// Fixed-length strided passes of the potential formulation along the outermost axis of the
// grid (axis 0 in 3-D, axis 1 in 2-D). A tile holds F frequency points of the other axes with
// all NC gradient components, so the scalar potential spectrum is expanded into the gradient
// (k_pot_cols_in) or the divergence is contracted back to a scalar (k_pot_cols_out) inside the
// pass that transforms that axis: the d-component spectrum of the gradient is written once and
// read once less per operator application than with separate elementwise kernels.
#pragma once
#include "fh_fixed.hpp"
#include "fh_pot.hpp"

namespace fh {
namespace pot {

// view frequencies of column point fp (the inner axes) at index i of the transformed axis
__device__ __forceinline__ void col_freq(int i, long long fp, const PotParams& g, int (&k)[3]) {
  if (g.n0 > 1) {
    const int i1 = (int)(fp / g.n2h), i2 = (int)(fp - (long long)i1 * g.n2h);
    k[0] = (2 * i >= g.n0) ? i - g.n0 : i;
    k[1] = (2 * i1 >= g.n1) ? i1 - g.n1 : i1;
    k[2] = ((g.n2 & 1) == 0 && 2 * i2 == g.n2) ? -i2 : i2;
  } else {
    const int i2 = (int)fp;
    k[0] = 0;
    k[1] = (2 * i >= g.n1) ? i - g.n1 : i;
    k[2] = ((g.n2 & 1) == 0 && 2 * i2 == g.n2) ? -i2 : i2;
  }
}

// p (scalar spectrum, reference layout) -> 2 pi i xi_c P p for the tile's columns, then the
// inverse transform along the axis; UPD 1 first sets p = r + beta p (beta from st, written back).
// data: the d-component spectrum (frequency-major, components adjacent), inner = columns per
// index of the axis (points * NC).
template <typename T, class FP, int F, int NC, int UPD>
__global__ void __launch_bounds__(FP::TPP * F * NC)
k_pot_cols_in(cpx<T>* __restrict__ p, const cpx<T>* __restrict__ r, const CGState* __restrict__ st,
              cpx<T>* __restrict__ data, long long inner, long long total, const cpx<T>* __restrict__ twg,
              PotParams g, const int* __restrict__ halt) {
  if (halt && *halt) return;
  constexpr int L = FP::L, P = F * NC, TPP = FP::TPP, NT = TPP * P;
  constexpr unsigned CS = sizeof(cpx<T>);
  static_assert(NT % F == 0, "sweep layout");
  __shared__ cpx<T> s[L * P];
  __shared__ cpx<T> twl[L];
  const int tid = threadIdx.x;
  const int q = tid % P, t = tid / P;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const long long col0 = (long long)tile * P;
  const int ncol = (int)min((long long)P, inner - col0);
  const long long plane = inner / NC;
  for (int e = tid; e < L; e += NT) twl[e] = twg[e];
  const T bta = UPD ? T(st->beta) : T(0);
  const int f = tid % F;
  const long long fp = col0 / NC + f;
  for (int i = tid / F; i < L; i += NT / F) {
    cpx<T> v[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] = mk(T(0), T(0));
    if (f * NC < ncol) {
      const long long idx = (long long)i * plane + fp;
      cpx<T> u = p[idx];
      if constexpr (UPD) {
        const cpx<T> rv = r[idx];
        u = mk(rv.x + bta * u.x, rv.y + bta * u.y);
        p[idx] = u;
      }
      int k[3];
      col_freq(i, fp, g, k);
      T qv[3] = {T(0), T(0), T(0)}, Pw = T(1);
      if (grad_symbol<T>(k, g, qv, Pw)) {
        const cpx<T> w = mk(Pw * u.x, Pw * u.y);
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = mk(-qv[c] * w.y, qv[c] * w.x);
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) s[i * P + f * NC + c] = v[c];
  }
  __syncthreads();
  auto at = [&](int j, int) -> cpx<T>& { return s[j * P + q]; };
  fx_lds_stages<T, FP, +1, 0, FP::NST - 1, 1>(t, at, twl);
  FxStage<T, FP, FP::NST - 1, +1> sl;
  sl.load(t, [&](int j) { return s[j * P + q]; });
  sl.compute(t, twl);
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(data + col0, (unsigned long long)(total - col0) * CS);
  const unsigned rowb = (unsigned)inner * CS, qb = (unsigned)q * CS;
  if (q < ncol) sl.store(t, [=](int j, cpx<T> v) { bstore_c<T>(rs, (unsigned)j * rowb + qb, v); });
}

// Forward transform along the axis of the tile's columns, then out = sgn P sum_c 2 pi i xi_c v_c
// into the scalar spectrum; with w != null per-workgroup partial sums of the weighted <out, w>.
template <typename T, class FP, int F, int NC>
__global__ void __launch_bounds__(FP::TPP * F * NC)
k_pot_cols_out(const cpx<T>* __restrict__ data, long long inner, long long total, const cpx<T>* __restrict__ twg,
               PotParams g, T sgn, cpx<T>* __restrict__ out, const cpx<T>* __restrict__ w,
               double* __restrict__ partial, const int* __restrict__ halt) {
  if (halt && *halt) return;
  constexpr int L = FP::L, P = F * NC, TPP = FP::TPP, NT = TPP * P;
  constexpr unsigned CS = sizeof(cpx<T>);
  static_assert(NT % F == 0, "sweep layout");
  __shared__ cpx<T> s[L * P];
  __shared__ cpx<T> twl[L];
  const int tid = threadIdx.x;
  const int q = tid % P, t = tid / P;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const long long col0 = (long long)tile * P;
  const int ncol = (int)min((long long)P, inner - col0);
  const long long plane = inner / NC;
  const __amdgpu_buffer_rsrc_t rs = make_rsrc(data + col0, (unsigned long long)(total - col0) * CS);
  const unsigned rowb = (unsigned)inner * CS, qb = (unsigned)q * CS;
  FxStage<T, FP, 0, -1> s0;
  s0.load(t, [=](int j) { return bload_c<T>(rs, (unsigned)j * rowb + qb); });
  // the output sweep's points (i, f): thread tid takes f = tid % F and i = tid / F + u * STEP;
  // the dot-product operand w of those points is loaded here, in flight during the transform
  constexpr int STEP = NT / F, IT = (L + STEP - 1) / STEP;
  const int f = tid % F;
  const long long fp = col0 / NC + f;
  cpx<T> wv[IT];
  if (w) {
#pragma unroll
    for (int u = 0; u < IT; ++u) {
      const int i = tid / F + u * STEP;
      if (i < L && f * NC < ncol) wv[u] = w[(long long)i * plane + fp];
    }
  }
  for (int e = tid; e < L; e += NT) twl[e] = twg[e];
  __syncthreads();
  s0.compute(t, twl);
  auto at = [&](int j, int) -> cpx<T>& { return s[j * P + q]; };
  s0.store(t, [&](int j, cpx<T> v) { s[j * P + q] = v; });
  __syncthreads();
  fx_lds_stages<T, FP, -1, 1, FP::NST, 1>(t, at, twl);
  double acc = 0.0;
#pragma unroll
  for (int u = 0; u < IT; ++u) {
    const int i = tid / F + u * STEP;
    if (i < L && f * NC < ncol) {
      int k[3];
      col_freq(i, fp, g, k);
      T qv[3] = {T(0), T(0), T(0)}, Pw = T(1);
      cpx<T> o = mk(T(0), T(0));
      if (grad_symbol<T>(k, g, qv, Pw)) {
        T re = T(0), im = T(0);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const cpx<T> v = s[i * P + f * NC + c];
          re += -qv[c] * v.y;
          im += qv[c] * v.x;
        }
        const T sc = sgn * Pw;
        o = mk(sc * re, sc * im);
      }
      const long long idx = (long long)i * plane + fp;
      out[idx] = o;
      if (w) {
        const double wt = (k[2] == 0 || ((g.n2 & 1) == 0 && 2 * k[2] == -g.n2)) ? 1.0 : 2.0;
        acc += wt * ((double)o.x * (double)wv[u].x + (double)o.y * (double)wv[u].y);
      }
    }
  }
  if (w) {
    const double bs = block_sum(acc);
    if (tid == 0) partial[blockIdx.x] = bs;
  }
}

}  // namespace pot
}  // namespace fh
