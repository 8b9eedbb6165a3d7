// Potential formulation of the scalar problem (ffthompy/tensorsLowRank/homogenisation.py:43-128):
// the unknown is the Fourier transform of a scalar potential u (rfftn layout, unnormalised),
// the field is e = grad u, and CG runs on
//     x -> P (-div) F A F^-1 grad (P x),     P(xi) = 1 / |2 pi xi|  (1 at xi = 0)
// (homogenisation.py:296-300). grad multiplies by 2 pi i xi_j, div contracts with it
// (ffthompy/tensors/operators.py:227-290). The kernels here are the spectral ends of that
// operator (scalar spectrum <-> d-component spectrum, with P and the band of a Ga solve) and the
// CG vector updates with the inner product of 'r' spectra: plane k2 = 0 (and the Nyquist plane of
// an even last axis) counted once, the other planes twice, divided by n^2
// (ffthompy/tensors/objects.py:618-631), which equals the grid-mean product of the real fields.
#pragma once
#include <hip/hip_runtime.h>

#include "fh_kernels.hpp"

namespace fh {
namespace pot {

struct PotParams {
  int n0, n1, n2, n2h;
  int dim;        // gradient components
  int ax0;        // view axis of component 0 (3 - dim)
  double Y[3];    // cell size per view axis
  double tpiY[3]; // 2 pi / Y per view axis (the symbol is 2 pi k / Y)
  int banded;     // 1: only |k_a| <= band[a] is kept (the coarse grid of a Ga solve)
  int band[3];
};

// Frequencies of half-spectrum entry t: the reference's index sets in ifftshift order, the
// Nyquist index of an even axis standing for -n/2 (ffthompy/trigpol.py:27-40).
__device__ __forceinline__ void freq_of(long long t, const PotParams& g, int (&k)[3]) {
  const int i2 = (int)(t % g.n2h);
  const long long r = t / g.n2h;
  const int i1 = (int)(r % g.n1), i0 = (int)(r / g.n1);
  k[0] = (2 * i0 >= g.n0) ? i0 - g.n0 : i0;
  k[1] = (2 * i1 >= g.n1) ? i1 - g.n1 : i1;
  k[2] = ((g.n2 & 1) == 0 && 2 * i2 == g.n2) ? -i2 : i2;
}

// q_j = 2 pi xi_j = k_j (2 pi / Y_j) on the gradient's axes and P = 1 / |q| (1 at xi = 0), in the
// arithmetic of T: double with an IEEE square root and division; float with the hardware
// reciprocal square root (the single-precision CG of the mixed solve); false outside the band.
template <typename T>
__device__ __forceinline__ bool grad_symbol(const int (&k)[3], const PotParams& g, T (&q)[3], T& P) {
  if (g.banded)
    for (int a = 0; a < 3; ++a)
      if (abs(k[a]) > g.band[a]) return false;
  T k2 = T(0);
  bool zero = true;
  for (int j = 0; j < g.dim; ++j) {
    const int a = g.ax0 + j;
    q[j] = T(k[a]) * T(g.tpiY[a]);
    k2 += q[j] * q[j];
    zero = zero && k[a] == 0;
  }
  if constexpr (sizeof(T) == 4)
    P = zero ? T(1) : __builtin_amdgcn_rsqf(k2);
  else
    P = zero ? T(1) : T(1) / sqrt(k2);
  return true;
}

// Weight of a half-spectrum entry in the inner product of 'r' spectra.
__device__ __forceinline__ double plane_weight(long long t, const PotParams& g) {
  const int i2 = (int)(t % g.n2h);
  return (i2 == 0 || ((g.n2 & 1) == 0 && 2 * i2 == g.n2)) ? 1.0 : 2.0;
}

// spec[t][j] = 2 pi i xi_j (P) u[t]: the gradient of the potential in the operator's spectral
// layout (frequency-major, components adjacent); with_P multiplies by P first.
template <typename T>
__global__ void __launch_bounds__(256)
k_pot_grad(const cpx<T>* __restrict__ u, cpx<T>* __restrict__ spec, PotParams g, int with_P,
           const int* __restrict__ halt) {
  if (halt && *halt) return;
  const long long nh = (long long)g.n0 * g.n1 * g.n2h;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    int k[3];
    freq_of(t, g, k);
    T q[3] = {T(0), T(0), T(0)}, P = T(1);
    const bool in = grad_symbol<T>(k, g, q, P);
    cpx<T> v = u[t];
    if (with_P) v = mk(T(P) * v.x, T(P) * v.y);
    for (int j = 0; j < g.dim; ++j)
      spec[t * g.dim + j] = in ? mk(-T(q[j]) * v.y, T(q[j]) * v.x) : mk(T(0), T(0));
  }
}

// out[t] = sgn P sum_j 2 pi i xi_j spec[t][j] (sgn = -1: the operator, +1: the right-hand side
// div F(A E)); with w != null also per-workgroup partial sums of the weighted <out, w>.
template <typename T>
__global__ void __launch_bounds__(256)
k_pot_div(const cpx<T>* __restrict__ spec, cpx<T>* __restrict__ out, PotParams g, T sgn,
          const cpx<T>* __restrict__ w, double* __restrict__ partial, const int* __restrict__ halt) {
  if (halt && *halt) return;
  const long long nh = (long long)g.n0 * g.n1 * g.n2h;
  double acc = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    int k[3];
    freq_of(t, g, k);
    T q[3] = {T(0), T(0), T(0)}, P = T(1);
    cpx<T> o = mk(T(0), T(0));
    if (grad_symbol<T>(k, g, q, P)) {
      T re = T(0), im = T(0);
      for (int j = 0; j < g.dim; ++j) {
        const cpx<T> s = spec[t * g.dim + j];
        re += -T(q[j]) * s.y;
        im += T(q[j]) * s.x;
      }
      const T f = sgn * T(P);
      o = mk(f * re, f * im);
    }
    out[t] = o;
    if (w) {
      const cpx<T> wv = w[t];
      acc += plane_weight(t, g) * ((double)o.x * (double)wv.x + (double)o.y * (double)wv.y);
    }
  }
  if (w) {
    const double s = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// out[t] = P u[t] inside the band, 0 outside (Fu = P x after the solve).
static __global__ void __launch_bounds__(256)
k_pot_precond(const cpx<double>* __restrict__ u, cpx<double>* __restrict__ out, PotParams g) {
  const long long nh = (long long)g.n0 * g.n1 * g.n2h;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    int k[3];
    freq_of(t, g, k);
    double q[3] = {0.0, 0.0, 0.0}, P = 1.0;
    const cpx<double> v = u[t];
    out[t] = grad_symbol(k, g, q, P) ? mk(P * v.x, P * v.y) : mk(0.0, 0.0);
  }
}

// x += alpha p, r -= alpha Ap, partial sums of the weighted <r, r>.
template <typename T>
__global__ void __launch_bounds__(256)
k_pot_update(cpx<T>* __restrict__ x, cpx<T>* __restrict__ r, const cpx<T>* __restrict__ p,
             const cpx<T>* __restrict__ Ap, PotParams g, CGState* __restrict__ st, double* __restrict__ partial) {
  if (st->halt) return;
  const T a = T(st->alpha);
  const long long nh = (long long)g.n0 * g.n1 * g.n2h;
  double acc = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    const cpx<T> pv = p[t], av = Ap[t];
    cpx<T> xv = x[t], rv = r[t];
    xv = mk(xv.x + a * pv.x, xv.y + a * pv.y);
    rv = mk(rv.x - a * av.x, rv.y - a * av.y);
    x[t] = xv;
    r[t] = rv;
    acc += plane_weight(t, g) * ((double)rv.x * (double)rv.x + (double)rv.y * (double)rv.y);
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// r = b - Ax (Ax null: r = b) in the residual precision T, partial sums of the weighted <r, r>
// of the fp64 values (the true residual of a reliable update).
template <typename T>
__global__ void __launch_bounds__(256)
k_pot_residual(const cpx<double>* __restrict__ b, const cpx<double>* __restrict__ Ax, cpx<T>* __restrict__ r,
               PotParams g, double* __restrict__ partial) {
  const long long nh = (long long)g.n0 * g.n1 * g.n2h;
  double acc = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    cpx<double> v = b[t];
    if (Ax) {
      const cpx<double> a = Ax[t];
      v = mk(v.x - a.x, v.y - a.y);
    }
    r[t] = mk(T(v.x), T(v.y));
    acc += plane_weight(t, g) * (v.x * v.x + v.y * v.y);
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// re[c][i] = E[c]: the constant macroscopic field (d <= 3 components).
static __global__ void __launch_bounds__(256)
k_pot_fill(double* __restrict__ re, long long n, int nc, double e0, double e1, double e2) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n * nc;
       i += (long long)gridDim.x * blockDim.x) {
    const long long c = i / n;
    re[i] = c == 0 ? e0 : (c == 1 ? e1 : e2);
  }
}

}  // namespace pot
}  // namespace fh
