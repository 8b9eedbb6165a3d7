// Stationary iterative solvers on the device: Richardson (ffthompy/general/solver.py:63-77) and the
// two-term Chebyshev iteration (solver.py:206-285). One iteration is a matvec plus a few
// elementwise kernels; the stop test runs in a one-workgroup kernel that raises `halt`, which
// every later kernel of the iteration checks, so the host can enqueue iterations ahead.
#pragma once
#include <hip/hip_runtime.h>

#include "fh_kernels.hpp"

namespace fh {
namespace it {

struct IterState {
  double tol;        // stop when the residual norm is <= tol (Richardson) / < tol (Chebyshev, relative)
  double inv_nvox;   // grid-mean inner product: sum / n
  double r0;         // Chebyshev: norm of the initial residual (the residual norms are relative to it)
  double norm;       // last residual norm (relative for Chebyshev)
  int kit, maxiter;
  int done;          // 1: stopped
  int halt;          // read by the queued kernels (1 once done)
  int relative;      // 1: norm = ||r|| / r0
  int hist_len;
  double* hist;      // hist[k] = norm after iteration k (k >= 1)
};

// r = b - Ax and per-workgroup partial sums of r.r
__global__ void __launch_bounds__(256)
k_it_residual(long long n, const double* __restrict__ b, const double* __restrict__ Ax, double* __restrict__ r,
              double* __restrict__ partial, const int* __restrict__ halt) {
  if (*halt) return;
  const long long stride = (long long)gridDim.x * blockDim.x;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double v = b[i] - Ax[i];
    r[i] = v;
    acc += v * v;
  }
  const double s = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Richardson step x += omega r
__global__ void __launch_bounds__(256)
k_it_richardson(long long n, double* __restrict__ x, const double* __restrict__ r, double omega,
                const int* __restrict__ halt) {
  if (*halt) return;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] += omega * r[i];
}

// Chebyshev step v = r - p v, x += w v
__global__ void __launch_bounds__(256)
k_it_chebyshev(long long n, double* __restrict__ x, double* __restrict__ v, const double* __restrict__ r,
               double p, double w, const int* __restrict__ halt) {
  if (*halt) return;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double vi = r[i] - p * v[i];
    v[i] = vi;
    x[i] += w * vi;
  }
}

// Close an iteration: norm of the residual from the partial sums, iteration count, history and
// the reference's loop condition (norm > tol and kit < maxiter continue).
__global__ void __launch_bounds__(256)
k_it_finish(IterState* st, const double* __restrict__ partial, int np) {
  if (st->halt) return;
  double acc = 0.0;
#pragma unroll 8  // loads issued together; the additions keep their order
  for (int i = threadIdx.x; i < np; i += blockDim.x) acc += partial[i];
  const double s = block_sum(acc);
  if (threadIdx.x == 0) {
    double nr = sqrt(s * st->inv_nvox);
    if (st->relative) nr = nr / st->r0;
    st->kit += 1;
    st->norm = nr;
    if (st->hist && st->kit < st->hist_len) st->hist[st->kit] = nr;
    if (!(nr > st->tol) || st->kit >= st->maxiter) {
      st->done = 1;
      st->halt = 1;
    }
  }
}

}  // namespace it
}  // namespace fh
