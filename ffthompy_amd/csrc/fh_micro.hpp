// Microstructure and coefficient generation on the device: nodal values of a material defined by
// inclusions (Material.evaluate / get_topologies, ffthompy/materials.py:171-301, through
// Material.get_A_GaNi, materials.py:115-123) and the spectral filter of a periodic Gaussian
// random field.
//
// The nodal evaluation reproduces the reference's floating-point expressions operation by
// operation (no contraction into fused multiply-adds), so membership tests at the nodes decide
// exactly as the reference's NumPy code does.
#pragma once
#include <hip/hip_runtime.h>

#include "fh_ga.hpp"

namespace fh {
namespace micro {

struct NodalGrid {
  int dim;
  int N[3];      // problem axes
  double Y[3];
  long long n;
  int ncoef;     // coefficient components per voxel in the output layout
  int layout;    // 1 iso (a I: one value), 2 Lame (lambda, mu), 3 full (ncoef = k^2)
  int k;         // components of a field (matrix size)
  int invert;    // 1: pointwise inverse (dual problem)
};

// Invert a k x k matrix in place (Gauss-Jordan, partial pivoting); returns false if singular.
__device__ inline bool invert_small(double* a, int k) {
  double inv[36];
  for (int i = 0; i < k * k; ++i) inv[i] = 0.0;
  for (int i = 0; i < k; ++i) inv[i * k + i] = 1.0;
  for (int c = 0; c < k; ++c) {
    int piv = c;
    double best = fabs(a[c * k + c]);
    for (int r = c + 1; r < k; ++r)
      if (fabs(a[r * k + c]) > best) {
        best = fabs(a[r * k + c]);
        piv = r;
      }
    if (!(best > 0.0)) return false;
    if (piv != c)
      for (int j = 0; j < k; ++j) {
        double t = a[c * k + j];
        a[c * k + j] = a[piv * k + j];
        a[piv * k + j] = t;
        t = inv[c * k + j];
        inv[c * k + j] = inv[piv * k + j];
        inv[piv * k + j] = t;
      }
    const double d = 1.0 / a[c * k + c];
    for (int j = 0; j < k; ++j) {
      a[c * k + j] *= d;
      inv[c * k + j] *= d;
    }
    for (int r = 0; r < k; ++r) {
      if (r == c) continue;
      const double f = a[r * k + c];
      if (f == 0.0) continue;
      for (int j = 0; j < k; ++j) {
        a[r * k + j] -= f * a[c * k + j];
        inv[r * k + j] -= f * inv[c * k + j];
      }
    }
  }
  for (int i = 0; i < k * k; ++i) a[i] = inv[i];
  return true;
}

// One output voxel per thread, origin-0 order (voxel i of an axis of N points is the node
// x = Y k / N with k = i for 2i < N, else i - N: the centred node set after ifftshift).
// err[0] = 1: 'otherwise' went negative (overlapping inclusions), err[1] = 1: singular inverse.
__global__ void __launch_bounds__(256)
k_nodal_coef(const ga::Incl* __restrict__ incl, int nincl, NodalGrid g, double* __restrict__ out,
             int* __restrict__ err) {
#pragma clang fp contract(off)
  for (long long vox = (long long)blockIdx.x * blockDim.x + threadIdx.x; vox < g.n;
       vox += (long long)gridDim.x * blockDim.x) {
    double x[3] = {0.0, 0.0, 0.0};
    long long rem = vox;
    for (int a = g.dim - 1; a >= 0; --a) {
      const int i = (int)(rem % g.N[a]);
      rem /= g.N[a];
      const int kk = (2 * i < g.N[a]) ? i : i - g.N[a];
      x[a] = g.Y[a] * (double)kk / (double)g.N[a];
    }
    double A[36];
    for (int c = 0; c < g.ncoef; ++c) A[c] = 0.0;
    double tsum = 0.0;  // sum of the topologies so far ('otherwise' = 1 - sum)
    for (int ii = 0; ii < nincl; ++ii) {
      const ga::Incl& in = incl[ii];
      double t = 0.0;
      if (in.kind == ga::INCL_ALL) {
        t = 1.0;
      } else if (in.kind == ga::INCL_OTHERWISE) {
        t = 1.0 - tsum;
        if (t < 0.0) err[0] = 1;
      } else {
        // periodic copies over the neighbouring cells, shifts in lexicographic order
        const int nsh = g.dim == 3 ? 27 : 9;
        for (int sidx = 0; sidx < nsh; ++sidx) {
          int s[3];
          int r = sidx;
          for (int a = g.dim - 1; a >= 0; --a) {
            s[a] = r % 3 - 1;
            r /= 3;
          }
          if (in.kind == ga::INCL_CUBE) {
            bool m = true;
            for (int a = 0; a < g.dim; ++a) {
              const double z = (x[a] - in.pos[a]) + g.Y[a] * (double)s[a];
              m = m && (z > -in.p[a] / 2) && (z <= in.p[a] / 2);
            }
            t += m ? 1.0 : 0.0;
          } else if (in.kind == ga::INCL_BALL) {
            double r2 = 0.0;
            for (int a = 0; a < g.dim; ++a) {
              const double z = (x[a] - in.pos[a]) - g.Y[a] * (double)s[a];
              r2 = r2 + z * z;
            }
            t += (sqrt(r2) < in.p[0] / 2) ? 1.0 : 0.0;
          } else {  // pyramid: the reference shifts by 2Y (materials.py:281)
            double v = 1.0;
            for (int a = 0; a < g.dim; ++a) {
              const double z = (x[a] - in.pos[a]) + (2.0 * g.Y[a]) * (double)s[a];
              v = v * fmax(1.0 - fabs(z) / (in.p[a] / 2.0), 0.0);
            }
            t = t + v;
          }
        }
      }
      tsum = tsum + t;
      for (int c = 0; c < g.ncoef; ++c) A[c] = A[c] + in.v[c] * t;
    }
    if (g.invert) {
      if (g.layout == 1) {
        A[0] = 1.0 / A[0];
      } else if (g.layout == 2) {
        const double lam = A[0], mu = A[1];
        const double dn = g.k == 6 ? 3.0 : 2.0;
        A[0] = -lam / (2.0 * mu * (2.0 * mu + dn * lam));
        A[1] = 1.0 / (4.0 * mu);
      } else if (!invert_small(A, g.k)) {
        err[1] = 1;
      }
    }
    for (int c = 0; c < g.ncoef; ++c) out[(long long)c * g.n + vox] = A[c];
  }
}

// Spectral filter of a Gaussian random field: spec(k) *= s exp(-pi^2 l^2 |xi|^2), xi_a = k_a / Y_a,
// on the half spectrum (n0, n1, n2/2+1) of the 3-D view.
__global__ void __launch_bounds__(256)
k_grf_filter(cpx<double>* __restrict__ spec, int n0, int n1, int n2, double iy0, double iy1, double iy2,
             double corr_len, double s) {
  const int n2h = n2 / 2 + 1;
  const long long nh = (long long)n0 * n1 * n2h;
  const double pi = 3.141592653589793;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    const int i2 = (int)(t % n2h);
    const long long r = t / n2h;
    const int i1 = (int)(r % n1), i0 = (int)(r / n1);
    const double k0 = (double)ga::sfreq(i0, n0) * iy0, k1 = (double)ga::sfreq(i1, n1) * iy1, k2 = (double)i2 * iy2;
    const double xi2 = k0 * k0 + k1 * k1 + k2 * k2;
    const double f = s * exp(-(pi * pi) * (corr_len * corr_len) * xi2);
    const cpx<double> v = spec[t];
    spec[t] = mk(f * v.x, f * v.y);
  }
}

__global__ void __launch_bounds__(256) k_count_nonfinite(long long n, const double* __restrict__ x, int* __restrict__ bad) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    if (!isfinite(x[i])) *bad = 1;
}

__global__ void __launch_bounds__(256) k_count_nonzero(long long n, const double* __restrict__ x, int* __restrict__ flag) {
  bool nz = false;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    nz = nz || (x[i] != 0.0);
  if (nz) *flag = 1;
}

__global__ void __launch_bounds__(256) k_exp_inplace(long long n, double* __restrict__ x) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    x[i] = exp(x[i]);
}

}  // namespace micro
}  // namespace fh
