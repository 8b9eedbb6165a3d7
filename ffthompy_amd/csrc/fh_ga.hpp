// Exactly integrated ("Ga") coefficients and trigonometric interpolation onto a quadrature grid.
//
// The Ga energy <A_Ga P(e_i), P(e_j)> of ffthompy/postprocess.py:20-41, 53-70 integrates, on the
// grid Nbar = 2N-1, the product of
//   * A_Ga, a trigonometric polynomial with frequencies |k_a| <= N_a - 1 (its Fourier
//     coefficients are the exact integrals of the inclusion shape functions,
//     ffthompy/materials.py:54-113, 125-169), and
//   * P(e_i), P(e_j), the trigonometric interpolants of the GaNi solutions (frequencies
//     |k_a| <= N_a / 2; Tensor.project/enlarge, ffthompy/tensors/objects.py:428-511).
// All three live on the quadrature grid M as spectra, built here from closed forms or from a
// coarse spectrum, and reach real space through the plan's inverse FFT passes. For odd N the
// product has frequencies |k_a| <= 2N_a - 2 < M_a, so the grid mean is the exact integral on
// any M_a >= 2N_a - 1; for even N it equals the reference's value only on M_a = 2N_a - 1.
//
// Kernels index the half spectrum (m0, m1, m2/2+1) of the 3-D view of the quadrature grid
// (2-D problems: m0 = 1), one thread per frequency.
#pragma once
#include <hip/hip_runtime.h>

#include "fh_kernels.hpp"

namespace fh {
namespace ga {

enum { INCL_CUBE = 1, INCL_BALL = 2, INCL_PYRAMID = 3, INCL_ALL = 4, INCL_OTHERWISE = 5 };
constexpr int kMaxCoef = 36;

// One inclusion of a material (ffthompy/materials.py:15-52), in the axes of the 3-D view
// (2-D problems: axis 0 is the unit dummy axis, size 1, position 0).
struct Incl {
  int kind;
  double p[3];        // box / pyramid side lengths per axis; ball: p[0] = diameter
  double pos[3];      // position (inclusion centre), reduced modulo Y
  double v[kMaxCoef]; // coefficient components of this phase (already inverted for the dual)
};

struct SpecGrid {
  int m0, m1, m2, m2h;  // quadrature grid (3-D view) and its half length on the last axis
  int band[3];          // |k_a| <= band[a]: the Nbar = 2N-1 index set (band[a] = N_a - 1)
  int nbar[3];          // Nbar per view axis (1 on the dummy axis)
  double Y[3];          // cell size per view axis (1 on the dummy axis)
  double vol;           // |Y| over the problem's axes
  double scale;         // factor applied to every coefficient
};

__device__ __forceinline__ int sfreq(int i, int m) { return (2 * i >= m) ? i - m : i; }

// Frequency of the ifftshift-ordered index set that the reference pairs with the centred
// frequency k when it multiplies the inclusion phase factor (built on the 'r' index set,
// ffthompy/materials.py:317-330, trigpol.py:11-24) with the weights (built on the centred
// set, materials.py:333-425): the same array position, read in the other order.
__device__ __forceinline__ int shifted_order_freq(int k, int nbar) {
  const int h = (nbar - 1) / 2;  // nbar is odd
  return (k <= 0) ? k + h : k - (nbar + 1) / 2;
}

// sinc(x) = sin(pi x) / (pi x), as numpy.sinc.
__device__ __forceinline__ double sinc_pi(double x) {
  if (x == 0.0) return 1.0;
  const double px = 3.141592653589793 * x;
  return sin(px) / px;
}

// Fourier coefficient of one inclusion's shape function at frequency k (view axes). The
// reference takes the real part of sum_k X(k) e^{2 pi i k.x/Y} with X(k) = W(k) S(k)
// (materials.py:125-169), so the coefficient is the Hermitian part (X(k) + conj X(-k)) / 2;
// it equals X(k) when the position is 0 (S = 1).
__device__ cpx<double> incl_coef(const Incl& in, const int k[3], const SpecGrid& g) {
  double W;
  if (in.kind == INCL_CUBE || in.kind == INCL_PYRAMID) {
    // get_weights_con (materials.py:333-359): prod_a h_a sinc(h_a k_a / Y_a) / |Y|, h = side;
    // get_weights_lin (362-389): prod_a h_a sinc^2(h_a k_a / Y_a) / |Y|, h = side / 2.
    W = 1.0 / g.vol;
    for (int a = 0; a < 3; ++a) {
      if (g.nbar[a] == 1) continue;
      const double h = in.kind == INCL_CUBE ? in.p[a] : in.p[a] / 2.0;
      const double s = sinc_pi(h * (double)k[a] / g.Y[a]);
      W *= h * (in.kind == INCL_CUBE ? s : s * s);
    }
  } else if (in.kind == INCL_BALL) {
    // get_weights_circ (materials.py:392-425): r^2 J1(2 pi |xi| r) / (|xi| r) / |Y|, pi r^2 / |Y|
    // at xi = 0, with |xi|^2 = sum_a (k_a / Y_a)^2 over the problem's axes.
    const double r = in.p[0] / 2.0;
    double c2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      const double q = (double)k[a] / g.Y[a];
      c2 += q * q;
    }
    if (r == 0.0) {
      W = 0.0;
    } else if (c2 == 0.0) {
      W = 3.141592653589793 * r * r / g.vol;
    } else {
      const double c = sqrt(c2);
      W = r * r * j1(2.0 * 3.141592653589793 * c * r) / (c * r) / g.vol;
    }
  } else {
    return mk(0.0, 0.0);
  }
  // phase exp(-2 pi i pos.k_s / Y) at k and at -k (k_s: the shifted-order frequency)
  double pr = 1.0, pi = 0.0, mr = 1.0, mi = 0.0;
  for (int a = 0; a < 3; ++a) {
    if (g.nbar[a] == 1 || in.pos[a] == 0.0) continue;
    const double tp = in.pos[a] * (double)shifted_order_freq(k[a], g.nbar[a]) / g.Y[a];
    const double tm = in.pos[a] * (double)shifted_order_freq(-k[a], g.nbar[a]) / g.Y[a];
    double sp, cp, sm, cm;
    sincos(-2.0 * 3.141592653589793 * tp, &sp, &cp);
    sincos(-2.0 * 3.141592653589793 * tm, &sm, &cm);
    const double r1 = pr * cp - pi * sp, i1 = pr * sp + pi * cp;
    pr = r1;
    pi = i1;
    const double r2 = mr * cm - mi * sm, i2 = mr * sm + mi * cm;
    mr = r2;
    mi = i2;
  }
  // (W e^{-i phi(k)} + conj(W e^{-i phi(-k)})) / 2
  return mk(0.5 * W * (pr + mr), 0.5 * W * (pi - mi));
}

// Spectrum of coefficient component c of A_Ga = sum_ii v_ii chi_ii (materials.py:54-73),
// zero outside the Nbar index set, scaled by g.scale, written to out (m0, m1, m2h).
__global__ void __launch_bounds__(256)
k_ga_incl_spec(const Incl* __restrict__ incl, int nincl, int c, SpecGrid g, cpx<double>* __restrict__ out) {
  const long long nh = (long long)g.m0 * g.m1 * g.m2h;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    const int i2 = (int)(t % g.m2h);
    const long long r = t / g.m2h;
    const int i1 = (int)(r % g.m1), i0 = (int)(r / g.m1);
    const int k[3] = {sfreq(i0, g.m0), sfreq(i1, g.m1), i2};
    cpx<double> acc = mk(0.0, 0.0);
    const bool inband = abs(k[0]) <= g.band[0] && abs(k[1]) <= g.band[1] && k[2] <= g.band[2];
    if (inband) {
      const bool zero = k[0] == 0 && k[1] == 0 && k[2] == 0;
      cpx<double> prev = mk(0.0, 0.0);  // sum of the inclusions before 'otherwise'
      for (int ii = 0; ii < nincl; ++ii) {
        const int kind = incl[ii].kind;
        cpx<double> chi;
        if (kind == INCL_ALL) {
          chi = mk(zero ? 1.0 : 0.0, 0.0);
        } else if (kind == INCL_OTHERWISE) {
          chi = mk((zero ? 1.0 : 0.0) - prev.x, -prev.y);
        } else {
          chi = incl_coef(incl[ii], k, g);
        }
        prev = prev + chi;
        const double v = incl[ii].v[c];
        acc = mk(acc.x + v * chi.x, acc.y + v * chi.y);
      }
    }
    out[t] = mk(g.scale * acc.x, g.scale * acc.y);
  }
}

// Spectrum of A_Ga from nodal values on a grid P (materials.py:74-113):
// W(k) hA0(k mod P), hA0 the unnormalised DFT of the nodal values (half spectrum
// (p0, p1, p2/2+1)), W the box (pw = 1) or tent (pw = 2) weights of h = Y / P.
__global__ void __launch_bounds__(256)
k_ga_grid_spec(const cpx<double>* __restrict__ hA0, int p0, int p1, int p2, int pw, SpecGrid g,
               cpx<double>* __restrict__ out) {
  const long long nh = (long long)g.m0 * g.m1 * g.m2h;
  const int p2h = p2 / 2 + 1;
  const int P[3] = {p0, p1, p2};
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    const int i2 = (int)(t % g.m2h);
    const long long r = t / g.m2h;
    const int i1 = (int)(r % g.m1), i0 = (int)(r / g.m1);
    const int k[3] = {sfreq(i0, g.m0), sfreq(i1, g.m1), i2};
    cpx<double> val = mk(0.0, 0.0);
    if (abs(k[0]) <= g.band[0] && abs(k[1]) <= g.band[1] && k[2] <= g.band[2]) {
      double W = 1.0 / g.vol;
      for (int a = 0; a < 3; ++a) {
        if (g.nbar[a] == 1) continue;
        const double h = g.Y[a] / (double)P[a];
        const double s = sinc_pi(h * (double)k[a] / g.Y[a]);
        W *= h * (pw == 1 ? s : s * s);
      }
      const int j0 = ((k[0] % p0) + p0) % p0, j1 = ((k[1] % p1) + p1) % p1, j2 = ((k[2] % p2) + p2) % p2;
      cpx<double> h0;
      if (j2 < p2h) {
        h0 = hA0[((long long)j0 * p1 + j1) * p2h + j2];
      } else {
        const cpx<double> m = hA0[((long long)((p0 - j0) % p0) * p1 + (p1 - j1) % p1) * p2h + (p2 - j2)];
        h0 = mk(m.x, -m.y);
      }
      val = mk(g.scale * W * h0.x, g.scale * W * h0.y);
    }
    out[t] = val;
  }
}

// Trigonometric interpolation of a coarse spectrum (n0, n1, n2/2+1) onto the quadrature
// grid's half spectrum: frequencies |k_a| <= n_a / 2 copied (times scale), the Nyquist
// coefficient of an even axis split in halves between +n/2 and -n/2 (Tensor.enlarge,
// ffthompy/tensors/objects.py:428-467), zero elsewhere.
__global__ void __launch_bounds__(256)
k_ga_pad(const cpx<double>* __restrict__ in, int n0, int n1, int n2, double scale, SpecGrid g,
         cpx<double>* __restrict__ out) {
  const long long nh = (long long)g.m0 * g.m1 * g.m2h;
  const int n2h = n2 / 2 + 1;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    const int i2 = (int)(t % g.m2h);
    const long long r = t / g.m2h;
    const int i1 = (int)(r % g.m1), i0 = (int)(r / g.m1);
    const int k0 = sfreq(i0, g.m0), k1 = sfreq(i1, g.m1), k2 = i2;
    cpx<double> val = mk(0.0, 0.0);
    if (2 * abs(k0) <= n0 && 2 * abs(k1) <= n1 && 2 * k2 <= n2) {
      double f = scale;
      if ((n0 & 1) == 0 && 2 * abs(k0) == n0) f *= 0.5;
      if ((n1 & 1) == 0 && 2 * abs(k1) == n1) f *= 0.5;
      if ((n2 & 1) == 0 && 2 * k2 == n2) f *= 0.5;
      const int j0 = (k0 + n0) % n0, j1 = (k1 + n1) % n1;
      const cpx<double> c = in[((long long)j0 * n1 + j1) * n2h + k2];
      val = mk(f * c.x, f * c.y);
    }
    out[t] = val;
  }
}

// Restriction of a quadrature-grid spectrum to the coarse grid (Tensor.decrease,
// ffthompy/tensors/objects.py:469-486): coarse frequency k (centred set, the Nyquist index of an
// even axis standing for -n/2) takes the fine coefficient at k, times scale. On the last axis
// the half spectrum holds +n2/2 at index n2/2; its value there is the fine coefficient at -n2/2,
// i.e. the conjugate of the fine one at (-k0, -k1, +n2/2). One thread per coarse frequency.
__global__ void __launch_bounds__(256)
k_ga_crop(const cpx<double>* __restrict__ in, int n0, int n1, int n2, double scale, SpecGrid g,
          cpx<double>* __restrict__ out) {
  const int n2h = n2 / 2 + 1;
  const long long nh = (long long)n0 * n1 * n2h;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    const int k2 = (int)(t % n2h);
    const long long r = t / n2h;
    const int j1 = (int)(r % n1), j0 = (int)(r / n1);
    const int k0 = sfreq(j0, n0), k1 = sfreq(j1, n1);
    cpx<double> v;
    if ((n2 & 1) == 0 && 2 * k2 == n2) {
      const int i0 = ((-k0) % g.m0 + g.m0) % g.m0, i1 = ((-k1) % g.m1 + g.m1) % g.m1;
      const cpx<double> c = in[((long long)i0 * g.m1 + i1) * g.m2h + k2];
      v = mk(c.x, -c.y);
    } else {
      const int i0 = (k0 % g.m0 + g.m0) % g.m0, i1 = (k1 % g.m1 + g.m1) % g.m1;
      v = in[((long long)i0 * g.m1 + i1) * g.m2h + k2];
    }
    out[t] = mk(scale * v.x, scale * v.y);
  }
}

}  // namespace ga
}  // namespace fh
