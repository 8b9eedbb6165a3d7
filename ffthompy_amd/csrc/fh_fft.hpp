// LDS-resident mixed-radix Stockham FFT engine for gfx950 (CDNA4).
//
// A workgroup holds a batch of pencils in LDS and transforms them together.
// Each stage is a radix-R Stockham step (natural-order output, no bit
// reversal): every thread loads its butterflies' R inputs into registers,
// applies twiddles and the radix-R DFT, then, after a barrier, stores the R
// outputs in place. Radices 2, 3, 4, 5, 7, 11, 13 are supported, which covers
// every grid length whose prime factors are at most 13 (the odd grids of the
// Fourier-Galerkin method: 5, 9, 15, 45, 63, 81, 99, 105, 135, 243, ...); any other
// length goes through Bluestein's chirp convolution on a 7-smooth length (fft_lds).
//
// The transform is unnormalised, forward sign exp(-2 pi i jk/L), like
// numpy.fft.fft, which the reference calls through ffthompy/tensors/fft.py:33-43.
#pragma once
#include <hip/hip_runtime.h>

namespace fh {

template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T a, T b) {
  cpx<T> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T>
__device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cpx<T> operator*(T s, cpx<T> a) { return mk(s * a.x, s * a.y); }
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__device__ __forceinline__ cpx<T> conj(cpx<T> a) { return mk(a.x, -a.y); }
// a * b + c with one rounding, in the precision of T
template <typename T>
__device__ __forceinline__ T fmadd(T a, T b, T c) {
  if constexpr (sizeof(T) == 4)
    return __builtin_fmaf(a, b, c);
  else
    return __builtin_fma(a, b, c);
}

// cos / sin of 2 pi m / R, m = 0..R-1 (40-digit values rounded to double).
template <int R>
struct RadixTab;
template <>
struct RadixTab<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
};
template <>
struct RadixTab<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410,
                                  -0.80901699437494742410, 0.30901699437494742410};
  static constexpr double s[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917,
                                  -0.58778525229247312917, -0.95105651629515357212};
};
template <>
struct RadixTab<7> {
  static constexpr double c[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429,
                                  -0.90096886790241912624, -0.90096886790241912624,
                                  -0.22252093395631440429, 0.62348980185873353053};
  static constexpr double s[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702,
                                  0.43388373911755812048, -0.43388373911755812048,
                                  -0.97492791218182360702, -0.78183148246802980871};
};
template <>
struct RadixTab<11> {
  static constexpr double c[11] = {1.0, 0.84125353283118116886, 0.41541501300188642553,
                                   -0.14231483827328514044, -0.65486073394528506406,
                                   -0.95949297361449738989, -0.95949297361449738989,
                                   -0.65486073394528506406, -0.14231483827328514044,
                                   0.41541501300188642553, 0.84125353283118116886};
  static constexpr double s[11] = {0.0, 0.54064081745559758211, 0.90963199535451837141,
                                   0.98982144188093273238, 0.75574957435425828377,
                                   0.28173255684142969771, -0.28173255684142969771,
                                   -0.75574957435425828377, -0.98982144188093273238,
                                   -0.90963199535451837141, -0.54064081745559758211};
};
template <>
struct RadixTab<13> {
  static constexpr double c[13] = {1.0, 0.88545602565320989590, 0.56806474673115580251,
                                   0.12053668025532305335, -0.35460488704253562597,
                                   -0.74851074817110109863, -0.97094181742605202716,
                                   -0.97094181742605202716, -0.74851074817110109863,
                                   -0.35460488704253562597, 0.12053668025532305335,
                                   0.56806474673115580251, 0.88545602565320989590};
  static constexpr double s[13] = {0.0, 0.46472317204376854565, 0.82298386589365639458,
                                   0.99270887409805399280, 0.93501624268541482344,
                                   0.66312265824079520238, 0.23931566428755776715,
                                   -0.23931566428755776715, -0.66312265824079520238,
                                   -0.93501624268541482344, -0.99270887409805399280,
                                   -0.82298386589365639458, -0.46472317204376854565};
};

// Radix-R DFT in registers. DIR = -1 forward, +1 inverse (unnormalised).
template <typename T, int R, int DIR>
__device__ __forceinline__ void dft(cpx<T> (&v)[R]) {
  if constexpr (R == 2) {
    const cpx<T> a = v[0], b = v[1];
    v[0] = a + b;
    v[1] = a - b;
  } else if constexpr (R == 4) {
    const cpx<T> s0 = v[0] + v[2], d0 = v[0] - v[2];
    const cpx<T> s1 = v[1] + v[3], d1 = v[1] - v[3];
    v[0] = s0 + s1;
    v[2] = s0 - s1;
    // forward: d0 - i d1 and d0 + i d1; inverse swaps them
    const cpx<T> m = mk(d0.x + d1.y, d0.y - d1.x);
    const cpx<T> p = mk(d0.x - d1.y, d0.y + d1.x);
    v[1] = DIR < 0 ? m : p;
    v[3] = DIR < 0 ? p : m;
  } else if constexpr (R == 3) {
    // a = x1 + x2, b = x1 - x2, y0 = x0 + a, A = x0 - a/2, y1,2 = A -+ i (sqrt3/2) b: the sine
    // term is folded into fused multiply-adds (12 operations per butterfly instead of 14)
    constexpr T s = T(0.86602540378443864676);
    const cpx<T> a = v[1] + v[2], b = v[1] - v[2];
    const cpx<T> x0 = v[0];
    const T Ax = fmadd(T(-0.5), a.x, x0.x), Ay = fmadd(T(-0.5), a.y, x0.y);
    const cpx<T> ym = mk(fmadd(s, b.y, Ax), fmadd(-s, b.x, Ay));  // A - i s b
    const cpx<T> yp = mk(fmadd(-s, b.y, Ax), fmadd(s, b.x, Ay));  // A + i s b
    v[0] = x0 + a;
    v[1] = DIR < 0 ? ym : yp;
    v[2] = DIR < 0 ? yp : ym;
  } else {
    // Odd R = 2h+1. With a_j = x_j + x_{R-j}, b_j = x_j - x_{R-j}:
    //   y_k = x_0 + sum_j a_j cos(2 pi jk/R)  -+  i sum_j b_j sin(2 pi jk/R)
    // (minus for the forward transform at y_k, plus at y_{R-k}).
    constexpr int H = R / 2;
    using Tab = RadixTab<R>;
    cpx<T> a[H], b[H];
    const cpx<T> x0 = v[0];
    cpx<T> y0 = x0;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      a[j - 1] = v[j] + v[R - j];
      b[j - 1] = v[j] - v[R - j];
      y0 = y0 + a[j - 1];
    }
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      cpx<T> A = x0;
      cpx<T> B = mk(T(0), T(0));
#pragma unroll
      for (int j = 1; j <= H; ++j) {
        const int m = (j * k) % R;
        A = A + T(Tab::c[m]) * a[j - 1];
        B = B + T(Tab::s[m]) * b[j - 1];
      }
      const cpx<T> ym = mk(A.x + B.y, A.y - B.x);  // A - iB
      const cpx<T> yp = mk(A.x - B.y, A.y + B.x);  // A + iB
      v[k] = DIR < 0 ? ym : yp;
      v[R - k] = DIR < 0 ? yp : ym;
    }
    v[0] = y0;
  }
}

constexpr int kMaxStages = 24;

// Radix schedule of one axis length; pre[s] is the product of the radices
// before stage s (the Stockham sub-transform length). A length L with a prime
// factor above 13 is transformed by Bluestein's algorithm (blue = 1): the DFT is the
// chirp-weighted circular convolution X_k = w_k sum_j (x_j w_j) conj(w_{k-j}),
// w_m = exp(-i pi m^2 / L), evaluated with forward and inverse transforms of a
// length M >= 2L-1 whose prime factors are at most 7; the stages then describe M.
// Otherwise M = L and blue = 0. Pencils in LDS are M elements long.
struct AxisPlan {
  int L;
  int M;
  int blue;
  int nst;
  int rad[kMaxStages];
  int pre[kMaxStages];
};

constexpr int kMaxElems = 16;  // register budget per thread and stage (complex values)

__host__ __device__ constexpr int max_butterflies(int R) {
  return (kMaxElems / R) > 0 ? (kMaxElems / R) : 1;
}

// One radix-R Stockham stage over npen pencils in LDS. Pencil q, element j
// lives at s[q*sp + j*sj]. rowmap: consecutive threads walk along a pencil
// (row layout, sj == 1); otherwise they walk across pencils (column layout).
// The host guarantees npen*(L/R) <= max_butterflies(R)*nthr.
template <typename T, int R, int DIR>
__device__ __forceinline__ void fft_stage(cpx<T>* s, int sp, int sj, int npen, int L, int p,
                                          const cpx<T>* __restrict__ tw, int tid, int nthr,
                                          bool rowmap) {
  constexpr int MAXM = max_butterflies(R);
  const int LR = L / R;
  const int total = npen * LR;
  const int tws = L / (p * R);
  cpx<T> v[MAXM][R];
  int ob[MAXM];
#pragma unroll
  for (int m = 0; m < MAXM; ++m) {
    const int w = tid + m * nthr;
    ob[m] = -1;
    if (w < total) {
      int q, i;
      if (rowmap) {
        q = w / LR;
        i = w - q * LR;
      } else {
        i = w / npen;
        q = w - i * npen;
      }
      const int k = i % p;
      const cpx<T>* src = s + q * sp + i * sj;
#pragma unroll
      for (int j = 0; j < R; ++j) v[m][j] = src[j * LR * sj];
      if (k != 0) {
#pragma unroll
        for (int j = 1; j < R; ++j) {
          cpx<T> t = tw[j * k * tws];
          if (DIR > 0) t.y = -t.y;
          v[m][j] = cmul(v[m][j], t);
        }
      }
      dft<T, R, DIR>(v[m]);
      ob[m] = q * sp + ((i - k) * R + k) * sj;
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MAXM; ++m) {
    if (ob[m] >= 0) {
#pragma unroll
      for (int j = 0; j < R; ++j) s[ob[m] + j * p * sj] = v[m][j];
    }
  }
  __syncthreads();
}

// The Stockham stages of ap (length ap.M) in direction DIR.
template <typename T, int DIR>
__device__ void fft_stages(cpx<T>* s, int sp, int sj, int npen, const AxisPlan& ap,
                           const cpx<T>* __restrict__ tw, int tid, int nthr, bool rowmap) {
  for (int st = 0; st < ap.nst; ++st) {
    const int R = ap.rad[st], p = ap.pre[st];
    switch (R) {
      case 2: fft_stage<T, 2, DIR>(s, sp, sj, npen, ap.M, p, tw, tid, nthr, rowmap); break;
      case 3: fft_stage<T, 3, DIR>(s, sp, sj, npen, ap.M, p, tw, tid, nthr, rowmap); break;
      case 4: fft_stage<T, 4, DIR>(s, sp, sj, npen, ap.M, p, tw, tid, nthr, rowmap); break;
      case 5: fft_stage<T, 5, DIR>(s, sp, sj, npen, ap.M, p, tw, tid, nthr, rowmap); break;
      case 7: fft_stage<T, 7, DIR>(s, sp, sj, npen, ap.M, p, tw, tid, nthr, rowmap); break;
      case 11: fft_stage<T, 11, DIR>(s, sp, sj, npen, ap.M, p, tw, tid, nthr, rowmap); break;
      case 13: fft_stage<T, 13, DIR>(s, sp, sj, npen, ap.M, p, tw, tid, nthr, rowmap); break;
      default: break;  // rejected at plan creation
    }
  }
}

// Length-L DFT (DIR -1 forward, +1 inverse, unnormalised) of npen pencils in LDS, in place;
// pencil q, element j at s[q*sp + j*sj]. The caller has synchronised the workgroup after writing
// the input; the output is visible to the whole workgroup on return.
// Twiddle table tw: W_M^m for m < M; with Bluestein (ap.blue) followed by the chirp w_m (m < L)
// and the spectra of the convolution kernel for both directions, FFT_M(conj w) / M and
// FFT_M(w) / M (M values each, the kernel extended periodically: index M - m holds index m).
template <typename T, int DIR>
__device__ void fft_lds(cpx<T>* s, int sp, int sj, int npen, const AxisPlan& ap,
                        const cpx<T>* __restrict__ tw, int tid, int nthr, bool rowmap) {
  if (!ap.blue) {
    fft_stages<T, DIR>(s, sp, sj, npen, ap, tw, tid, nthr, rowmap);
    return;
  }
  const int L = ap.L, M = ap.M;
  const cpx<T>* chirp = tw + M;
  const cpx<T>* kern = tw + M + L + (DIR < 0 ? 0 : M);
  // a_j = x_j w_j (conj w_j for the inverse), zero-padded to M
  for (int e = tid; e < npen * M; e += nthr) {
    const int q = e / M, j = e - q * M;
    cpx<T>* a = s + q * sp + j * sj;
    if (j < L) {
      cpx<T> w = chirp[j];
      if (DIR > 0) w.y = -w.y;
      *a = cmul(*a, w);
    } else {
      *a = mk(T(0), T(0));
    }
  }
  __syncthreads();
  fft_stages<T, -1>(s, sp, sj, npen, ap, tw, tid, nthr, rowmap);
  for (int e = tid; e < npen * M; e += nthr) {
    const int q = e / M, k = e - q * M;
    cpx<T>* a = s + q * sp + k * sj;
    *a = cmul(*a, kern[k]);
  }
  __syncthreads();
  fft_stages<T, +1>(s, sp, sj, npen, ap, tw, tid, nthr, rowmap);
  for (int e = tid; e < npen * L; e += nthr) {
    const int q = e / L, k = e - q * L;
    cpx<T>* a = s + q * sp + k * sj;
    cpx<T> w = chirp[k];
    if (DIR > 0) w.y = -w.y;
    *a = cmul(*a, w);
  }
  __syncthreads();
}

}  // namespace fh
