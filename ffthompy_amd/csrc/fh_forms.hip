// Spectrum-form, shift, resize, projection and pointwise-inverse kernels and their launchers (on
// the null stream); layouts and entry points in fh_forms.hpp.
#include "fh_forms.hpp"

namespace fh {
namespace forms {

// Index of the centred array that holds entry i of the ifftshift-ordered one: fftshift moves
// index i to (i + n/2) mod n.
__device__ __forceinline__ int shifted(int i, int n) {
  const int j = i + n / 2;
  return j >= n ? j - n : j;
}

__device__ __forceinline__ void unflatten(long long r, int n1, int n2, int& i0, int& i1, int& i2) {
  i2 = (int)(r % n2);
  const long long q = r / n2;
  i1 = (int)(q % n1);
  i0 = (int)(q / n1);
}

// full[k] (form 0, or 'c' when centred) from the 'r' half spectrum: scale * half[k] for
// k2 <= n2/2, scale * conj(half[-k]) otherwise (ffthompy/tensors/objects.py:143-154).
template <typename T>
__global__ void __launch_bounds__(256)
k_half_to_full(const cpx<T>* __restrict__ half, cpx<T>* __restrict__ full, Grid3 g, long long count, T scale,
               int centred) {
  const long long nf = (long long)g.n0 * g.n1 * g.n2, nh = (long long)g.n0 * g.n1 * g.n2h;
  const long long total = count * nf;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long c = t / nf, r = t - c * nf;
    int i0, i1, i2;
    unflatten(r, g.n1, g.n2, i0, i1, i2);
    cpx<T> v;
    if (i2 < g.n2h) {
      v = half[c * nh + ((long long)i0 * g.n1 + i1) * g.n2h + i2];
    } else {
      const int j0 = i0 ? g.n0 - i0 : 0, j1 = i1 ? g.n1 - i1 : 0;
      v = conj(half[c * nh + ((long long)j0 * g.n1 + j1) * g.n2h + (g.n2 - i2)]);
    }
    long long o = r;
    if (centred) o = ((long long)shifted(i0, g.n0) * g.n1 + shifted(i1, g.n1)) * g.n2 + shifted(i2, g.n2);
    full[c * nf + o] = mk(scale * v.x, scale * v.y);
  }
}

// 'r' half spectrum from a full one (form 0, or 'c' when centred): scale * full[k] for
// k2 <= n2/2 (ffthompy/tensors/objects.py:155-163). hermitian = 1 takes the Hermitian part
// (full[k] + conj(full[-k])) / 2 instead: the inverse real transform of that half spectrum is the
// real part of the inverse complex transform of the full one (ffthompy/tensors/fft.py:24-37).
template <typename T>
__global__ void __launch_bounds__(256)
k_full_to_half(const cpx<T>* __restrict__ full, cpx<T>* __restrict__ half, Grid3 g, long long count, T scale,
               int centred, int hermitian) {
  const long long nf = (long long)g.n0 * g.n1 * g.n2, nh = (long long)g.n0 * g.n1 * g.n2h;
  const long long total = count * nh;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long c = t / nh, r = t - c * nh;
    int i0, i1, i2;
    unflatten(r, g.n1, g.n2h, i0, i1, i2);
    auto at = [&](int a0, int a1, int a2) -> cpx<T> {
      if (centred) {
        a0 = shifted(a0, g.n0);
        a1 = shifted(a1, g.n1);
        a2 = shifted(a2, g.n2);
      }
      return full[c * nf + ((long long)a0 * g.n1 + a1) * g.n2 + a2];
    };
    cpx<T> v = at(i0, i1, i2);
    if (hermitian) {
      const cpx<T> w = at(i0 ? g.n0 - i0 : 0, i1 ? g.n1 - i1 : 0, i2 ? g.n2 - i2 : 0);
      v = mk(T(0.5) * (v.x + w.x), T(0.5) * (v.y - w.y));
    }
    half[t] = mk(scale * v.x, scale * v.y);
  }
}

// fftshift (inverse = 0) or ifftshift (inverse = 1) over the grid axes of count fields whose
// elements are `words` 32-bit words (real or complex, either precision): form 0 <-> 'c' of a
// spectrum, origin 0 <-> 'c' of a real field (ffthompy/tensors/objects.py:164-186).
__global__ void __launch_bounds__(256)
k_shift(const unsigned* __restrict__ in, unsigned* __restrict__ out, Grid3 g, long long count, int words,
        int inverse) {
  const long long nf = (long long)g.n0 * g.n1 * g.n2;
  const long long total = count * nf;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long c = t / nf, r = t - c * nf;
    int i0, i1, i2;
    unflatten(r, g.n1, g.n2, i0, i1, i2);
    // fftshift: out[shifted(i)] = in[i]; ifftshift: out[i] = in[shifted(i)]
    const long long s = ((long long)shifted(i0, g.n0) * g.n1 + shifted(i1, g.n1)) * g.n2 + shifted(i2, g.n2);
    const long long src = inverse ? s : r, dst = inverse ? r : s;
    for (int w = 0; w < words; ++w) out[(c * nf + dst) * words + w] = in[(c * nf + src) * words + w];
  }
}

// Centred spectra from grid a (N) to grid b (M), per output entry; the entry of output index j
// has the integer frequency f = j - M/2 (floor), that of input index i has i - N/2.
//  enlarge (M >= N + [N even] per axis, Tensor.enlarge, ffthompy/tensors/objects.py:428-467):
//    each even axis is first extended by one entry, axis after axis in increasing order: the
//    entry at -N/2 is halved and a new entry at +N/2 holds half the conjugate of the entry at
//    -N/2 mirrored through the origin on the other axes (an axis of even current length keeps its
//    -n/2 entry in place); then the extended array is embedded into zeros. The gather below
//    unwinds those extensions from the last axis to the first.
//  decrease (M <= N per axis, Tensor.decrease, objects.py:469-486, ffthompy/trigpol.py:191-214):
//    the entries whose frequencies lie in M's index set, i.e. a centred crop.
template <typename T>
__global__ void __launch_bounds__(256)
k_resize_centred(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out, Grid3 a, Grid3 b, long long count,
                 int enlarge) {
  const long long na = (long long)a.n0 * a.n1 * a.n2, nb = (long long)b.n0 * b.n1 * b.n2;
  const long long total = count * nb;
  const int N[3] = {a.n0, a.n1, a.n2}, M[3] = {b.n0, b.n1, b.n2};
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const long long c = t / nb, r = t - c * nb;
    int j[3];
    unflatten(r, b.n1, b.n2, j[0], j[1], j[2]);
    int f[3];
    bool inside = true;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      f[ax] = j[ax] - M[ax] / 2;
      // frequencies present in the input (after the even-axis extensions when enlarging)
      const int lo = -(N[ax] / 2), hi = enlarge ? N[ax] / 2 : (N[ax] - 1) / 2;
      if (f[ax] < lo || f[ax] > hi) inside = false;
    }
    cpx<T> v = mk(T(0), T(0));
    if (inside) {
      T fac = T(1);
      bool cj = false;
      if (enlarge) {
#pragma unroll
        for (int ax = 2; ax >= 0; --ax) {
          if (N[ax] & 1) continue;
          const int h = N[ax] / 2;
          if (f[ax] == -h) {
            fac *= T(0.5);
          } else if (f[ax] == h) {
            fac *= T(0.5);
            cj = !cj;
            f[ax] = -h;
#pragma unroll
            for (int o = 0; o < 3; ++o) {
              if (o == ax) continue;
              // axes extended before this one have odd length by then: plain mirror; later axes
              // of even length keep their -n/2 entry
              if (o > ax && (N[o] & 1) == 0 && f[o] == -(N[o] / 2)) continue;
              f[o] = -f[o];
            }
          }
        }
      }
      const long long i = ((long long)(f[0] + N[0] / 2) * a.n1 + (f[1] + N[1] / 2)) * a.n2 + (f[2] + N[2] / 2);
      v = in[c * na + i];
      v = mk(fac * v.x, cj ? -fac * v.y : fac * v.y);
    }
    out[t] = v;
  }
}

// Signed integer frequency of index i on an axis of n points in a layout: ifftshift order
// (forms 'r' and 0: the Nyquist index n/2 of an even axis is -n/2, ffthompy/trigpol.py:27-40) or
// centred ('c': i - n//2). `nyq` is set at -n/2 of an even axis.
__device__ __forceinline__ int form_freq(int i, int n, int form, bool& nyq) {
  int s;
  if (form == FORM_CENTRED)
    s = i - n / 2;
  else
    s = (2 * i >= n) ? i - n : i;
  if ((n & 1) == 0 && 2 * s == -n) nyq = true;
  return s;
}

// G1 or G2 (g.dual) applied to g.ncomp component spectra of any form (ffthompy/projections.py:
// 9-112, 114-267, applied as the einsum of ffthompy/tensors/objects.py:223-232), times g.scale.
// null_nyq = 1 gives the reference's NyqNul=True (G built on the odd grid and zero-padded: zero
// wherever an even axis sits at its Nyquist index); with 0 every mode keeps G of its frequency,
// -n/2 at the Nyquist index. `l2` is the layout's last-axis length (n2/2+1 or n2).
template <typename T>
__global__ void __launch_bounds__(256)
k_project_form(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out, long long nspec, int l2, int form,
               int null_nyq, GParams g) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nspec;
       i += (long long)gridDim.x * blockDim.x) {
    int i0, i1, i2;
    unflatten(i, g.n1, l2, i0, i1, i2);
    bool nyq = false;
    const int s0 = form_freq(i0, g.n0, form, nyq);
    const int s1 = form_freq(i1, g.n1, form, nyq);
    const int s2 = form_freq(i2, g.n2, form, nyq);
    cpx<T> v[kMaxComp];
    for (int c = 0; c < g.ncomp; ++c) v[c] = in[c * nspec + i];
    apply_G_at<T>(v, s0, s1, s2, null_nyq && nyq, g);
    for (int c = 0; c < g.ncomp; ++c) out[c * nspec + i] = v[c];
  }
}

// Gradient of a scalar 'r' spectrum, out[j][t] = 2 pi i xi_j in[t] for the d grid axes
// (Tensor grad of a scalar, ffthompy/tensors/operators.py:227-260, with the frequencies of
// ffthompy/trigpol.py:27-40: the Nyquist index of an even axis stands for -n/2). Component-major
// output in the reference layout (d,) + half grid. tpiY[a] = 2 pi / Y of view axis a.
template <typename T>
__global__ void __launch_bounds__(256)
k_spectrum_grad(const cpx<T>* __restrict__ in, cpx<T>* __restrict__ out, Grid3 g, int dim, double t0, double t1,
                double t2) {
  const long long nh = (long long)g.n0 * g.n1 * g.n2h;
  const double tp[3] = {t0, t1, t2};
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nh;
       t += (long long)gridDim.x * blockDim.x) {
    int i0, i1, i2;
    unflatten(t, g.n1, g.n2h, i0, i1, i2);
    bool nyq = false;
    const int k[3] = {form_freq(i0, g.n0, FORM_FULL, nyq), form_freq(i1, g.n1, FORM_FULL, nyq),
                      form_freq(i2, g.n2, FORM_FULL, nyq)};
    const cpx<T> v = in[t];
    for (int j = 0; j < dim; ++j) {
      const int a = 3 - dim + j;
      const T q = T((double)k[a] * tp[a]);
      out[j * nh + t] = mk(-q * v.y, q * v.x);
    }
  }
}

// Pointwise products of Tensor.__mul__ (ffthompy/tensors/objects.py:223-245, the einsums of
// objects.py:599-604) for fields of nvox grid points, component-major:
//  mode 0: out[a][v] = sum_b X[a][b][v] * Y[b][v]   (multype 21: A = i, B = j; 42: A = ij, B = kl;
//          'grad': B = 1; 'div': A = 1)
//  mode 1: out[a][v] = X[a][v] * Y[a][v]             (multype 00, elementwise)
// V is the element type (real or complex of either precision).
template <typename V>
__device__ __forceinline__ V vmul(V a, V b) {
  return a * b;
}
template <>
__device__ __forceinline__ cpx<float> vmul(cpx<float> a, cpx<float> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <>
__device__ __forceinline__ cpx<double> vmul(cpx<double> a, cpx<double> b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename V>
__global__ void __launch_bounds__(256)
k_pointwise_product(const V* __restrict__ X, const V* __restrict__ Y, V* __restrict__ out, int A, int B,
                    long long nvox, int mode) {
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
       v += (long long)gridDim.x * blockDim.x) {
    for (int a = 0; a < A; ++a) {
      if (mode == 1) {
        out[(long long)a * nvox + v] = vmul(X[(long long)a * nvox + v], Y[(long long)a * nvox + v]);
        continue;
      }
      V acc = vmul(X[(long long)a * B * nvox + v], Y[v]);
      for (int b = 1; b < B; ++b) acc = acc + vmul(X[((long long)a * B + b) * nvox + v], Y[(long long)b * nvox + v]);
      out[(long long)a * nvox + v] = acc;
    }
  }
}

// Pointwise inverse of a (K, K) matrix field, entry (i, j) of voxel v at [(i*K + j)*nvox + v]:
// Gauss-Jordan elimination without pivoting in the order of ffthompy/trigpol.py:120-159 (row m
// normalised by its pivot and subtracted from the rows below, then a backward sweep).
template <typename T, int K>
__global__ void __launch_bounds__(256)
k_pointwise_inverse(const T* __restrict__ in, T* __restrict__ out, long long nvox) {
  for (long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
       v += (long long)gridDim.x * blockDim.x) {
    T B[K][K], I[K][K];
#pragma unroll
    for (int r = 0; r < K; ++r)
#pragma unroll
      for (int c = 0; c < K; ++c) {
        B[r][c] = in[(long long)(r * K + c) * nvox + v];
        I[r][c] = T(r == c);
      }
#pragma unroll
    for (int m = 0; m < K; ++m) {
      const T piv = B[m][m];
      B[m][m] = T(1);
#pragma unroll
      for (int c = m + 1; c < K; ++c) B[m][c] = B[m][c] / piv;
#pragma unroll
      for (int c = 0; c < K; ++c) I[m][c] = I[m][c] / piv;
#pragma unroll
      for (int r = m + 1; r < K; ++r) {
        const T f = B[r][m];
#pragma unroll
        for (int c = 0; c < K; ++c) {
          B[r][c] = B[r][c] - B[m][c] * f;
          I[r][c] = I[r][c] - I[m][c] * f;
        }
      }
    }
#pragma unroll
    for (int m = K - 1; m >= 0; --m) {
#pragma unroll
      for (int r = m - 1; r >= 0; --r) {
        const T f = B[r][m];
#pragma unroll
        for (int c = 0; c < K; ++c) {
          B[r][c] = B[r][c] - B[m][c] * f;
          I[r][c] = I[r][c] - I[m][c] * f;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < K; ++r)
#pragma unroll
      for (int c = 0; c < K; ++c) out[(long long)(r * K + c) * nvox + v] = I[r][c];
  }
}

namespace {

dim3 blocks_for(long long total) {
  const long long b = (total + 255) / 256;
  return dim3((unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b)));
}

}  // namespace

hipError_t half_to_full(int dbl, Grid3 g, long long count, const void* half, void* full, double scale, int centred) {
  const dim3 grid = blocks_for(count * g.n0 * g.n1 * g.n2);
  if (dbl)
    hipLaunchKernelGGL(k_half_to_full<double>, grid, dim3(256), 0, 0, (const cpx<double>*)half, (cpx<double>*)full,
                       g, count, scale, centred);
  else
    hipLaunchKernelGGL(k_half_to_full<float>, grid, dim3(256), 0, 0, (const cpx<float>*)half, (cpx<float>*)full, g,
                       count, (float)scale, centred);
  return hipGetLastError();
}

hipError_t full_to_half(int dbl, Grid3 g, long long count, const void* full, void* half, double scale, int centred,
                        int hermitian) {
  const dim3 grid = blocks_for(count * g.n0 * g.n1 * g.n2h);
  if (dbl)
    hipLaunchKernelGGL(k_full_to_half<double>, grid, dim3(256), 0, 0, (const cpx<double>*)full, (cpx<double>*)half,
                       g, count, scale, centred, hermitian);
  else
    hipLaunchKernelGGL(k_full_to_half<float>, grid, dim3(256), 0, 0, (const cpx<float>*)full, (cpx<float>*)half, g,
                       count, (float)scale, centred, hermitian);
  return hipGetLastError();
}

hipError_t shift(Grid3 g, long long count, int words, const void* in, void* out, int inverse) {
  const dim3 grid = blocks_for(count * g.n0 * g.n1 * g.n2);
  hipLaunchKernelGGL(k_shift, grid, dim3(256), 0, 0, (const unsigned*)in, (unsigned*)out, g, count, words, inverse);
  return hipGetLastError();
}

hipError_t resize_centred(int dbl, Grid3 a, Grid3 b, long long count, const void* in, void* out, int enlarge) {
  const dim3 grid = blocks_for(count * b.n0 * b.n1 * b.n2);
  if (dbl)
    hipLaunchKernelGGL(k_resize_centred<double>, grid, dim3(256), 0, 0, (const cpx<double>*)in, (cpx<double>*)out,
                       a, b, count, enlarge);
  else
    hipLaunchKernelGGL(k_resize_centred<float>, grid, dim3(256), 0, 0, (const cpx<float>*)in, (cpx<float>*)out, a,
                       b, count, enlarge);
  return hipGetLastError();
}

hipError_t project_form(int dbl, const GParams& g, int form, int null_nyq, const void* in, void* out) {
  const int l2 = form == FORM_R ? g.n2 / 2 + 1 : g.n2;
  const long long nspec = (long long)g.n0 * g.n1 * l2;
  const dim3 grid = blocks_for(nspec);
  if (dbl)
    hipLaunchKernelGGL(k_project_form<double>, grid, dim3(256), 0, 0, (const cpx<double>*)in, (cpx<double>*)out,
                       nspec, l2, form, null_nyq, g);
  else
    hipLaunchKernelGGL(k_project_form<float>, grid, dim3(256), 0, 0, (const cpx<float>*)in, (cpx<float>*)out, nspec,
                       l2, form, null_nyq, g);
  return hipGetLastError();
}

hipError_t pointwise_product(int dbl, int cplx, int mode, int A, int B, long long nvox, const void* X, const void* Y,
                             void* out) {
  const dim3 grid = blocks_for(nvox);
  if (cplx && dbl)
    hipLaunchKernelGGL(k_pointwise_product<cpx<double>>, grid, dim3(256), 0, 0, (const cpx<double>*)X,
                       (const cpx<double>*)Y, (cpx<double>*)out, A, B, nvox, mode);
  else if (cplx)
    hipLaunchKernelGGL(k_pointwise_product<cpx<float>>, grid, dim3(256), 0, 0, (const cpx<float>*)X,
                       (const cpx<float>*)Y, (cpx<float>*)out, A, B, nvox, mode);
  else if (dbl)
    hipLaunchKernelGGL(k_pointwise_product<double>, grid, dim3(256), 0, 0, (const double*)X, (const double*)Y,
                       (double*)out, A, B, nvox, mode);
  else
    hipLaunchKernelGGL(k_pointwise_product<float>, grid, dim3(256), 0, 0, (const float*)X, (const float*)Y,
                       (float*)out, A, B, nvox, mode);
  return hipGetLastError();
}

hipError_t spectrum_grad(int dbl, Grid3 g, int dim, const double* tpiY, const void* in, void* out) {
  const dim3 grid = blocks_for((long long)g.n0 * g.n1 * g.n2h);
  if (dbl)
    hipLaunchKernelGGL(k_spectrum_grad<double>, grid, dim3(256), 0, 0, (const cpx<double>*)in, (cpx<double>*)out, g,
                       dim, tpiY[0], tpiY[1], tpiY[2]);
  else
    hipLaunchKernelGGL(k_spectrum_grad<float>, grid, dim3(256), 0, 0, (const cpx<float>*)in, (cpx<float>*)out, g,
                       dim, tpiY[0], tpiY[1], tpiY[2]);
  return hipGetLastError();
}

template <int K>
hipError_t inverse_k(int dbl, long long nvox, const void* in, void* out) {
  const dim3 grid = blocks_for(nvox);
  if (dbl)
    hipLaunchKernelGGL((k_pointwise_inverse<double, K>), grid, dim3(256), 0, 0, (const double*)in, (double*)out,
                       nvox);
  else
    hipLaunchKernelGGL((k_pointwise_inverse<float, K>), grid, dim3(256), 0, 0, (const float*)in, (float*)out, nvox);
  return hipGetLastError();
}

hipError_t pointwise_inverse(int dbl, int k, long long nvox, const void* in, void* out) {
  switch (k) {
    case 1: return inverse_k<1>(dbl, nvox, in, out);
    case 2: return inverse_k<2>(dbl, nvox, in, out);
    case 3: return inverse_k<3>(dbl, nvox, in, out);
    case 4: return inverse_k<4>(dbl, nvox, in, out);
    case 5: return inverse_k<5>(dbl, nvox, in, out);
    case 6: return inverse_k<6>(dbl, nvox, in, out);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace forms
}  // namespace fh
