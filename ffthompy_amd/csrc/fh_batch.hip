// Persistent batched solve kernel (2-D scalar GaNi tasks); see fh_batch.hpp.
#include "fh_batch.hpp"

namespace fh {
namespace batch {

__global__ void k_batch_reset(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

namespace {

struct Slot {
  double *x[2], *b, *r64, *p64, *q64;
  cpx<double>* spec64;
  float *r32, *p32, *q32, *xc32;
};

__device__ __forceinline__ Slot slot_of(const Args& A) {
  Slot s;
  double* w = A.ws64 + (long long)blockIdx.x * A.ws64_stride;
  s.x[0] = w;
  s.x[1] = w + A.nvec;
  s.b = w + 2 * A.nvec;
  s.r64 = w + 3 * A.nvec;
  s.p64 = w + 4 * A.nvec;
  s.q64 = w + 5 * A.nvec;
  s.spec64 = reinterpret_cast<cpx<double>*>(w + 6 * A.nvec);
  float* f = A.ws32 + (long long)blockIdx.x * A.ws32_stride;
  s.r32 = f;
  s.p32 = f + A.nvec;
  s.q32 = f + 2 * A.nvec;
  s.xc32 = f + 3 * A.nvec;
  return s;
}

template <typename T>
__device__ __forceinline__ InOp<T> in_op(int mode, int coef, int nv, const T* a) {
  InOp<T> o;
  o.mode = mode;
  o.coef = coef;
  o.nv = nv;
  o.e0 = T(0);
  o.e1 = T(0);
  o.beta = T(0);
  o.v = nullptr;
  o.r = nullptr;
  o.a = a;
  return o;
}

template <typename T>
__device__ __forceinline__ OutOp<T> out_op(int mode, int nv) {
  OutOp<T> o;
  o.mode = mode;
  o.nv = nv;
  o.y = nullptr;
  o.p = nullptr;
  o.b = nullptr;
  o.r32 = nullptr;
  o.r64 = nullptr;
  o.acc = 0.0;
  return o;
}

// Phases of one load case. Every matrix-vector product of the task goes through one of
// two call sites (fp64 or fp32) so that each pass is instantiated once per length.
enum Phase : int {
  PH_RHS = 0,   // b = G(-c E_L)                        (fp64)
  PH_ITER,      // q = G(c p), <p, q>                    (fp32 mixed, fp64 plain)
  PH_TRUE,      // r = b - G(c x)                        (fp64, mixed only)
  PH_PROJ,      // x = G x                               (fp64, mixed only)
  PH_FOLD,      // d = G d, x += d: the fp32 correction   (fp64, mixed only)
  PH_DONE
};

// CG state of the load case in progress (uniform over the workgroup)
struct CG {
  double rr, rr_old, rmax, beta;
  int kit, nupd, status, first;
};

}  // namespace

template <int VAR>
__global__ void __launch_bounds__(NT) k_batch_solve(Args A) {
  __shared__ int sh_task;
  const int tid = threadIdx.x;
  const Params P = *A.par;
  const Slot s = slot_of(A);
  for (;;) {
    if (tid == 0) sh_task = atomicAdd(A.counter, 1);
    __syncthreads();
    const int task = sh_task;
    __syncthreads();
    if (task >= A.ntasks) break;
    const int code = A.order[task];
    const int si = code >> 1;
    const Sample sm = A.smp[si];
    Task tk;
    tk.n0 = sm.n0;
    tk.n1 = sm.n1;
    tk.h = sm.n1 / 2 + 1;
    tk.nv = sm.n0 * sm.n1;
    tk.pd = code & 1;
    tk.iY0 = sm.iY0;
    tk.iY1 = sm.iY1;
    tk.scale = 1.0 / (double)tk.nv;
    // twiddle tables of both axes into LDS
    {
      const int id0 = len_id(tk.n0), id1 = len_id(tk.n1);
      cpx<double>* t64 = lds_tw<double>(0);
      cpx<float>* t32 = lds_tw<float>(0);
      for (int e = tid; e < 2 * TWS; e += NT) {
        const int ax = e / TWS, m = e - ax * TWS;
        const int id = ax ? id1 : id0;
        t64[e] = A.tw64[id * TWS + m];
        t32[e] = A.tw32[id * TWS + m];
      }
    }
    const double* a64 = A.a64 + sm.off;
    const float* a32 = A.a32 + sm.off;
    const int base = si * 2 + tk.pd;
    const int coef = tk.pd + 1;
    const int nv = tk.nv, N = 2 * nv;
    const double inv_n = tk.scale;
    int status = 0;
#pragma unroll 1
    for (int L = 0; L < 2; ++L) {
      double* x = L ? s.x[1] : s.x[0];
      CG cg;
      cg.kit = 0;
      cg.nupd = 1;
      cg.status = 0;
      cg.first = 1;
      cg.beta = 0.0;
      cg.rr = cg.rr_old = cg.rmax = 0.0;
      int ph = PH_RHS;
#pragma unroll 1
      while (ph != PH_DONE) {
        // ---- the matrix-vector product of this phase
        double acc;
        if (ph == PH_ITER && P.mixed) {
          InOp<float> in = in_op<float>(cg.first ? 1 : 2, coef, nv, a32);
          in.v = s.p32;
          in.r = s.r32;
          in.beta = (float)cg.beta;
          OutOp<float> out = out_op<float>(1, nv);
          out.y = s.q32;
          out.p = s.p32;
          matvec<float, true, VAR>(tk, nullptr, in, out);
          acc = out.acc;
        } else {
          InOp<double> in = in_op<double>(1, coef, nv, a64);
          OutOp<double> out = out_op<double>(0, nv);
          if (ph == PH_RHS) {
            in.mode = 0;
            in.e0 = L == 0 ? -1.0 : 0.0;
            in.e1 = L == 1 ? -1.0 : 0.0;
            out.y = s.b;
          } else if (ph == PH_ITER) {
            in.mode = cg.first ? 1 : 2;
            in.v = s.p64;
            in.r = s.r64;
            in.beta = cg.beta;
            out.mode = 1;
            out.y = s.q64;
            out.p = s.p64;
          } else if (ph == PH_TRUE) {
            in.v = x;
            out.mode = 2;
            out.b = s.b;
            out.r32 = s.r32;
          } else if (ph == PH_FOLD) {
            in.coef = 0;
            in.v = s.q64;
            out.y = s.q64;
          } else {  // PH_PROJ
            in.coef = 0;
            in.v = x;
            out.y = x;
          }
          matvec<double, false, VAR>(tk, s.spec64, in, out);
          acc = out.acc;
        }
        // ---- what follows the product
        if (ph == PH_RHS) {
          double a2 = 0.0;
          for (int e = tid; e < N; e += NT) {
            const double v = s.b[e];
            x[e] = 0.0;
            if (P.mixed) {
              s.r32[e] = (float)v;
              s.p32[e] = (float)v;
              s.xc32[e] = 0.0f;
            } else {
              s.r64[e] = v;
              s.p64[e] = v;
            }
            a2 += v * v;
          }
          cg.rr = wg_sum(a2) * inv_n;
          cg.rmax = sqrt(cg.rr);
          const bool done = !(cg.rmax > P.tol) || P.maxiter <= 0;
          ph = done ? PH_DONE : PH_ITER;
        } else if (ph == PH_ITER) {
          cg.first = 0;
          const double pAp = wg_sum(acc) * inv_n;
          if (!(pAp > 0.0) || !isfinite(pAp)) {
            cg.status = 1;
            if (P.mixed) {
              for (int e = tid; e < N; e += NT) x[e] += (double)s.xc32[e];
            }
            ph = P.mixed ? PH_PROJ : PH_DONE;
            continue;
          }
          const double alpha = cg.rr / pAp;
          double a2 = 0.0;
          if (P.mixed) {
            const float af = (float)alpha;
            for (int e = tid; e < N; e += NT) {
              s.xc32[e] += af * s.p32[e];
              const float r = s.r32[e] - af * s.q32[e];
              s.r32[e] = r;
              a2 += (double)r * (double)r;
            }
          } else {
            for (int e = tid; e < N; e += NT) {
              x[e] += alpha * s.p64[e];
              const double r = s.r64[e] - alpha * s.q64[e];
              s.r64[e] = r;
              a2 += r * r;
            }
          }
          const double rn = wg_sum(a2) * inv_n;
          cg.kit += 1;
          cg.beta = rn / cg.rr;
          cg.rr_old = cg.rr;
          cg.rr = rn;
          const double nr = sqrt(rn);
          const bool conv = !(nr > P.tol), maxed = cg.kit >= P.maxiter;
          if (P.mixed) {
            if (conv || nr < P.delta * cg.rmax || maxed) {
              // reliable update: the fp32 correction, projected in fp64 (q64 is free in mixed
              // precision), is added to x; b - A x follows in fp64. Rounding leaves the correction
              // outside the range of G by about 1e-5 of its size: added unprojected, that part
              // would enter the later true residuals and be offset by an error inside the range
              for (int e = tid; e < N; e += NT) {
                s.q64[e] = (double)s.xc32[e];
                s.xc32[e] = 0.0f;
              }
              ph = PH_FOLD;
            }
          } else if (conv || maxed) {
            ph = PH_DONE;
          }
        } else if (ph == PH_FOLD) {
          __syncthreads();  // the projected correction of every thread is stored
          for (int e = tid; e < N; e += NT) x[e] += s.q64[e];
          ph = PH_TRUE;
        } else if (ph == PH_TRUE) {
          // the true residual differs from the recurrence by more than rounding (the projection
          // removed a part of the correction): the search direction restarts from it (p = r)
          const double rt = wg_sum(acc) * inv_n;
          cg.beta = 0.0;
          cg.rr = rt;
          cg.rmax = sqrt(rt);
          cg.nupd += 1;
          ph = (!(cg.rmax > P.tol) || cg.kit >= P.maxiter) ? PH_DONE : PH_ITER;
        } else {  // PH_PROJ done
          ph = PH_DONE;
        }
      }
      if (tid == 0) {
        A.kit[base * 2 + L] = cg.kit;
        A.nupd[base * 2 + L] = cg.nupd;
        A.normres[base * 2 + L] = sqrt(cg.rr);
      }
      status |= cg.status;
      __syncthreads();
    }
    // energies <c e_i, e_j>, e_i = E_i + x_i (ffthompy/postprocess.py:53-70)
    dd e00, e01, e11;
    e00.hi = e00.lo = e01.hi = e01.lo = e11.hi = e11.lo = 0.0;
    for (int v = tid; v < nv; v += NT) {
      double c = a64[v];
      if (tk.pd) c = 1.0 / c;
      const double u0 = s.x[0][v] + 1.0, u1 = s.x[0][nv + v];
      const double w0 = s.x[1][v], w1 = s.x[1][nv + v] + 1.0;
      dd t;
      t.lo = 0.0;
      t.hi = c * (u0 * u0 + u1 * u1);
      e00 = dd_add(e00, t);
      t.hi = c * (u0 * w0 + u1 * w1);
      e01 = dd_add(e01, t);
      t.hi = c * (w0 * w0 + w1 * w1);
      e11 = dd_add(e11, t);
    }
    e00 = wg_sum_dd(e00);
    e01 = wg_sum_dd(e01);
    e11 = wg_sum_dd(e11);
    if (tid == 0) {
      double* E = A.energy + base * 4;
      E[0] = (e00.hi + e00.lo) * tk.scale;
      E[1] = (e01.hi + e01.lo) * tk.scale;
      E[2] = E[1];
      E[3] = (e11.hi + e11.lo) * tk.scale;
      A.status[base] = status;
    }
    __syncthreads();
  }
}

template __global__ void k_batch_solve<0>(Args);
template __global__ void k_batch_solve<1>(Args);

hipError_t allow_lds(size_t lds) {
  hipError_t e = hipFuncSetAttribute((const void*)k_batch_solve<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void*)k_batch_solve<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  return e;
}

hipError_t enqueue(const Args& A, int nslots, size_t lds, hipStream_t stream, int var) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, stream, A.counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (var == 0)
    hipLaunchKernelGGL(k_batch_solve<0>, dim3(nslots), dim3(NT), lds, stream, A);
  else
    hipLaunchKernelGGL(k_batch_solve<1>, dim3(nslots), dim3(NT), lds, stream, A);
  return hipGetLastError();
}

}  // namespace batch
}  // namespace fh
