// Launchers of the fixed-length kernels, shared by the per-precision translation units
// (fh_fixed_f32.hip, fh_fixed_f64.hip) and the host plan code (fh_api.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include "fh_kernels.hpp"
#include "fh_pot.hpp"

namespace fh {

// What a launcher needs from a plan (no host types from fh_api.hip).
struct LaunchCtx {
  hipStream_t stream;
  int n0, n1, n2, n2h;
  long long n, nh;
  const void* tw[3];  // per-axis twiddle table W_L^m, m < L, in the launch precision
  hipEvent_t t0 = nullptr, t1 = nullptr;  // if set: start/stop time of the launched kernel
  int proj_var = 0;  // projection pass schedule variant (FxShape::proj_f)
  int cols_var = 0;  // strided pass schedule variant (FxShape::cols_p)
  int rows_var = 0;  // contiguous-axis pass schedule variant (FxShape::rows_rpb)
  // fused CG iteration on a small grid (alpha and beta without their one-workgroup kernels):
  // the forward row pass takes the deferred beta from the inverse pass's partials bpart[0..bnp),
  // the inverse row pass marks beta as deferred, and block 0 of the strided pass given `fin`
  // finalises beta and alpha (cg_fin_block)
  const double* bpart = nullptr;
  int bnp = 0;
  int defer_beta = 0;
  CgFin fin{nullptr, nullptr, nullptr, 0, 0};
  // inverse row pass epilogue 3 (double precision): the single-precision residual it writes
  float* r32 = nullptr;
};

// Number of schedule variants of the projection, strided and contiguous-axis passes.
constexpr int kProjVariants = 4, kColsVariants = 2, kRowsVariants = 3;

// Return 1 if a specialised kernel was launched, 0 if the length has none (caller falls
// back to the generic kernels), or a negative hipError_t on launch failure.
// rvec != null: `in` is the CG direction p, replaced by r + beta p (beta in *st) first.
// xc != null (with rvec, COEF_ISO/NONE): fused iteration, see k_rows_fwd_fx UPD 2; the
// partial sums of <p, a p> go to partial, the workgroup count to *nblocks.
template <typename T>
int fixed_rows_fwd(const LaunchCtx& c, int nc, const T* in, const T* coef, int kind, cpx<T>* out,
                   const int* halt, const T* rvec, const CGState* st, T* xc, double* partial,
                   size_t partial_cap, int* nblocks);
// epi 0: out = y; epi 1: also partial sums of <out, w>; epi 2: out is r, r -= alpha y and
// partial sums of <r, r>; epi 3 (double precision; out unused): c.r32 = float(w - y), the
// residual of b = w, and partial sums of <w - y, w - y> in double (k_rows_inv_st).
template <typename T>
int fixed_rows_inv(const LaunchCtx& c, int nc, const cpx<T>* in, T* out, T scale, int epi, const T* w,
                   double* partial, size_t partial_cap, const int* halt, int* nblocks, CGState* st);
// In place on the AoS half spectrum: C2R (times scale), isotropic a(x), R2C along the
// contiguous axis (k_rows_rt_st); nc 2 or 3.
template <typename T>
int fixed_rows_rt(const LaunchCtx& c, int nc, cpx<T>* spec, const T* coef, T scale, const int* halt);
// Potential formulation, strided pass along the outermost axis (0 in 3-D, 1 in 2-D) fused with
// the scalar <-> gradient spectra: dir +1 expands p (after p = r + beta p when r != null) and
// transforms back into data; dir -1 transforms data forward and writes sgn P div into out, with
// partial sums of the weighted <out, w> (w != null; *nblocks workgroups).
template <typename T>
int fixed_pot_cols(const LaunchCtx& c, int axis, int dir, int nc, cpx<T>* data, cpx<T>* p, const cpx<T>* r,
                   const CGState* st, T sgn, cpx<T>* out, const cpx<T>* w, double* partial, size_t partial_cap,
                   const pot::PotParams& g, const int* halt, int* nblocks);
template <typename T>
int fixed_cols(const LaunchCtx& c, int dir, int axis, int nc, cpx<T>* data, const int* halt);
template <typename T>
int fixed_cols_g(const LaunchCtx& c, int axis, int nc, cpx<T>* data, const GParams& g, const int* halt);
// Lengths with a specialised schedule.
bool fixed_length_supported(int L);

extern template int fixed_rows_fwd<float>(const LaunchCtx&, int, const float*, const float*, int, cpx<float>*,
                                          const int*, const float*, const CGState*, float*, double*, size_t,
                                          int*);
extern template int fixed_rows_fwd<double>(const LaunchCtx&, int, const double*, const double*, int,
                                           cpx<double>*, const int*, const double*, const CGState*, double*,
                                           double*, size_t, int*);
extern template int fixed_rows_inv<float>(const LaunchCtx&, int, const cpx<float>*, float*, float, int,
                                          const float*, double*, size_t, const int*, int*, CGState*);
extern template int fixed_rows_inv<double>(const LaunchCtx&, int, const cpx<double>*, double*, double, int,
                                           const double*, double*, size_t, const int*, int*, CGState*);
extern template int fixed_rows_rt<float>(const LaunchCtx&, int, cpx<float>*, const float*, float, const int*);
extern template int fixed_rows_rt<double>(const LaunchCtx&, int, cpx<double>*, const double*, double, const int*);
extern template int fixed_pot_cols<float>(const LaunchCtx&, int, int, int, cpx<float>*, cpx<float>*,
                                          const cpx<float>*, const CGState*, float, cpx<float>*, const cpx<float>*,
                                          double*, size_t, const pot::PotParams&, const int*, int*);
extern template int fixed_pot_cols<double>(const LaunchCtx&, int, int, int, cpx<double>*, cpx<double>*,
                                           const cpx<double>*, const CGState*, double, cpx<double>*,
                                           const cpx<double>*, double*, size_t, const pot::PotParams&, const int*,
                                           int*);
extern template int fixed_cols<float>(const LaunchCtx&, int, int, int, cpx<float>*, const int*);
extern template int fixed_cols<double>(const LaunchCtx&, int, int, int, cpx<double>*, const int*);
extern template int fixed_cols_g<float>(const LaunchCtx&, int, int, cpx<float>*, const GParams&, const int*);
extern template int fixed_cols_g<double>(const LaunchCtx&, int, int, cpx<double>*, const GParams&, const int*);

}  // namespace fh
