// Fixed-length kernels and launchers, single precision.
#include "fh_fixed_impl.hpp"

namespace fh {
template int fixed_rows_fwd<float>(const LaunchCtx&, int, const float*, const float*, int, cpx<float>*, const int*,
                                   const float*, const CGState*, float*, double*, size_t, int*);
template int fixed_rows_inv<float>(const LaunchCtx&, int, const cpx<float>*, float*, float, int, const float*, double*,
                                   size_t, const int*, int*, CGState*);
template int fixed_rows_rt<float>(const LaunchCtx&, int, cpx<float>*, const float*, float, const int*);
template int fixed_pot_cols<float>(const LaunchCtx&, int, int, int, cpx<float>*, cpx<float>*, const cpx<float>*,
                                   const CGState*, float, cpx<float>*, const cpx<float>*, double*, size_t,
                                   const pot::PotParams&, const int*, int*);
template int fixed_cols<float>(const LaunchCtx&, int, int, int, cpx<float>*, const int*);

bool fixed_length_supported(int L) {
  switch (L) {
#define FH_CASE(LL, FP, FPR) case LL:
    FH_FIXED_LENGTHS(FH_CASE)
#undef FH_CASE
    return true;
    default: return false;
  }
}
}  // namespace fh
