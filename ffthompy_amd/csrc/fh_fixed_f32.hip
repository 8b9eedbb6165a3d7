// Strided passes of the fixed lengths (and those of the potential formulation), single precision;
// the contiguous-axis and projection passes are compiled in their own translation units.
#include "fh_fixed_impl.hpp"

namespace fh {
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
