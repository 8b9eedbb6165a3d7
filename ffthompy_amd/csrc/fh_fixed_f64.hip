// Strided passes of the fixed lengths (and those of the potential formulation), double precision;
// the contiguous-axis and projection passes are compiled in their own translation units.
#include "fh_fixed_impl.hpp"

namespace fh {
template int fixed_pot_cols<double>(const LaunchCtx&, int, int, int, cpx<double>*, cpx<double>*, const cpx<double>*,
                                    const CGState*, double, cpx<double>*, const cpx<double>*, double*, size_t,
                                    const pot::PotParams&, const int*, int*);
template int fixed_cols<double>(const LaunchCtx&, int, int, int, cpx<double>*, const int*);
}  // namespace fh
