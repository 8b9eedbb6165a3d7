// Forward contiguous-axis pass (coefficient, CG direction update, R2C) of the fixed lengths, double
// precision, every row-count variant: its own translation unit so that the build compiles it in
// parallel with the other passes.
#include "fh_fixed_impl.hpp"

namespace fh {
template int fixed_rows_fwd<double>(const LaunchCtx&, int, const double*, const double*, int, cpx<double>*, const int*,
                                   const double*, const CGState*, double*, double*, size_t, int*);
}  // namespace fh
