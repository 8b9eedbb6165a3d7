// Forward contiguous-axis pass (coefficient, CG direction update, R2C) of the fixed lengths, single
// precision, every row-count variant: its own translation unit so that the build compiles it in
// parallel with the other passes.
#include "fh_fixed_impl.hpp"

namespace fh {
template int fixed_rows_fwd<float>(const LaunchCtx&, int, const float*, const float*, int, cpx<float>*, const int*,
                                   const float*, const CGState*, float*, double*, size_t, int*);
}  // namespace fh
