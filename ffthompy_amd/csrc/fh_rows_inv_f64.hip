// Inverse contiguous-axis pass (C2R and its epilogues) and the in-place round trip of the
// potential formulation, fixed lengths, double precision: their own translation unit so that the
// build compiles them in parallel with the other passes.
#include "fh_fixed_impl.hpp"

namespace fh {
template int fixed_rows_inv<double>(const LaunchCtx&, int, const cpx<double>*, double*, double, int, const double*, double*,
                                   size_t, const int*, int*, CGState*);
template int fixed_rows_rt<double>(const LaunchCtx&, int, cpx<double>*, const double*, double, const int*);
}  // namespace fh
