// Inverse contiguous-axis pass (C2R and its epilogues) and the in-place round trip of the
// potential formulation, fixed lengths, single precision: their own translation unit so that the
// build compiles them in parallel with the other passes.
#include "fh_fixed_impl.hpp"

namespace fh {
template int fixed_rows_inv<float>(const LaunchCtx&, int, const cpx<float>*, float*, float, int, const float*, double*,
                                   size_t, const int*, int*, CGState*);
template int fixed_rows_rt<float>(const LaunchCtx&, int, cpx<float>*, const float*, float, const int*);
}  // namespace fh
