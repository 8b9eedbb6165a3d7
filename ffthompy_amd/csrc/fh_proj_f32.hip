// Projection pass (forward FFT along the outermost axis, G(xi), inverse FFT) of the fixed
// lengths, single precision, with every schedule variant: its own translation unit so that the
// build compiles it in parallel with the other passes.
#include "fh_fixed_impl.hpp"

namespace fh {
template int fixed_cols_g<float>(const LaunchCtx&, int, int, cpx<float>*, const GParams&, const int*);
}  // namespace fh
