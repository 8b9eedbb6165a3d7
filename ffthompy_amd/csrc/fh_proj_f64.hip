// Projection pass (forward FFT along the outermost axis, G(xi), inverse FFT) of the fixed
// lengths, double precision, with every schedule variant: its own translation unit so that the
// build compiles it in parallel with the other passes.
#include "fh_fixed_impl.hpp"

namespace fh {
template int fixed_cols_g<double>(const LaunchCtx&, int, int, cpx<double>*, const GParams&, const int*);
}  // namespace fh
