// Fixed-length kernels and launchers, double precision.
#include "fh_fixed_impl.hpp"

namespace fh {
template int fixed_rows_fwd<double>(const LaunchCtx&, int, const double*, const double*, int, cpx<double>*, const int*,
                                   const double*, const CGState*, double*, double*, size_t, int*);
template int fixed_rows_inv<double>(const LaunchCtx&, int, const cpx<double>*, double*, double, int, const double*, double*,
                                   size_t, const int*, int*, CGState*);
template int fixed_rows_rt<double>(const LaunchCtx&, int, cpx<double>*, const double*, double, const int*);
template int fixed_pot_cols<double>(const LaunchCtx&, int, int, int, cpx<double>*, cpx<double>*, const cpx<double>*,
                                    const CGState*, double, cpx<double>*, const cpx<double>*, double*, size_t,
                                    const pot::PotParams&, const int*, int*);
template int fixed_cols<double>(const LaunchCtx&, int, int, int, cpx<double>*, const int*);
}  // namespace fh
