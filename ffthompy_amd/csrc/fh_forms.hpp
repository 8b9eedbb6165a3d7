// Spectrum forms, shifts, resizing, projections of any form and pointwise inverses of fields that
// live on the GPU (kernels in fh_forms.hip): the data-movement side of the reference's Tensor
// (ffthompy/tensors/objects.py:135-186, 247-251, 406-511) and its projections
// (ffthompy/projections.py:9-112, 114-267). Grid-stride loops over the output, one element per
// thread and iteration; memory-bound gathers with index arithmetic, off the solver's hot path.
//
// Layouts (component-major, C order, last axis contiguous; grids of dimension 1 or 2 are the
// 3-D view with leading axes of length 1):
//  'r' spectrum: (count,) + (n0, n1, n2/2+1) complex, the unnormalised rfftn;
//  form 0:       (count,) + (n0, n1, n2) complex, fftn / prod(n);
//  form 'c':     form 0 with fftshift over the grid axes (zero frequency at index n//2).
#pragma once
#include <hip/hip_runtime.h>

#include "fh_kernels.hpp"

namespace fh {
namespace forms {

struct Grid3 {
  int n0, n1, n2, n2h;
};

enum Form { FORM_R = 0, FORM_FULL = 1, FORM_CENTRED = 2 };

// Host launchers (fh_forms.hip) on the null stream; dbl selects double or float. Each returns the
// launch status.
hipError_t half_to_full(int dbl, Grid3 g, long long count, const void* half, void* full, double scale, int centred);
hipError_t full_to_half(int dbl, Grid3 g, long long count, const void* full, void* half, double scale, int centred,
                        int hermitian);
hipError_t shift(Grid3 g, long long count, int words, const void* in, void* out, int inverse);
hipError_t resize_centred(int dbl, Grid3 a, Grid3 b, long long count, const void* in, void* out, int enlarge);
hipError_t project_form(int dbl, const GParams& g, int form, int null_nyq, const void* in, void* out);
hipError_t pointwise_inverse(int dbl, int k, long long nvox, const void* in, void* out);
// tpiY: 2 pi / Y per view axis (3 entries)
hipError_t spectrum_grad(int dbl, Grid3 g, int dim, const double* tpiY, const void* in, void* out);
hipError_t pointwise_product(int dbl, int cplx, int mode, int A, int B, long long nvox, const void* X, const void* Y,
                             void* out);

}  // namespace forms
}  // namespace fh
