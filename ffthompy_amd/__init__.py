"""MI355X-native Fourier-Galerkin homogenization (the GaNi solve of FFTHomPy).

The operator y = F^-1 G(xi) F (A(x) v(x)), the conjugate-gradient solve and the
energy forms run as hand-written HIP kernels for gfx950 behind a C ABI
(include/ffthom.h), bound here with ctypes. The package mirrors the reference's
Problem / Tensor / linear_solver interface so the GPU operator is a drop-in.
"""
__version__ = '0.1.0'

from .tensors import DFT, GpuOperator, Tensor, scalar_product  # noqa: E402,F401
from .solver import CallBack, linear_solver  # noqa: E402,F401
from .problem import Problem, import_file  # noqa: E402,F401
