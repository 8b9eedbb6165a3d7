"""ctypes binding of the gfx950 C ABI (include/ffthom.h) and device arrays.

The HIP library is the only compute path: if it is missing or no gfx950 GPU is
present, every entry point raises ``FhError``. There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'lib', 'libffthom.so')

FH_OK, FH_ERR_ARG, FH_ERR_HIP, FH_ERR_NODEV, FH_ERR_UNSUPPORTED, FH_ERR_STATE = range(6)
FH_SCALAR, FH_ELASTICITY = 0, 1
FH_PRIMAL, FH_DUAL = 0, 1
FH_F64, FH_F32 = 0, 1
FH_SOLVE_FP64, FH_SOLVE_MIXED = 0, 1
FH_COEF_ISO, FH_COEF_LAME, FH_COEF_FULL = 1, 2, 3
FH_ITER_RICHARDSON, FH_ITER_CHEBYSHEV = 1, 2
FH_CG_OK, FH_CG_BREAKDOWN = 0, 1
FH_FORM_R, FH_FORM_FULL, FH_FORM_CENTRED = 0, 1, 2

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p

# name -> (argtypes); every function returns int except fh_last_error
SIGNATURES = {
    'fh_init': [ctypes.c_char_p, ctypes.c_int],
    'fh_malloc': [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t],
    'fh_free': [_vp],
    'fh_memcpy_htod': [_vp, _vp, ctypes.c_size_t],
    'fh_memcpy_dtoh': [_vp, _vp, ctypes.c_size_t],
    'fh_memcpy_dtod': [_vp, _vp, ctypes.c_size_t],
    'fh_memset_zero': [_vp, ctypes.c_size_t],
    'fh_plan_create': [ctypes.c_int, _c_int_p, _c_double_p, ctypes.c_int, ctypes.c_int,
                       ctypes.POINTER(ctypes.c_void_p)],
    'fh_plan_destroy': [_vp],
    'fh_plan_ncomp': [_vp],
    'fh_plan_set_primaldual': [_vp, ctypes.c_int],
    'fh_plan_use_generic': [_vp, ctypes.c_int],
    'fh_plan_set_cg_variant': [_vp, ctypes.c_int],
    'fh_plan_use_graphs': [_vp, ctypes.c_int],
    'fh_last_solve_graph_batches': [_vp, _c_int_p],
    'fh_set_coeff': [_vp, ctypes.c_int, _c_double_p],
    'fh_matvec': [_vp, ctypes.c_int, _vp, _vp],
    'fh_project': [_vp, ctypes.c_int, _vp, _vp],
    'fh_rfftn': [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp],
    'fh_irfftn': [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp],
    'fh_apply_projection_hat': [_vp, ctypes.c_int, _vp, _vp],
    'fh_cg_solve': [_vp, ctypes.c_int, _vp, _vp, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                    _c_int_p, _c_double_p, _c_double_p, ctypes.c_int, _c_int_p],
    'fh_energy_matrix': [_vp, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, _c_double_p],
    'fh_energy_matrix_noalloc': [ctypes.c_longlong, ctypes.c_int, _vp, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                 _c_double_p],
    'fh_energy_gradient': [_vp, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, _c_double_p, _vp, _vp],
    'fh_field_gradient': [_vp, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, _vp,
                          _vp],
    'fh_axpby': [ctypes.c_int, ctypes.c_size_t, ctypes.c_double, _vp, ctypes.c_double, _vp],
    'fh_dot': [ctypes.c_int, ctypes.c_size_t, _vp, _vp, _c_double_p],
    'fh_add_constant': [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, _c_double_p, _vp],
    'fh_spectrum_half_to_full': [ctypes.c_int, ctypes.c_int, _c_int_p, ctypes.c_longlong, _vp, _vp, ctypes.c_double,
                                 ctypes.c_int],
    'fh_spectrum_full_to_half': [ctypes.c_int, ctypes.c_int, _c_int_p, ctypes.c_longlong, _vp, _vp, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int],
    'fh_fftshift': [ctypes.c_int, ctypes.c_int, _c_int_p, ctypes.c_longlong, _vp, _vp, ctypes.c_int],
    'fh_spectrum_resize': [ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p, ctypes.c_longlong, _vp, _vp, ctypes.c_int],
    'fh_project_spectrum': [ctypes.c_int, ctypes.c_int, _c_int_p, _c_double_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, _vp, _vp],
    'fh_pointwise_inverse': [ctypes.c_int, ctypes.c_int, ctypes.c_longlong, _vp, _vp],
    'fh_spectrum_grad': [ctypes.c_int, ctypes.c_int, _c_int_p, _c_double_p, _vp, _vp],
    'fh_pointwise_product': [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_longlong,
                             _vp, _vp, _vp],
    'fh_last_timing': [_vp, _c_double_p, _c_double_p],
    'fh_last_solve_updates': [_vp, _c_int_p],
    'fh_matvec_profile': [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, _c_double_p, _c_double_p,
                          _c_int_p],
    'fh_plan_set_schedule': [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
    'fh_plan_get_schedule': [_vp, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p],
    'fh_plan_autotune': [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _c_double_p, _c_double_p, _c_double_p, _c_int_p,
                         _c_int_p, _c_int_p],
    'fh_batch_create': [ctypes.c_int, _c_int_p, _c_double_p, ctypes.POINTER(ctypes.c_void_p)],
    'fh_batch_destroy': [_vp],
    'fh_batch_set_coeff': [_vp, _c_double_p],
    'fh_batch_solve': [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double, _c_double_p, _c_double_p,
                       _c_int_p, _c_int_p],
    'fh_batch_stats': [_vp, _c_int_p, _c_double_p, _c_double_p],
    'fh_batch_set_schedule': [_vp, ctypes.c_int, ctypes.c_int],
    'fh_batch_autotune': [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _c_double_p, _c_int_p, _c_int_p],
    'fh_ga_create': [_vp, _c_int_p, ctypes.POINTER(ctypes.c_void_p)],
    'fh_ga_set_plan_coeff': [_vp, _vp],
    'fh_ga_energy_on_grid': [_vp, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, _c_double_p],
    'fh_plan_set_band': [_vp, _c_int_p, ctypes.c_double],
    'fh_ga_destroy': [_vp],
    'fh_ga_grid': [_vp, _c_int_p],
    'fh_ga_set_inclusions': [_vp, ctypes.c_int, _c_int_p, _c_double_p, _c_double_p, ctypes.c_int, _c_double_p],
    'fh_ga_set_grid_coeff': [_vp, ctypes.c_int, _c_int_p, ctypes.c_int, _c_double_p],
    'fh_ga_coefficient': [_vp, _c_double_p],
    'fh_ga_interpolate': [_vp, _vp, _vp],
    'fh_ga_restrict': [_vp, _vp, _vp],
    'fh_ga_energy': [_vp, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, _c_double_p],
    'fh_iter_solve': [_vp, ctypes.c_int, _vp, _vp, _c_double_p, ctypes.c_double, ctypes.c_int, _c_int_p, _c_double_p,
                      _c_double_p, ctypes.c_int],
    'fh_plan_set_residual_log': [_vp, ctypes.c_int],
    'fh_last_residual_log': [_vp, _c_int_p, _c_double_p, ctypes.c_int, _c_int_p],
    'fh_plan_set_iterate_callback': [_vp, _vp, _vp, ctypes.c_int],
    'fh_potential_solve': [_vp, ctypes.c_int, _c_double_p, ctypes.c_double, ctypes.c_int, ctypes.c_double, _vp, _vp,
                           _c_int_p, _c_double_p, _c_double_p, ctypes.c_int, _c_int_p],
    'fh_set_coeff_inclusions': [_vp, ctypes.c_int, _c_int_p, _c_double_p, _c_double_p, ctypes.c_int, _c_double_p,
                                ctypes.c_int],
    'fh_coeff_to_host': [_vp, _c_double_p],
    'fh_set_coeff_device': [_vp, ctypes.c_int, _vp],
    'fh_grf': [_vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _vp],
    'fh_synchronize': [],
}


# int fn(void* user, int kit, const void* x): the library's host callback on the iterate
ITERATE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p)


class FhError(RuntimeError):
    """Failure reported by the HIP library (or the library is missing)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


_LIB = None


def load_library():
    """Load the built library and declare every exported symbol. Raises if absent."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise FhError("HIP library %s is not built; run __graft_entry__.build()" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    lib.fh_last_error.restype = ctypes.c_char_p
    lib.fh_last_error.argtypes = []
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    _LIB = lib
    return lib


def call(name, *args):
    """Call an fh_* entry point; raise FhError with the library's message on failure."""
    lib = load_library()
    rc = getattr(lib, name)(*args)
    if rc != FH_OK:
        msg = lib.fh_last_error().decode(errors='replace')
        raise FhError('%s failed (code %d): %s' % (name, rc, msg), rc)
    return rc


_DEVICE_NAME = None


def init():
    """Initialise the GPU (device 0, must be gfx950). Returns the device name."""
    global _DEVICE_NAME
    if _DEVICE_NAME is None:
        buf = ctypes.create_string_buffer(256)
        call('fh_init', buf, 256)
        _DEVICE_NAME = buf.value.decode()
    return _DEVICE_NAME


_DTYPE_CODE = {np.dtype(np.float64): FH_F64, np.dtype(np.float32): FH_F32}


def dtype_code(dtype):
    try:
        return _DTYPE_CODE[np.dtype(dtype)]
    except KeyError:
        raise FhError('unsupported dtype %s (float64 or float32)' % dtype)


class DeviceArray:
    """A device allocation with a NumPy-like shape and dtype (float64, float32,
    complex128, complex64). Freed when garbage collected."""

    def __init__(self, shape, dtype=np.float64):
        init()
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        call('fh_malloc', ctypes.byref(p), ctypes.c_size_t(max(self.nbytes, 1)))
        self._ptr = p

    @property
    def ptr(self):
        return self._ptr

    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def ndim(self):
        return len(self.shape)

    @classmethod
    def from_host(cls, arr, dtype=None):
        arr = np.ascontiguousarray(arr if dtype is None else np.asarray(arr, dtype=dtype))
        out = cls(arr.shape, arr.dtype)
        call('fh_memcpy_htod', out.ptr, arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(arr.nbytes))
        return out

    @classmethod
    def zeros(cls, shape, dtype=np.float64):
        out = cls(shape, dtype)
        call('fh_memset_zero', out.ptr, ctypes.c_size_t(out.nbytes))
        return out

    def to_host(self):
        out = np.empty(self.shape, dtype=self.dtype)
        call('fh_memcpy_dtoh', out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes))
        return out

    def copy(self):
        out = DeviceArray(self.shape, self.dtype)
        call('fh_memcpy_dtod', out.ptr, self.ptr, ctypes.c_size_t(self.nbytes))
        return out

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        call('fh_memcpy_dtod', self.ptr, other.ptr, ctypes.c_size_t(self.nbytes))
        return self

    @classmethod
    def borrow(cls, ptr, shape, dtype=np.float64):
        """A non-owning view of device memory the library owns (e.g. the iterate handed to an
        iterate callback): valid only while the library keeps it."""
        v = object.__new__(cls)
        v.shape = tuple(int(s) for s in shape)
        v.dtype = np.dtype(dtype)
        v.nbytes = int(np.prod(v.shape, dtype=np.int64)) * v.dtype.itemsize
        v._ptr = ctypes.c_void_p(ptr)
        v._owner = True  # not ours to free
        return v

    def reshape(self, shape):
        """A view with another shape (same memory, keeps the owner alive)."""
        v = object.__new__(DeviceArray)
        v.shape = tuple(int(s) for s in shape)
        assert int(np.prod(v.shape)) == self.size
        v.dtype = self.dtype
        v.nbytes = self.nbytes
        v._ptr = self._ptr
        v._owner = self
        return v

    def __del__(self):
        if getattr(self, '_owner', None) is not None:
            return
        p = getattr(self, '_ptr', None)
        if p is not None and p.value and _LIB is not None:
            try:
                _LIB.fh_free(p)
            except Exception:
                pass
            self._ptr = None

    def __repr__(self):
        return 'DeviceArray(shape=%s, dtype=%s)' % (self.shape, self.dtype)


def synchronize():
    call('fh_synchronize')


def _ptr(x):
    return x.ptr if isinstance(x, DeviceArray) else ctypes.c_void_p(int(x))


def axpby(a, x, b, y):
    """y <- a x + b y on device vectors of equal size and dtype."""
    call('fh_axpby', dtype_code(x.dtype), ctypes.c_size_t(x.size), ctypes.c_double(a), x.ptr,
         ctypes.c_double(b), y.ptr)
    return y


def dot(x, y):
    """sum(x*y) over two device vectors, accumulated in float64."""
    out = ctypes.c_double()
    call('fh_dot', dtype_code(x.dtype), ctypes.c_size_t(x.size), x.ptr, y.ptr, ctypes.byref(out))
    return out.value


def add_constant(y, E):
    """y[c] += E[c] on a component-major device field y of shape (ncomp,)+N."""
    E = np.ascontiguousarray(E, dtype=np.float64)
    nvox = y.size // E.size
    call('fh_add_constant', dtype_code(y.dtype), ctypes.c_size_t(nvox), ctypes.c_int(E.size),
         E.ctypes.data_as(_c_double_p), y.ptr)
    return y


_MODES = {'fp64': FH_SOLVE_FP64, 'mixed': FH_SOLVE_MIXED}


def solve_mode(mode):
    """CG arithmetic code of a precision string: 'fp64' or 'mixed' (anything else raises)."""
    try:
        return _MODES[mode]
    except (KeyError, TypeError):
        raise FhError("precision must be 'fp64' or 'mixed', got %r" % (mode,), FH_ERR_ARG)


def check_array(arr, size, dtypes, what):
    """Raise FhError(FH_ERR_ARG) on the host unless `arr` is a DeviceArray of `size` elements
    whose dtype is one of `dtypes`: every pointer handed to the library covers what the kernels
    touch."""
    if not isinstance(arr, DeviceArray):
        raise FhError('%s must be a DeviceArray, got %s' % (what, type(arr).__name__), FH_ERR_ARG)
    dts = tuple(np.dtype(d) for d in dtypes)
    if arr.dtype not in dts:
        raise FhError('%s has dtype %s, expected %s' % (what, arr.dtype, ' or '.join(str(d) for d in dts)),
                      FH_ERR_ARG)
    if arr.size != int(size):
        raise FhError('%s has %d elements (shape %s), expected %d' % (what, arr.size, arr.shape, int(size)),
                      FH_ERR_ARG)
    return arr


_REAL = (np.float64, np.float32)
_COMPLEX_OF = {np.dtype(np.float64): np.complex128, np.dtype(np.float32): np.complex64}


# ---- spectrum forms and grid changes (ffthompy/tensors/objects.py:135-186, 247-251, 406-511) ----

FORM_CODES = {'r': FH_FORM_R, 0: FH_FORM_FULL, 'c': FH_FORM_CENTRED}
_REAL_OF = {np.dtype(np.complex128): np.float64, np.dtype(np.complex64): np.float32}


def form_code(fft_form):
    """FH_FORM_* of the reference's fft_form 'r', 0 or 'c'."""
    if isinstance(fft_form, str) and fft_form in ('r', 'c'):
        return FORM_CODES[fft_form]
    if isinstance(fft_form, (int, np.integer)) and not isinstance(fft_form, bool) and fft_form == 0:
        return FH_FORM_FULL
    raise FhError("fft_form must be 'r', 0 or 'c', got %r" % (fft_form,), FH_ERR_ARG)


def _grid(N):
    N = [int(n) for n in np.atleast_1d(N)]
    if not 1 <= len(N) <= 3 or min(N) < 1:
        raise FhError('grid must have 1 to 3 positive lengths, got %s' % (N,), FH_ERR_ARG)
    return N, (ctypes.c_int * len(N))(*N)


def _half(N):
    return int(np.prod(N[:-1], dtype=np.int64)) * (N[-1] // 2 + 1)


def _cplx_code(arr, what):
    if arr.dtype not in _REAL_OF:
        raise FhError('%s must be complex128 or complex64, got %s' % (what, arr.dtype), FH_ERR_ARG)
    return dtype_code(_REAL_OF[arr.dtype])


def spectrum_half_to_full(half, N, count, scale=1., centred=False):
    """Full spectrum (form 0, or 'c' when centred) of `count` 'r' half spectra, times scale."""
    N, cN = _grid(N)
    code = _cplx_code(half, 'half spectrum')
    check_array(half, count * _half(N), (half.dtype,), 'half spectrum')
    out = DeviceArray((int(count) * int(np.prod(N)),), half.dtype)
    call('fh_spectrum_half_to_full', code, len(N), cN, ctypes.c_longlong(count), half.ptr, out.ptr,
         ctypes.c_double(scale), int(bool(centred)))
    return out


def spectrum_full_to_half(full, N, count, scale=1., centred=False, hermitian=False):
    """'r' half spectra (times scale) of `count` full spectra (form 0, or 'c' when centred);
    hermitian takes the Hermitian part first."""
    N, cN = _grid(N)
    code = _cplx_code(full, 'full spectrum')
    check_array(full, count * int(np.prod(N)), (full.dtype,), 'full spectrum')
    out = DeviceArray((int(count) * _half(N),), full.dtype)
    call('fh_spectrum_full_to_half', code, len(N), cN, ctypes.c_longlong(count), full.ptr, out.ptr,
         ctypes.c_double(scale), int(bool(centred)), int(bool(hermitian)))
    return out


def fftshift(x, N, count, inverse=False):
    """numpy.fft.fftshift (or ifftshift) over the grid axes of `count` fields of grid N."""
    N, cN = _grid(N)
    if x.dtype.itemsize not in (4, 8, 16):
        raise FhError('unsupported dtype %s' % x.dtype, FH_ERR_ARG)
    check_array(x, count * int(np.prod(N)), (x.dtype,), 'field')
    out = DeviceArray(x.shape, x.dtype)
    call('fh_fftshift', x.dtype.itemsize, len(N), cN, ctypes.c_longlong(count), x.ptr, out.ptr, int(bool(inverse)))
    return out


def spectrum_resize(x, N, M, count, enlarge):
    """Centred spectra of grid N zero-padded (enlarge) or cropped to grid M."""
    N, cN = _grid(N)
    M, cM = _grid(M)
    if len(M) != len(N):
        raise FhError('grids N and M differ in dimension', FH_ERR_ARG)
    for n, m in zip(N, M):
        if (enlarge and m < n + (1 - n % 2)) or (not enlarge and m > n):
            raise FhError('%s from grid %s to %s: every axis needs %s' % (
                'enlarging' if enlarge else 'decreasing', tuple(N), tuple(M),
                'M >= N (odd N) or M >= N + 1 (even N)' if enlarge else 'M <= N'), FH_ERR_ARG)
    code = _cplx_code(x, 'spectrum')
    check_array(x, count * int(np.prod(N)), (x.dtype,), 'spectrum')
    out = DeviceArray((int(count) * int(np.prod(M)),), x.dtype)
    call('fh_spectrum_resize', code, len(N), cN, cM, ctypes.c_longlong(count), x.ptr, out.ptr, int(bool(enlarge)))
    return out


def project_spectrum(x, N, Y, physics, which, fft_form='r', null_nyq=True):
    """G1 (which 1) or G2 (which 2) of the scalar or elasticity problem on component spectra of
    any fft_form, NyqNul as null_nyq (fh_project_spectrum)."""
    N, cN = _grid(N)
    d = len(N)
    ph = {'scalar': FH_SCALAR, 'elasticity': FH_ELASTICITY}.get(physics)
    if ph is None or d not in (2, 3):
        raise FhError('physics must be scalar or elasticity on a 2-D or 3-D grid', FH_ERR_ARG)
    form = form_code(fft_form)
    nc = d if ph == FH_SCALAR else d * (d + 1) // 2
    nspec = _half(N) if form == FH_FORM_R else int(np.prod(N))
    code = _cplx_code(x, 'spectrum')
    check_array(x, nc * nspec, (x.dtype,), 'spectrum')
    Yd = np.ascontiguousarray(np.atleast_1d(Y), dtype=np.float64)
    if Yd.size != d:
        raise FhError('Y must have %d entries' % d, FH_ERR_ARG)
    out = DeviceArray(x.shape, x.dtype)
    call('fh_project_spectrum', code, d, cN, Yd.ctypes.data_as(_c_double_p), ph, int(which), form,
         int(bool(null_nyq)), x.ptr, out.ptr)
    return out


def spectrum_grad(x, N, Y):
    """Gradient (d,) + half grid of a scalar 'r' spectrum of grid N (fh_spectrum_grad)."""
    N, cN = _grid(N)
    code = _cplx_code(x, 'spectrum')
    check_array(x, _half(N), (x.dtype,), 'spectrum')
    Yd = np.ascontiguousarray(np.atleast_1d(Y), dtype=np.float64)
    if Yd.size != len(N):
        raise FhError('Y must have %d entries' % len(N), FH_ERR_ARG)
    out = DeviceArray((len(N),) + tuple(N[:-1]) + (N[-1] // 2 + 1,), x.dtype)
    call('fh_spectrum_grad', code, len(N), cN, Yd.ctypes.data_as(_c_double_p), x.ptr, out.ptr)
    return out


_ELEM = {np.dtype(np.float64): (FH_F64, 0), np.dtype(np.float32): (FH_F32, 0),
         np.dtype(np.complex128): (FH_F64, 1), np.dtype(np.complex64): (FH_F32, 1)}


def pointwise_product(X, Y, A, B, nvox, mode=0):
    """out[a] = sum_b X[a, b] Y[b] (mode 0) or X[a] Y[a] (mode 1) at every grid point, on the
    device (fh_pointwise_product). X, Y: DeviceArrays of one dtype, component-major."""
    if not isinstance(X, DeviceArray) or not isinstance(Y, DeviceArray) or X.dtype != Y.dtype:
        raise FhError('pointwise_product needs two DeviceArrays of one dtype', FH_ERR_ARG)
    if X.dtype not in _ELEM:
        raise FhError('unsupported dtype %s' % X.dtype, FH_ERR_ARG)
    code, cplx = _ELEM[X.dtype]
    A, B, nvox = int(A), int(B), int(nvox)
    check_array(X, A * (B if mode == 0 else 1) * nvox, (X.dtype,), 'X')
    check_array(Y, (B if mode == 0 else A) * nvox, (X.dtype,), 'Y')
    out = DeviceArray((A * nvox,), X.dtype)
    call('fh_pointwise_product', code, cplx, int(mode), A, B, ctypes.c_longlong(nvox), X.ptr, Y.ptr, out.ptr)
    return out


def pointwise_inverse(A):
    """Per-voxel inverse of a (k, k) + grid device field (1 <= k <= 6)."""
    if not isinstance(A, DeviceArray) or len(A.shape) < 3 or A.shape[0] != A.shape[1]:
        raise FhError('pointwise_inverse needs a (k, k) + grid DeviceArray', FH_ERR_ARG)
    k = A.shape[0]
    if k > 6:
        raise FhError('pointwise_inverse handles k <= 6, got %d' % k, FH_ERR_ARG)
    dtype_code(A.dtype)
    nvox = A.size // (k * k)
    out = DeviceArray(A.shape, A.dtype)
    call('fh_pointwise_inverse', dtype_code(A.dtype), k, ctypes.c_longlong(nvox), A.ptr, out.ptr)
    return out


def energy_matrix_noalloc(kind, coef, fields):
    """AH[i, j] = <A e_i, e_j> without a plan or any device allocation (fh_energy_matrix_noalloc):
    kind 'iso' / 'lame' / 'full', coef a fp64 DeviceArray of that layout, fields nf fp64 DeviceArrays
    of nf components each (nf = 2, 3 or 6)."""
    nf = len(fields)
    if nf not in (2, 3, 6):
        raise FhError('nf must be 2, 3 or 6', FH_ERR_ARG)
    nvox = fields[0].size // nf
    per = {'iso': 1, 'lame': 2, 'full': nf * nf}[kind]
    check_array(coef, per * nvox, (np.float64,), 'coef')
    for i, f in enumerate(fields):
        check_array(f, nf * nvox, (np.float64,), 'fields[%d]' % i)
    arr = (ctypes.c_void_p * nf)(*[f.ptr.value for f in fields])
    AH = np.zeros((nf, nf))
    call('fh_energy_matrix_noalloc', ctypes.c_longlong(nvox), ctypes.c_int(_COEF_CODES[kind]), coef.ptr, arr,
         ctypes.c_int(nf), AH.ctypes.data_as(_c_double_p))
    return AH


class Plan:
    """Owner of an fh_plan: grid, cell, physics, projection and coefficient."""

    def __init__(self, N, Y, physics='scalar', primaldual='primal'):
        init()
        N = [int(n) for n in np.atleast_1d(N)]
        Y = [float(y) for y in np.atleast_1d(Y)]
        self.N = tuple(N)
        self.Y = tuple(Y)
        self.dim = len(N)
        self.physics = physics
        self.primaldual = primaldual
        Na = (ctypes.c_int * self.dim)(*N)
        Ya = (ctypes.c_double * self.dim)(*Y)
        h = ctypes.c_void_p()
        call('fh_plan_create', ctypes.c_int(self.dim), Na, Ya,
             ctypes.c_int(FH_SCALAR if physics == 'scalar' else FH_ELASTICITY),
             ctypes.c_int(FH_PRIMAL if primaldual == 'primal' else FH_DUAL), ctypes.byref(h))
        self._h = h
        self.ncomp = load_library().fh_plan_ncomp(h)
        self.nvox = int(np.prod(N))
        self.Nhalf = tuple(N[:-1]) + (N[-1] // 2 + 1,)
        self.coef_kind = None

    @property
    def handle(self):
        return self._h

    def _field(self, arr, what, dtypes=_REAL, ncomp=None):
        """Check a real field of ncomp (default: the plan's) components on the plan grid."""
        return check_array(arr, (self.ncomp if ncomp is None else ncomp) * self.nvox, dtypes, what)

    def _spectrum(self, arr, what, dtypes, ncomp=None):
        """Check an 'r'-form spectrum of ncomp (default: the plan's) components."""
        nh = int(np.prod(self.Nhalf))
        return check_array(arr, (self.ncomp if ncomp is None else ncomp) * nh, dtypes, what)

    def set_primaldual(self, primaldual):
        call('fh_plan_set_primaldual', self._h, ctypes.c_int(FH_PRIMAL if primaldual == 'primal' else FH_DUAL))
        self.primaldual = primaldual

    def set_cg_variant(self, fused=True, defer=True):
        """Mixed-precision CG: fused iteration (default) or the operator-form denominator; with
        defer=False the fused iteration keeps its one-workgroup alpha/beta kernels on small grids
        too (fh_plan_set_cg_variant 1 or 2)."""
        call('fh_plan_set_cg_variant', self._h, ctypes.c_int((1 if defer else 2) if fused else 0))

    def set_band(self, band, scale=1.0):
        """Projection restricted to |k_a| <= band[a] (None: unrestricted); `scale` multiplies
        the operator G A (see fh_plan_set_band)."""
        B = None if band is None else (ctypes.c_int * self.dim)(*[int(b) for b in band])
        call('fh_plan_set_band', self._h, B, ctypes.c_double(scale))

    def use_graphs(self, flag=True):
        """Mixed CG: replay batches of fused iterations from a captured HIP graph (default)."""
        call('fh_plan_use_graphs', self._h, ctypes.c_int(1 if flag else 0))

    def last_graph_batches(self):
        n = ctypes.c_int()
        call('fh_last_solve_graph_batches', self._h, ctypes.byref(n))
        return n.value

    def set_residual_log(self, every):
        """fp64 CG: log ||b - A x|| every `every` iterations (0: off), fh_plan_set_residual_log."""
        call('fh_plan_set_residual_log', self._h, ctypes.c_int(int(every)))

    def set_iterate_callback(self, fn, every=1):
        """fn(kit, x) after every `every`-th iteration of fh_cg_solve (fp64) / fh_iter_solve and after
        the last one, x a borrowed DeviceArray view of the current iterate ((ncomp,)+N, fp64); fn=None
        clears it. An exception in fn stops the solve and is re-raised by the solve's caller through
        pop_callback_error()."""
        if fn is None or every <= 0:
            call('fh_plan_set_iterate_callback', self._h, None, None, ctypes.c_int(0))
            self._cfn = None
            return
        shape = (self.ncomp,) + self.N
        self._cb_error = None

        def tramp(user, kit, xptr):
            try:
                fn(int(kit), DeviceArray.borrow(xptr, shape))
                return 0
            except BaseException as e:  # handed back to Python after the solve returns
                self._cb_error = e
                return 1

        self._cfn = ITERATE_FN(tramp)
        call('fh_plan_set_iterate_callback', self._h, ctypes.cast(self._cfn, ctypes.c_void_p), None,
             ctypes.c_int(int(every)))

    def pop_callback_error(self):
        e, self._cb_error = getattr(self, '_cb_error', None), None
        return e

    def residual_log(self):
        """(iterations, norms) of the last solve's residual log."""
        cnt = ctypes.c_int()
        call('fh_last_residual_log', self._h, None, None, ctypes.c_int(0), ctypes.byref(cnt))
        k = np.zeros(max(cnt.value, 1), dtype=np.int32)
        v = np.zeros(max(cnt.value, 1))
        call('fh_last_residual_log', self._h, k.ctypes.data_as(_c_int_p), v.ctypes.data_as(_c_double_p),
             ctypes.c_int(cnt.value), ctypes.byref(cnt))
        return k[:cnt.value].copy(), v[:cnt.value].copy()

    def use_generic(self, flag=True):
        call('fh_plan_use_generic', self._h, ctypes.c_int(1 if flag else 0))

    def set_coeff(self, kind, vals):
        """vals: host array, or a DeviceArray (fp64) of the same layout."""
        code = {'iso': FH_COEF_ISO, 'lame': FH_COEF_LAME, 'full': FH_COEF_FULL}[kind]
        if isinstance(vals, DeviceArray):
            per = {'iso': 1, 'lame': 2, 'full': self.ncomp ** 2}[kind]
            if vals.size != per * self.nvox or np.dtype(vals.dtype) != np.float64:
                raise FhError('device coefficient needs %d fp64 values' % (per * self.nvox), FH_ERR_ARG)
            call('fh_set_coeff_device', self._h, ctypes.c_int(code), vals.ptr)
        else:
            vals = np.ascontiguousarray(vals, dtype=np.float64)
            call('fh_set_coeff', self._h, ctypes.c_int(code), vals.ctypes.data_as(_c_double_p))
        self.coef_kind = kind

    def set_coeff_inclusions(self, kinds, params, positions, kind, vals, invert=False):
        """Nodal coefficient of an inclusion material evaluated on the device (see
        Ga.set_inclusions for the arguments)."""
        n = len(kinds)
        codes = np.ascontiguousarray([GA_INCLUSION_CODES[k] for k in kinds], dtype=np.int32)
        P = np.zeros((n, 3))
        X = np.zeros((n, 3))
        P[:, :self.dim] = np.asarray(params, dtype=np.float64).reshape(n, self.dim)
        X[:, :self.dim] = np.asarray(positions, dtype=np.float64).reshape(n, self.dim)
        ncoef = {'iso': 1, 'lame': 2, 'full': self.ncomp ** 2}[kind]
        V = np.ascontiguousarray(np.asarray(vals, dtype=np.float64).reshape(n, ncoef))
        call('fh_set_coeff_inclusions', self._h, ctypes.c_int(n), codes.ctypes.data_as(_c_int_p),
             np.ascontiguousarray(P).ctypes.data_as(_c_double_p), np.ascontiguousarray(X).ctypes.data_as(_c_double_p),
             ctypes.c_int(_COEF_CODES[kind]), V.ctypes.data_as(_c_double_p), ctypes.c_int(1 if invert else 0))
        self.coef_kind = kind

    def coeff_to_host(self):
        ncoef = {'iso': 1, 'lame': 2, 'full': self.ncomp ** 2}[self.coef_kind]
        out = np.zeros((ncoef,) + self.N)
        call('fh_coeff_to_host', self._h, out.ctypes.data_as(_c_double_p))
        return out

    def grf(self, white, corr_len, std, exponentiate=False, out=None):
        """Gaussian random field on this plan's grid from device white noise (fh_grf)."""
        self._field(white, 'white noise', (np.float64,), ncomp=1)
        if out is None:
            out = DeviceArray(self.N)
        self._field(out, 'out', (np.float64,), ncomp=1)
        call('fh_grf', self._h, white.ptr, ctypes.c_double(corr_len), ctypes.c_double(std),
             ctypes.c_int(1 if exponentiate else 0), out.ptr)
        return out

    def _pair(self, x, y):
        self._field(x, 'x')
        self._field(y, 'y', (x.dtype,))

    def matvec(self, x, y):
        self._pair(x, y)
        call('fh_matvec', self._h, dtype_code(x.dtype), x.ptr, y.ptr)
        return y

    def project(self, x, y):
        self._pair(x, y)
        call('fh_project', self._h, dtype_code(x.dtype), x.ptr, y.ptr)
        return y

    def rfftn(self, x, xhat, ncomp):
        self._field(x, 'x', ncomp=ncomp)
        self._spectrum(xhat, 'xhat', (_COMPLEX_OF[x.dtype],), ncomp=ncomp)
        call('fh_rfftn', self._h, dtype_code(x.dtype), ctypes.c_int(ncomp), x.ptr, xhat.ptr)
        return xhat

    def irfftn(self, xhat, x, ncomp):
        self._field(x, 'x', ncomp=ncomp)
        self._spectrum(xhat, 'xhat', (_COMPLEX_OF[x.dtype],), ncomp=ncomp)
        call('fh_irfftn', self._h, dtype_code(x.dtype), ctypes.c_int(ncomp), xhat.ptr, x.ptr)
        return x

    def apply_projection_hat(self, xin, xout):
        self._spectrum(xin, 'in_hat', (np.complex128, np.complex64))
        self._spectrum(xout, 'out_hat', (xin.dtype,))
        real = np.float64 if xin.dtype == np.complex128 else np.float32
        call('fh_apply_projection_hat', self._h, dtype_code(real), xin.ptr, xout.ptr)
        return xout

    def cg_solve(self, b, x, tol, maxiter, mode='fp64', delta=1e-4, hist_len=0):
        code = solve_mode(mode)
        self._field(b, 'b', (np.float64,))
        self._field(x, 'x', (np.float64,))
        kit = ctypes.c_int()
        nr = ctypes.c_double()
        st = ctypes.c_int()
        hist = np.zeros(max(hist_len, 1))
        call('fh_cg_solve', self._h, ctypes.c_int(code),
             b.ptr, x.ptr, ctypes.c_double(tol), ctypes.c_int(int(maxiter)), ctypes.c_double(delta),
             ctypes.byref(kit), ctypes.byref(nr), hist.ctypes.data_as(_c_double_p), ctypes.c_int(hist_len),
             ctypes.byref(st))
        k = kit.value
        return {'kit': k, 'norm_res': nr.value, 'status': st.value,
                'res_log': hist[:min(hist_len, k + 1)].copy() if hist_len > 0 else None}

    def iter_solve(self, method, b, x, coefs, tol, maxiter, hist_len=0):
        """Richardson (coefs = [alpha]) or Chebyshev (coefs = eigenvalue range) on the device."""
        code = {'richardson': FH_ITER_RICHARDSON, 'chebyshev': FH_ITER_CHEBYSHEV}[method]
        self._field(b, 'b', (np.float64,))
        self._field(x, 'x', (np.float64,))
        c = np.ascontiguousarray(np.asarray(coefs, dtype=np.float64).reshape(-1))
        kit = ctypes.c_int()
        nres = ctypes.c_double()
        hist = np.zeros(max(hist_len, 1))
        call('fh_iter_solve', self._h, ctypes.c_int(code), b.ptr, x.ptr, c.ctypes.data_as(_c_double_p),
             ctypes.c_double(tol), ctypes.c_int(int(maxiter)), ctypes.byref(kit), ctypes.byref(nres),
             hist.ctypes.data_as(_c_double_p) if hist_len > 0 else None, ctypes.c_int(hist_len))
        out = {'kit': kit.value, 'norm_res': nres.value}
        if hist_len > 0:
            out['res_log'] = hist[:min(hist_len, kit.value + 1)].copy()
        return out

    def potential_solve(self, E, tol, maxiter, Fu=None, X=None, hist_len=0, mode='fp64', delta=1e-4):
        """CG of the potential formulation (fh_potential_solve), mode 'fp64' or 'mixed': Fu, the
        potential's rfftn spectrum (complex, N[:-1] + (N[-1]//2+1,)), and X = F^-1 grad Fu
        ((ncomp,) + N, optional)."""
        code = solve_mode(mode)
        E = np.ascontiguousarray(np.asarray(E, dtype=np.float64).reshape(-1))
        if E.size != self.ncomp:
            raise FhError('E needs %d values' % self.ncomp, FH_ERR_ARG)
        if Fu is None:
            Fu = DeviceArray(self.N[:-1] + (self.N[-1] // 2 + 1,), np.complex128)
        self._spectrum(Fu, 'Fu', (np.complex128,), ncomp=1)
        if X is not None:
            self._field(X, 'X', (np.float64,))
        kit, nres, status = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        hist = np.zeros(max(hist_len, 1))
        call('fh_potential_solve', self._h, ctypes.c_int(code),
             E.ctypes.data_as(_c_double_p), ctypes.c_double(tol), ctypes.c_int(int(maxiter)), ctypes.c_double(delta),
             Fu.ptr, X.ptr if X is not None else None, ctypes.byref(kit),
             ctypes.byref(nres), hist.ctypes.data_as(_c_double_p) if hist_len > 0 else None,
             ctypes.c_int(hist_len), ctypes.byref(status))
        out = {'kit': kit.value, 'norm_res': nres.value, 'status': status.value}
        if hist_len > 0:
            out['res_log'] = hist[:min(hist_len, kit.value + 1)].copy()
        return Fu, out

    def _fields(self, fields, what='fields'):
        if len(fields) != self.ncomp:
            raise FhError('need %d %s (one per unit load), got %d' % (self.ncomp, what, len(fields)), FH_ERR_ARG)
        for i, f in enumerate(fields):
            self._field(f, '%s[%d]' % (what, i), (np.float64,))

    def energy_matrix(self, fields):
        self._fields(fields)
        nf = len(fields)
        arr = (ctypes.c_void_p * nf)(*[f.ptr.value for f in fields])
        AH = np.zeros((nf, nf))
        call('fh_energy_matrix', self._h, arr, ctypes.c_int(nf), AH.ctypes.data_as(_c_double_p))
        return AH

    def energy_gradient(self, fields, W, chain=None, out=None):
        """Device gradient of sum_ij W_ij AH_ij w.r.t. the coefficient (see fh_energy_gradient)."""
        self._fields(fields)
        nf = len(fields)
        arr = (ctypes.c_void_p * nf)(*[f.ptr.value for f in fields])
        W = np.ascontiguousarray(W, dtype=np.float64)
        if W.size != nf * nf:
            raise FhError('W needs %d values' % (nf * nf), FH_ERR_ARG)
        nvals = 1 if self.coef_kind == 'iso' else 2
        if chain is not None:
            self._field(chain, 'chain', (np.float64,), ncomp=1)
        if out is None:
            out = DeviceArray((nvals,) + self.N if nvals > 1 else self.N, np.float64)
        self._field(out, 'out', (np.float64,), ncomp=nvals)
        call('fh_energy_gradient', self._h, arr, ctypes.c_int(nf), W.ctypes.data_as(_c_double_p),
             chain.ptr if chain is not None else None, out.ptr)
        return out

    def field_gradient(self, fields, adjoints, chain=None, out=None):
        """Device gradient -chain * sum_i adjoints_i . fields_i (see fh_field_gradient)."""
        if len(adjoints) != len(fields):
            raise FhError('one adjoint per field', FH_ERR_ARG)
        for i, (f, a) in enumerate(zip(fields, adjoints)):
            self._field(f, 'fields[%d]' % i, (np.float64,))
            self._field(a, 'adjoints[%d]' % i, (np.float64,))
        if chain is not None:
            self._field(chain, 'chain', (np.float64,), ncomp=1)
        nf = len(fields)
        fa = (ctypes.c_void_p * nf)(*[f.ptr.value for f in fields])
        aa = (ctypes.c_void_p * nf)(*[f.ptr.value for f in adjoints])
        if out is None:
            out = DeviceArray(self.N, np.float64)
        self._field(out, 'out', (np.float64,), ncomp=1)
        call('fh_field_gradient', self._h, fa, aa, ctypes.c_int(nf), chain.ptr if chain is not None else None,
             out.ptr)
        return out

    def last_timing(self):
        mv = ctypes.c_double()
        sv = ctypes.c_double()
        call('fh_last_timing', self._h, ctypes.byref(mv), ctypes.byref(sv))
        return mv.value, sv.value

    def last_updates(self):
        n = ctypes.c_int()
        call('fh_last_solve_updates', self._h, ctypes.byref(n))
        return n.value

    def matvec_profile(self, x, y, nrep=10):
        """Average ms and algorithmic bytes per kernel of one matvec (HIP events)."""
        self._pair(x, y)
        ms = np.zeros(8)
        by = np.zeros(8)
        nk = ctypes.c_int()
        call('fh_matvec_profile', self._h, dtype_code(x.dtype), x.ptr, y.ptr, ctypes.c_int(nrep), ctypes.c_int(8),
             ms.ctypes.data_as(_c_double_p), by.ctypes.data_as(_c_double_p), ctypes.byref(nk))
        return ms[:nk.value].copy(), by[:nk.value].copy()

    def set_schedule(self, dtype, proj_variant, cols_variant, rows_variant=0):
        """Schedule variants of the fixed-length passes for one arithmetic (fh_plan_set_schedule)."""
        call('fh_plan_set_schedule', self._h, ctypes.c_int(dtype_code(dtype)), ctypes.c_int(int(proj_variant)),
             ctypes.c_int(int(cols_variant)), ctypes.c_int(int(rows_variant)))

    def schedule(self, dtype):
        p, c, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        call('fh_plan_get_schedule', self._h, ctypes.c_int(dtype_code(dtype)), ctypes.byref(p), ctypes.byref(c),
             ctypes.byref(r))
        return p.value, c.value, r.value

    def autotune(self, dtype=np.float32, nrep=5, seed=0):
        """Time every schedule variant of the projection, strided and row passes on this plan (random
        field, 3 interleaved rounds of nrep matvecs per variant, best round; in single precision the
        row counts on the fused CG iteration's row passes) and keep the fastest: fh_plan_autotune.
        Returns
        {'proj_ms': [...], 'cols_ms': [...], 'rows_ms': [...], 'proj_variant': p, 'cols_variant': c,
        'rows_variant': r}."""
        dt = np.dtype(dtype)
        x = DeviceArray.from_host(np.random.default_rng(seed).standard_normal((self.ncomp,) + self.N).astype(dt))
        y = DeviceArray((self.ncomp,) + self.N, dt)
        tp, tc, tr = np.zeros(4), np.zeros(2), np.zeros(3)
        p, c, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        call('fh_plan_autotune', self._h, ctypes.c_int(dtype_code(dt)), x.ptr, y.ptr, ctypes.c_int(int(nrep)),
             tp.ctypes.data_as(_c_double_p), tc.ctypes.data_as(_c_double_p), tr.ctypes.data_as(_c_double_p),
             ctypes.byref(p), ctypes.byref(c), ctypes.byref(r))
        return {'proj_ms': tp.tolist(), 'cols_ms': tc.tolist(), 'rows_ms': tr.tolist(), 'proj_variant': p.value,
                'cols_variant': c.value, 'rows_variant': r.value}

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and _LIB is not None:
            try:
                _LIB.fh_plan_destroy(h)
            except Exception:
                pass
            self._h = None


class Batch:
    """Owner of an fh_batch: many 2-D scalar samples on their own grids, solved together."""

    def __init__(self, dims, Y=None):
        init()
        dims = np.ascontiguousarray(np.asarray(dims, dtype=np.int32).reshape(-1, 2))
        self.dims = [tuple(int(v) for v in d) for d in dims]
        self.S = len(self.dims)
        Ya = None
        if Y is not None:
            Ya = np.ascontiguousarray(np.broadcast_to(np.asarray(Y, dtype=np.float64), (self.S, 2)))
            self._Y = Ya
        h = ctypes.c_void_p()
        call('fh_batch_create', ctypes.c_int(self.S), dims.ctypes.data_as(_c_int_p),
             Ya.ctypes.data_as(_c_double_p) if Ya is not None else None, ctypes.byref(h))
        self._h = h
        self.nvox = sum(a * b for a, b in self.dims)

    def set_coeff(self, samples):
        """samples: list of 2-D arrays (shape dims[s]) of positive conductivities, origin 0."""
        flat = np.concatenate([np.asarray(a, dtype=np.float64).reshape(-1) for a in samples])
        if flat.size != self.nvox:
            raise FhError('expected %d voxel values, got %d' % (self.nvox, flat.size), FH_ERR_ARG)
        flat = np.ascontiguousarray(flat)
        call('fh_batch_set_coeff', self._h, flat.ctypes.data_as(_c_double_p))

    def solve(self, tol=1e-8, maxiter=10000, mode='mixed', delta=1e-4):
        up = np.zeros((self.S, 2, 2))
        lo = np.zeros((self.S, 2, 2))
        kit = np.zeros((self.S, 4), dtype=np.int32)
        st = np.zeros(self.S, dtype=np.int32)
        call('fh_batch_solve', self._h, ctypes.c_int(solve_mode(mode)),
             ctypes.c_double(tol), ctypes.c_int(int(maxiter)), ctypes.c_double(delta),
             up.ctypes.data_as(_c_double_p), lo.ctypes.data_as(_c_double_p), kit.ctypes.data_as(_c_int_p),
             st.ctypes.data_as(_c_int_p))
        return up, lo, kit, st

    def set_schedule(self, variant, order_mode=0):
        call('fh_batch_set_schedule', self._h, ctypes.c_int(variant), ctypes.c_int(order_mode))

    def autotune(self, mode='mixed', delta=1e-4, iters=12):
        """Time every kernel schedule on this batch; keep the fastest. Returns (times_ms[order][variant],
        variant, order_mode)."""
        t = np.zeros(4)
        v = ctypes.c_int()
        o = ctypes.c_int()
        call('fh_batch_autotune', self._h, ctypes.c_int(solve_mode(mode)),
             ctypes.c_double(delta), ctypes.c_int(iters), t.ctypes.data_as(_c_double_p), ctypes.byref(v),
             ctypes.byref(o))
        return t.reshape(2, 2), v.value, o.value

    def stats(self):
        nupd = np.zeros((self.S, 4), dtype=np.int32)
        nres = np.zeros((self.S, 4))
        ms = ctypes.c_double()
        call('fh_batch_stats', self._h, nupd.ctypes.data_as(_c_int_p), nres.ctypes.data_as(_c_double_p),
             ctypes.byref(ms))
        return nupd, nres, ms.value

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and _LIB is not None:
            try:
                _LIB.fh_batch_destroy(h)
            except Exception:
                pass
            self._h = None


GA_INCLUSION_CODES = {'cube': 1, 'square': 1, 'ball': 2, 'circle': 2, 'bilinear_pyramid': 3, 'pyramid': 3,
                      'all': 4, 'otherwise': 5}
_COEF_CODES = {'iso': FH_COEF_ISO, 'lame': FH_COEF_LAME, 'full': FH_COEF_FULL}


class Ga:
    """Owner of an fh_ga: exactly integrated coefficient and interpolated solutions on the
    quadrature grid of a plan's solutions (Ga post-processing)."""

    def __init__(self, plan, grid=None):
        init()
        self.plan = plan
        h = ctypes.c_void_p()
        G = None if grid is None else (ctypes.c_int * plan.dim)(*[int(m) for m in grid])
        call('fh_ga_create', plan.handle, G, ctypes.byref(h))
        self._h = h
        M = (ctypes.c_int * plan.dim)()
        call('fh_ga_grid', h, M)
        self.M = tuple(int(m) for m in M)
        self.nvox = int(np.prod(self.M))
        self.coef_kind = None
        self.ncoef = 0

    def _ncoef(self, kind):
        return {'iso': 1, 'lame': 2, 'full': self.plan.ncomp ** 2}[kind]

    def set_inclusions(self, kinds, params, positions, kind, vals):
        """kinds: inclusion names; params, positions: (nincl, dim) (a ball's diameter in
        params[:, 0]); vals: (nincl, ncoef) phase coefficients in layout `kind`."""
        n = len(kinds)
        codes = np.ascontiguousarray([GA_INCLUSION_CODES[k] for k in kinds], dtype=np.int32)
        P = np.zeros((n, 3))
        X = np.zeros((n, 3))
        d = self.plan.dim
        P[:, :d] = np.asarray(params, dtype=np.float64).reshape(n, d)
        X[:, :d] = np.asarray(positions, dtype=np.float64).reshape(n, d)
        V = np.ascontiguousarray(np.asarray(vals, dtype=np.float64).reshape(n, self._ncoef(kind)))
        P = np.ascontiguousarray(P)
        X = np.ascontiguousarray(X)
        call('fh_ga_set_inclusions', self._h, ctypes.c_int(n), codes.ctypes.data_as(_c_int_p),
             P.ctypes.data_as(_c_double_p), X.ctypes.data_as(_c_double_p), ctypes.c_int(_COEF_CODES[kind]),
             V.ctypes.data_as(_c_double_p))
        self.coef_kind, self.ncoef = kind, self._ncoef(kind)

    def set_grid_coeff(self, order, P, kind, vals):
        """vals: (ncoef,) + P nodal values in origin-0 order (dual: pointwise inverse)."""
        P = np.ascontiguousarray(np.asarray(P, dtype=np.int32).reshape(-1))
        V = np.ascontiguousarray(np.asarray(vals, dtype=np.float64))
        if V.size != self._ncoef(kind) * int(np.prod(P)):
            raise FhError('expected %d coefficient values, got %d' % (self._ncoef(kind) * int(np.prod(P)), V.size),
                          FH_ERR_ARG)
        call('fh_ga_set_grid_coeff', self._h, ctypes.c_int(int(order)), P.ctypes.data_as(_c_int_p),
             ctypes.c_int(_COEF_CODES[kind]), V.ctypes.data_as(_c_double_p))
        self.coef_kind, self.ncoef = kind, self._ncoef(kind)

    def coefficient(self):
        out = np.zeros((self.ncoef,) + self.M)
        call('fh_ga_coefficient', self._h, out.ctypes.data_as(_c_double_p))
        return out

    def interpolate(self, x, out=None):
        if out is None:
            out = DeviceArray((self.plan.ncomp,) + self.M)
        call('fh_ga_interpolate', self._h, x.ptr, out.ptr)
        return out

    def restrict(self, x, out=None):
        """Field on the quadrature grid -> the coarse plan's grid (spectrum cropped)."""
        if out is None:
            out = DeviceArray((self.plan.ncomp,) + self.plan.N)
        call('fh_ga_restrict', self._h, x.ptr, out.ptr)
        return out

    def energy_matrix(self, fields, on_grid=False):
        """fields: solutions on the plan's grid (interpolated here), or with on_grid=True
        already on the quadrature grid."""
        nf = len(fields)
        arr = (ctypes.c_void_p * nf)(*[f.ptr for f in fields])
        AH = np.zeros((nf, nf))
        call('fh_ga_energy_on_grid' if on_grid else 'fh_ga_energy', self._h, arr, ctypes.c_int(nf),
             AH.ctypes.data_as(_c_double_p))
        return AH

    def set_plan_coeff(self, plan):
        """Make A_Ga the coefficient of `plan` (a plan on the quadrature grid)."""
        call('fh_ga_set_plan_coeff', self._h, plan.handle)
        plan.coef_kind = self.coef_kind

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and _LIB is not None:
            try:
                _LIB.fh_ga_destroy(h)
            except Exception:
                pass
            self._h = None


def smooth_length(L):
    """Smallest length >= L whose prime factors are all <= 13 (the FFT radices)."""
    L = int(L)
    while True:
        r = L
        for p in (2, 3, 5, 7, 11, 13):
            while r % p == 0:
                r //= p
        if r == 1:
            return L
        L += 1
