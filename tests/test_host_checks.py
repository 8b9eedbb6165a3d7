"""Host-side argument checks of the binding (no GPU): every Plan method that hands device
pointers to the library checks their sizes and dtypes against the plan first and raises
FhError(FH_ERR_ARG) before any library call; precision strings are validated; x0 must match B."""
import ctypes

import numpy as np
import pytest

from ffthompy_amd import _native as nat
from ffthompy_amd import solver


def fake_array(shape, dtype=np.float64):
    """A DeviceArray-shaped object with no device memory (its pointer is never used: the checks
    must raise first)."""
    a = object.__new__(nat.DeviceArray)
    a.shape = tuple(shape)
    a.dtype = np.dtype(dtype)
    a.nbytes = int(np.prod(shape)) * a.dtype.itemsize
    a._ptr = ctypes.c_void_p(0)
    a._owner = True  # nothing to free
    return a


def fake_plan(N, physics='scalar'):
    p = object.__new__(nat.Plan)
    p.N = tuple(N)
    p.dim = len(N)
    p.ncomp = len(N) if physics == 'scalar' else len(N) * (len(N) + 1) // 2
    p.nvox = int(np.prod(N))
    p.Nhalf = tuple(N[:-1]) + (N[-1] // 2 + 1,)
    p.coef_kind = 'iso'
    p._h = None
    return p


@pytest.fixture
def no_library_calls(monkeypatch):
    def refuse(name, *args):
        raise AssertionError('library call %s reached: the host check did not fire' % name)
    monkeypatch.setattr(nat, 'call', refuse)


def _arg_error(fn, *a, **kw):
    with pytest.raises(nat.FhError) as ei:
        fn(*a, **kw)
    assert ei.value.code == nat.FH_ERR_ARG, ei.value


def test_matvec_project_checks(no_library_calls):
    pl = fake_plan((9, 9, 9))
    good = fake_array((3, 9, 9, 9))
    _arg_error(pl.matvec, fake_array((2, 9, 9, 9)), good)            # wrong component count
    _arg_error(pl.matvec, good, fake_array((3, 9, 9, 9), np.float32))  # y dtype differs from x
    _arg_error(pl.project, fake_array((3, 9, 9, 7)), good)           # another grid
    _arg_error(pl.matvec, fake_array((3, 9, 9, 9), np.complex128), good)
    _arg_error(pl.matvec, np.zeros((3, 9, 9, 9)), good)              # host array


def test_fft_and_projection_hat_checks(no_library_calls):
    pl = fake_plan((9, 9, 9))
    x = fake_array((3, 9, 9, 9))
    _arg_error(pl.rfftn, x, fake_array((3, 9, 9, 5), np.complex64), 3)  # precision of the spectrum
    _arg_error(pl.rfftn, x, fake_array((3, 9, 9, 4), np.complex128), 3)  # size of the spectrum
    _arg_error(pl.irfftn, fake_array((2, 9, 9, 5), np.complex128), x, 3)
    _arg_error(pl.apply_projection_hat, fake_array((3, 9, 9, 5), np.complex128),
               fake_array((3, 9, 9, 5), np.complex64))


def test_solver_entry_checks(no_library_calls):
    pl = fake_plan((9, 9, 9))
    b = fake_array((3, 9, 9, 9))
    _arg_error(pl.cg_solve, b, fake_array((3, 9, 9, 9), np.float32), 1e-8, 10)
    _arg_error(pl.cg_solve, fake_array((3, 7, 9, 9)), b, 1e-8, 10)
    _arg_error(pl.cg_solve, b, b, 1e-8, 10, mode='FP64')
    _arg_error(pl.cg_solve, b, b, 1e-8, 10, mode='double')
    _arg_error(pl.iter_solve, 'richardson', fake_array((6, 9, 9, 9)), b, [1.], 1e-6, 10)
    _arg_error(pl.potential_solve, np.ones(3), 1e-8, 10, Fu=fake_array((9, 9, 5), np.complex128),
               X=fake_array((2, 9, 9, 9)))
    _arg_error(pl.potential_solve, np.ones(3), 1e-8, 10, Fu=fake_array((9, 9, 9), np.complex128))
    _arg_error(pl.potential_solve, np.ones(3), 1e-8, 10, Fu=fake_array((9, 9, 5), np.complex128), mode='fp32')


def test_energy_and_gradient_checks(no_library_calls):
    pl = fake_plan((9, 9, 9))
    f = [fake_array((3, 9, 9, 9)) for _ in range(3)]
    _arg_error(pl.energy_matrix, f[:2])                                  # one field per load
    _arg_error(pl.energy_matrix, f[:2] + [fake_array((3, 9, 9, 8))])
    _arg_error(pl.energy_gradient, f, np.eye(2))                         # W of the wrong size
    _arg_error(pl.energy_gradient, f, np.eye(3), chain=fake_array((9, 9, 8)))
    _arg_error(pl.field_gradient, f, f[:2])
    _arg_error(pl.field_gradient, f, f, out=fake_array((9, 9, 9), np.float32))
    _arg_error(pl.grf, fake_array((9, 9, 8)), 0.1, 1.)


def test_precision_strings():
    assert nat.solve_mode('fp64') == nat.FH_SOLVE_FP64
    assert nat.solve_mode('mixed') == nat.FH_SOLVE_MIXED
    for bad in ('FP64', 'double', 'fp32', None, 1):
        with pytest.raises(nat.FhError):
            nat.solve_mode(bad)


def test_batch_rejects_bad_precision_before_device(monkeypatch):
    from ffthompy_amd import batch

    def no_device(*a, **k):
        raise AssertionError('the batch was created before the precision was checked')

    monkeypatch.setattr(nat, 'Batch', no_device)
    for bad in ('FP64', 'double', 'fp32'):
        with pytest.raises(nat.FhError):
            batch.homogenize_batch([np.ones((27, 27))], precision=bad)


def test_cg_rejects_bad_precision_and_x0_before_device(monkeypatch):
    from ffthompy_amd.tensors import Tensor

    class Op:
        plan = None

    monkeypatch.setattr(solver, 'GpuOperator', Op)
    B = Tensor(name='B', val=np.zeros((2, 5, 5)), order=1)
    with pytest.raises(nat.FhError):
        solver.CG(Op(), B, par={'precision': 'double'})
    x0 = Tensor(name='x0', val=np.zeros((2, 5, 7)), order=1)
    monkeypatch.setattr(Tensor, 'to_device', lambda self: self)
    with pytest.raises(ValueError):
        solver.CG(Op(), B, x0=x0, par={'precision': 'fp64'})


def test_form_and_grid_changes_check_on_the_host(no_library_calls):
    """Spectrum-form entry points reject bad forms, grids, sizes and dtypes before any call."""
    for bad in ('x', 1, None, True):
        with pytest.raises(nat.FhError):
            nat.form_code(bad)
    assert nat.form_code(np.int64(0)) == nat.FH_FORM_FULL and nat.form_code('c') == nat.FH_FORM_CENTRED
    half = fake_array((2, 6, 4), np.complex128)  # N = (6, 7): half (6, 4)
    with pytest.raises(nat.FhError):
        nat.spectrum_half_to_full(half, (6, 7), 3)  # wrong count
    with pytest.raises(nat.FhError):
        nat.spectrum_half_to_full(fake_array((2, 6, 4)), (6, 7), 2)  # real, not complex
    with pytest.raises(nat.FhError):
        nat.spectrum_full_to_half(half, (6, 7), 2)  # a half spectrum is not a full one
    with pytest.raises(nat.FhError):
        nat.fftshift(fake_array((2, 5, 5), np.int16), (5, 5), 2)
    full = fake_array((2, 6, 7), np.complex128)
    with pytest.raises(nat.FhError):
        nat.spectrum_resize(full, (6, 7), (6, 9), 2, enlarge=True)  # even axis needs M >= N + 1
    with pytest.raises(nat.FhError):
        nat.spectrum_resize(full, (6, 7), (7, 8), 2, enlarge=False)  # decreasing needs M <= N
    with pytest.raises(nat.FhError):
        nat.spectrum_resize(full, (6, 7), (7, 9, 3), 2, enlarge=True)
    with pytest.raises(nat.FhError):
        nat.project_spectrum(fake_array((3, 6, 4), np.complex128), (6, 7), (1., 1.), 'scalar', 1, 'r')
    with pytest.raises(nat.FhError):
        nat.project_spectrum(fake_array((2, 6, 7), np.complex128), (6, 7), (1., 1.), 'scalar', 1, 'r')
    with pytest.raises(nat.FhError):
        nat.project_spectrum(fake_array((2, 6, 4), np.complex128), (6, 7), (1.,), 'scalar', 1, 'r')
    with pytest.raises(nat.FhError):
        nat.project_spectrum(fake_array((1, 9), np.complex128), (9,), (1.,), 'scalar', 1, 0)
    with pytest.raises(nat.FhError):
        nat.pointwise_inverse(fake_array((7, 7, 4, 4)))
    with pytest.raises(nat.FhError):
        nat.pointwise_inverse(fake_array((3, 3, 4, 4), np.int32))
