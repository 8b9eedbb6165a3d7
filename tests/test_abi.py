"""CPU checks of the C ABI: the library loads, exports every symbol include/ffthom.h
declares, and fails cleanly (error code + message, no crash) without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from ffthompy_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'ffthom.h')


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'\b(fh_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = nat.load_library()
    names = declared_symbols()
    assert 'fh_cg_solve' in names and 'fh_matvec' in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    names = set(declared_symbols()) - {'fh_last_error'}
    assert names == set(nat.SIGNATURES), names ^ set(nat.SIGNATURES)


def _have_gpu():
    try:
        nat.init()
        return True
    except nat.FhError:
        return False


@pytest.mark.skipif(_have_gpu(), reason='checks the no-GPU error path')
def test_no_device_is_a_clean_error():
    with pytest.raises(nat.FhError) as ei:
        nat.init()
    assert ei.value.code in (nat.FH_ERR_NODEV, nat.FH_ERR_UNSUPPORTED, nat.FH_ERR_HIP)
    with pytest.raises(nat.FhError):
        nat.Plan((5, 5, 5), (1., 1., 1.))
    lib = nat.load_library()
    h = ctypes.c_void_p()
    N = (ctypes.c_int * 3)(5, 5, 5)
    Y = (ctypes.c_double * 3)(1., 1., 1.)
    assert lib.fh_plan_create(3, N, Y, 0, 0, ctypes.byref(h)) != 0
    assert len(lib.fh_last_error()) > 0


def test_argument_validation_without_device():
    lib = nat.load_library()
    h = ctypes.c_void_p()
    N = (ctypes.c_int * 3)(5, 5, 5)
    Y = (ctypes.c_double * 3)(1., 1., 1.)
    assert lib.fh_plan_create(4, N, Y, 0, 0, ctypes.byref(h)) == nat.FH_ERR_ARG
    assert lib.fh_plan_create(3, N, Y, 7, 0, ctypes.byref(h)) == nat.FH_ERR_ARG
    Yb = (ctypes.c_double * 3)(1., -1., 1.)
    assert lib.fh_plan_create(3, N, Yb, 0, 0, ctypes.byref(h)) == nat.FH_ERR_ARG
    assert lib.fh_matvec(None, 0, None, None) == nat.FH_ERR_ARG
    assert lib.fh_cg_solve(None, 0, None, None, 1e-6, 10, 0.1, None, None, None, 0, None) == nat.FH_ERR_ARG


def test_argument_validation_of_later_entry_points_without_device():
    """Null handles and pointers are rejected with FH_ERR_ARG before any device work."""
    lib = nat.load_library()
    h = ctypes.c_void_p()
    assert lib.fh_ga_create(None, None, ctypes.byref(h)) == nat.FH_ERR_ARG
    assert lib.fh_ga_energy(None, None, 3, None) == nat.FH_ERR_ARG
    assert lib.fh_ga_energy_on_grid(None, None, 3, None) == nat.FH_ERR_ARG
    assert lib.fh_ga_set_inclusions(None, 1, None, None, None, 1, None) == nat.FH_ERR_ARG
    assert lib.fh_ga_set_grid_coeff(None, 0, None, 1, None) == nat.FH_ERR_ARG
    assert lib.fh_ga_interpolate(None, None, None) == nat.FH_ERR_ARG
    assert lib.fh_ga_set_plan_coeff(None, None) == nat.FH_ERR_ARG
    assert lib.fh_iter_solve(None, 1, None, None, None, 1e-6, 10, None, None, None, 0) == nat.FH_ERR_ARG
    assert lib.fh_potential_solve(None, 0, None, 1e-6, 10, 1e-4, None, None, None, None, None, 0, None) == nat.FH_ERR_ARG
    assert lib.fh_set_coeff_inclusions(None, 1, None, None, None, 1, None, 0) == nat.FH_ERR_ARG
    assert lib.fh_set_coeff_device(None, 1, None) == nat.FH_ERR_ARG
    assert lib.fh_grf(None, None, 0.1, 1.0, 0, None) == nat.FH_ERR_ARG
    assert lib.fh_plan_set_band(None, None, 1.0) == nat.FH_ERR_ARG
    assert lib.fh_plan_use_graphs(None, 1) == nat.FH_ERR_ARG
    assert lib.fh_plan_set_residual_log(None, 1) == nat.FH_ERR_ARG
    assert lib.fh_last_residual_log(None, None, None, 0, None) == nat.FH_ERR_ARG
    assert len(lib.fh_last_error()) > 0


def test_missing_library_fails_loudly(monkeypatch):
    """Without the built library every product entry raises FhError: there is no CPU fallback."""
    monkeypatch.setattr(nat, '_LIB', None)
    monkeypatch.setattr(nat, 'LIB_PATH', '/nonexistent/libffthom.so')
    with pytest.raises(nat.FhError):
        nat.load_library()
    from ffthompy_amd import potential
    with pytest.raises(nat.FhError):
        nat.Plan((5, 5), (1., 1.), 'scalar', 'primal')
    with pytest.raises(nat.FhError):
        potential.homog_GaNi_full_potential(
            potential.Tensor(name='A', val=np.einsum('ij,...->ij...', np.eye(2), np.ones((5, 5))), order=2),
            None, {'solver': {'tol': 1e-8, 'maxiter': 10}})
