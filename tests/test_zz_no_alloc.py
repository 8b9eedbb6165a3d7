"""GPU test of fh_energy_matrix_noalloc: the allocation-free energy entry point against
fh_energy_matrix (same kernels, same partial-sum grid: equal to the last bit), through the C ABI,
on two calls in a row (the second on a smaller input) and on the elasticity layouts."""
import ctypes

import numpy as np
import pytest

from ffthompy_amd import _native as nat

pytestmark = pytest.mark.gpu


def _fields(rng, nf, N):
    return [nat.DeviceArray.from_host(rng.standard_normal((nf,) + N)) for _ in range(nf)]


def _noalloc_abi(nvox, kind, coef, fields):
    """The raw C call, as a caller without the Python wrapper makes it."""
    lib = nat.load_library()
    nf = len(fields)
    arr = (ctypes.c_void_p * nf)(*[f.ptr.value for f in fields])
    AH = np.zeros((nf, nf))
    rc = lib.fh_energy_matrix_noalloc(ctypes.c_longlong(nvox), ctypes.c_int(kind), coef.ptr, arr, ctypes.c_int(nf),
                                      AH.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == nat.FH_OK, lib.fh_last_error()
    return AH


def _with_plan(N, physics, kind, coef_host, fields):
    pl = nat.Plan(N, np.ones(len(N)), physics, 'primal')
    pl.set_coeff(kind, coef_host)
    return pl.energy_matrix(fields)


def test_noalloc_energy_equals_plan_energy_two_calls():
    nat.init()
    rng = np.random.default_rng(7)
    # first call: 3-D scalar 45^3, isotropic coefficient
    N = (45, 45, 45)
    a = 1. + 99. * (rng.random(N) > 0.8)
    F = _fields(rng, 3, N)
    ref = _with_plan(N, 'scalar', 'iso', a, F)
    got = _noalloc_abi(int(np.prod(N)), nat.FH_COEF_ISO, nat.DeviceArray.from_host(a), F)
    print('45^3 iso: max |diff| %.3e' % np.abs(got - ref).max(), flush=True)
    assert np.array_equal(got, ref)
    # second call, smaller input: 2-D scalar 27 x 15 (the scratch keeps the first call's contents)
    N2 = (27, 15)
    a2 = np.exp(rng.standard_normal(N2))
    F2 = _fields(rng, 2, N2)
    ref2 = _with_plan(N2, 'scalar', 'iso', a2, F2)
    got2 = _noalloc_abi(int(np.prod(N2)), nat.FH_COEF_ISO, nat.DeviceArray.from_host(a2), F2)
    print('27x15 iso: max |diff| %.3e' % np.abs(got2 - ref2).max(), flush=True)
    assert np.array_equal(got2, ref2)
    # and the first input again: still equal
    got3 = nat.energy_matrix_noalloc('iso', nat.DeviceArray.from_host(a), F)
    assert np.array_equal(got3, ref)


def test_noalloc_energy_elasticity_layouts():
    nat.init()
    rng = np.random.default_rng(8)
    N = (15, 15, 15)
    n = int(np.prod(N))
    lame = np.stack([1. + rng.random(N), 0.5 + rng.random(N)])
    F = _fields(rng, 6, N)
    ref = _with_plan(N, 'elasticity', 'lame', lame, F)
    got = nat.energy_matrix_noalloc('lame', nat.DeviceArray.from_host(lame), F)
    assert np.array_equal(got, ref)
    full = rng.standard_normal((6, 6) + N)
    ref = _with_plan(N, 'elasticity', 'full', full, F)
    got = _noalloc_abi(n, nat.FH_COEF_FULL, nat.DeviceArray.from_host(full), F)
    assert np.array_equal(got, ref)


def test_noalloc_energy_rejects_bad_sizes():
    nat.init()
    lib = nat.load_library()
    rng = np.random.default_rng(9)
    N = (9, 9, 9)
    F = _fields(rng, 3, N)
    a = nat.DeviceArray.from_host(np.ones(N))
    arr = (ctypes.c_void_p * 7)(*([f.ptr.value for f in F] + [F[0].ptr.value] * 4))
    AH = np.zeros(49)
    dp = AH.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.fh_energy_matrix_noalloc(729, nat.FH_COEF_ISO, a.ptr, arr, 7, dp) == nat.FH_ERR_ARG
    assert b'at most 6' in lib.fh_last_error()
    assert lib.fh_energy_matrix_noalloc(729, 9, a.ptr, arr, 3, dp) == nat.FH_ERR_ARG
    assert lib.fh_energy_matrix_noalloc(0, nat.FH_COEF_ISO, a.ptr, arr, 3, dp) == nat.FH_ERR_ARG
    assert lib.fh_energy_matrix_noalloc(729, nat.FH_COEF_LAME, a.ptr, arr, 2, dp) == nat.FH_ERR_ARG
