"""Coefficients and random fields generated on the device.

* fh_set_coeff_inclusions against the host evaluation of the same material (the reference's
  get_A_GaNi rules, which the golden problems pin): bit-identical nodal values for the
  isotropic and Lame layouts (every example material, odd and even grids, inclusions whose
  boundaries pass through nodes), the dual inverse to rounding, overlapping inclusions rejected.
* fh_grf against the NumPy spectral synthesis of the same white noise (1e-12).
"""
import numpy as np
import pytest

from oracle import examples

pytestmark = pytest.mark.gpu
ffa = pytest.importorskip('ffthompy_amd')
from ffthompy_amd import _native as nat  # noqa: E402
from ffthompy_amd import microstructure as ms  # noqa: E402
from ffthompy_amd.materials import Material  # noqa: E402
from ffthompy_amd.tensors import classify_coefficient  # noqa: E402


def _materials():
    out = []
    for fam, dim, physics in ((examples.scalar_2d_materials, 2, 'scalar'), (examples.scalar_3d_materials, 3, 'scalar'),
                              (examples.linelas_3d_materials, 3, 'elasticity')):
        for name in fam():
            out.append((fam, name, dim, physics))
    I2 = np.eye(2)
    ties = {'inclusions': ['square', 'circle', 'pyramid', 'otherwise'],
            'params': [np.array([0.4, 0.4]), 0.6, np.array([0.4, 0.4]), ''],
            'positions': [np.array([0., 0.]), np.array([0.5, 0.5]), np.array([0.5, 0.]), ''],
            'vals': [3. * I2, 5. * I2, 2. * I2, 1. * I2], 'Y': np.ones(2)}
    out.append((lambda: {'ties': ties}, 'ties', 2, 'scalar'))
    return out


MATS = _materials()


@pytest.mark.parametrize('case', MATS, ids=['%s' % c[1] for c in MATS])
@pytest.mark.parametrize('pd', ['primal', 'dual'])
def test_device_nodal_coefficient_equals_host(case, pd):
    fam, name, dim, physics = case
    grids = [(5, 5), (7, 6), (10, 10)] if dim == 2 else [(5, 5, 5), (6, 7, 5)]
    for N in grids:
        print('case micro', name, pd, N, flush=True)
        conf = fam()[name]
        host = Material(dict(conf, positions=list(conf['positions']))).get_A_GaNi(N, pd).val
        kind, ref = classify_coefficient(host, physics)
        plan = nat.Plan(N, conf['Y'], physics, pd)
        Material(dict(conf, positions=list(conf['positions']))).set_A_GaNi_device(plan, pd)
        assert plan.coef_kind == kind
        got = plan.coeff_to_host().reshape(np.shape(ref) if kind != 'iso' else (1,) + N)
        ref = np.asarray(ref).reshape(got.shape)
        if pd == 'primal' or kind == 'iso':
            assert np.array_equal(got, ref), (name, N)
        else:
            assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max(), (name, N)


def test_overlapping_inclusions_are_rejected():
    I2 = np.eye(2)
    conf = {'inclusions': ['square', 'square', 'otherwise'], 'params': [0.6 * np.ones(2), 0.6 * np.ones(2), ''],
            'positions': [np.zeros(2), np.zeros(2), ''], 'vals': [2. * I2, 3. * I2, I2], 'Y': np.ones(2)}
    plan = nat.Plan((5, 5), (1., 1.))
    with pytest.raises(nat.FhError, match='Overlapping'):
        Material(conf).set_A_GaNi_device(plan)


@pytest.mark.parametrize('N,Y', [((63, 81), (1., 1.)), ((45, 27, 33), (1., 1.5, 0.75))])
@pytest.mark.parametrize('expo', [False, True])
def test_device_grf_equals_numpy(N, Y, expo):
    got = ms.gaussian_random_field_device(N, 0.12, 0.9, seed=11, Y=Y, exponentiate=expo).to_host()
    ref = ms.gaussian_random_field(N, 0.12, 0.9, seed=11, Y=Y)
    if expo:
        ref = np.exp(ref)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_device_coefficient_from_device_grf_matches_host_upload():
    N = (27, 27, 27)
    a = ms.gaussian_random_field_device(N, 0.15, 0.8, seed=3, exponentiate=True)
    p1 = nat.Plan(N, (1., 1., 1.))
    p1.set_coeff('iso', a)
    p2 = nat.Plan(N, (1., 1., 1.))
    p2.set_coeff('iso', a.to_host())
    assert np.array_equal(p1.coeff_to_host(), p2.coeff_to_host())
    x = nat.DeviceArray.from_host(np.random.default_rng(0).standard_normal((3,) + N))
    y1, y2 = nat.DeviceArray((3,) + N), nat.DeviceArray((3,) + N)
    p1.matvec(x, y1)
    p2.matvec(x, y2)
    assert np.array_equal(y1.to_host(), y2.to_host())
    bad = a.to_host()
    bad[0, 0, 0] = np.nan
    with pytest.raises(nat.FhError):
        p1.set_coeff('iso', nat.DeviceArray.from_host(bad))
