"""CPU checks of the configuration fixtures and of the host-side elastic constants.

* The oracle reproduces tests/golden/oracle_configs.json (configuration 1 at 45^2 and
  configuration 3 at 45^3, written by tests/golden/make_oracle_configs.py).
* The product's Lame parameters of an (E, nu) phase give the oracle's ElasticTensor
  (ffthompy/mechanics/matcoef.py:12-71) in Mandel form; the closed-form Lame compliance equals
  the oracle's pointwise Gauss-Jordan inverse (ffthompy/trigpol.py:120-159); the Mandel -> Voigt
  map turns the Mandel stiffness into the engineering-shear Voigt matrix.
"""
import json
import os

import numpy as np

from oracle.grid import pointwise_inverse
from oracle.homogenize import Problem as OracleProblem
from oracle.matcoef import ElasticTensor as OElastic
from ffthompy_amd import microstructure as ms
from ffthompy_amd.materials import lame_compliance, mandel_to_voigt

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, 'golden', 'oracle_configs.json')))


def _load_maker():
    import importlib.util
    spec = importlib.util.spec_from_file_location('make_oracle_configs',
                                                  os.path.join(HERE, 'golden', 'make_oracle_configs.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _mandel_lame(lam, mu):
    m = np.array([1., 1., 1., 0., 0., 0.])
    return lam * np.outer(m, m) + 2. * mu * np.eye(6)


def test_lame_from_young_matches_oracle_elastic_tensor():
    for E, nu in ((1.0, 0.3), (10.0, 0.2), (3.5, 0.45)):
        lam, mu = ms.lame_from_young(E, nu)
        assert np.abs(_mandel_lame(lam, mu) - OElastic.from_young(E, nu).mandel).max() < 1e-13 * E


def test_lame_compliance_is_the_pointwise_inverse():
    rng = np.random.default_rng(3)
    lam = rng.uniform(-0.3, 5., (4, 3))
    mu = rng.uniform(0.2, 4., (4, 3))
    C = np.einsum('ij,...->ij...', _mandel_lame(1., 0.), lam) + np.einsum('ij,...->ij...', _mandel_lame(0., 1.), mu)
    lc, mc = lame_compliance(lam, mu)
    S = np.einsum('ij,...->ij...', _mandel_lame(1., 0.), lc) + np.einsum('ij,...->ij...', _mandel_lame(0., 1.), mc)
    Sref = pointwise_inverse(C)
    assert np.abs(S - Sref).max() < 1e-13 * np.abs(Sref).max()


def test_mandel_to_voigt():
    """sigma = C eps in Mandel (shears weighted sqrt2) equals sigma_voigt = C_voigt gamma for
    engineering shears gamma = 2 eps_ij, with the stress's shear components unweighted."""
    C = OElastic.from_young(2., 0.25).mandel
    Cv = mandel_to_voigt(C)
    lam, mu = ms.lame_from_young(2., 0.25)
    ref = np.zeros((6, 6))
    ref[:3, :3] = lam
    ref[:3, :3] += 2 * mu * np.eye(3)
    ref[3:, 3:] = mu * np.eye(3)
    assert np.abs(Cv - ref).max() < 1e-14


def test_config1_fixture_reproduced_by_oracle():
    mk = _load_maker()
    pb = OracleProblem(mk.config1_problem())
    pb.calculate()
    c1 = FIX['config1']
    for pd in ('primal', 'dual'):
        assert [r['info']['kit'] for r in pb.output['res_' + pd]] == c1['kit_' + pd]
        for key, AH in pb.output['mat_' + pd].items():
            assert np.abs(np.asarray(AH) - np.asarray(c1[key])).max() < 1e-13, key
    # the Ga post-processing pair brackets: lower < GaNi-type values < upper here
    assert c1['AH_Ga_dual'][0][0] < c1['AH_GaNi_dual'][0][0] <= c1['AH_GaNi_primal'][0][0] + 1e-12
    assert c1['AH_GaNi_primal'][0][0] < c1['AH_Ga_primal'][0][0]


def test_config3_coefficient_fixture():
    """The fixture's microstructure is the one the generator makes (count and volume fraction);
    the solves themselves (about 25 s) are rerun by the fixture script."""
    mk = _load_maker()
    ind, count, C = mk.config3_coefficient(27)
    assert ind.mean() >= 0.25 and C.shape == (6, 6, 27, 27, 27)
    c3 = FIX['config3']
    ind45, count45 = ms.ellipsoid_indicator((45, 45, 45), target_fraction=0.25, seed=c3['seed'])
    assert count45 == c3['ellipsoids'] and abs(ind45.mean() - c3['volume_fraction']) < 1e-15
    # symmetric bound matrices, GaNi primal = dual (pointwise-inverse dual) to CG accuracy
    up, lo = np.asarray(c3['AH_primal']), np.asarray(c3['AH_dual'])
    assert np.abs(up - up.T).max() < 1e-10 and np.abs(up - lo).max() < 1e-6 * np.abs(up).max()
