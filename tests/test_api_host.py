"""Host-side parts of the drop-in API (no GPU): material evaluation, coefficient
classification and problem parsing, checked against the oracle's restatement."""
import numpy as np
import pytest

from oracle import examples
from oracle.materials import Material as OMaterial
from ffthompy_amd.materials import Material, ElasticTensor, pointwise_inverse
from ffthompy_amd.tensors import classify_coefficient
from ffthompy_amd.problem import Problem


def _all_materials():
    out = []
    for f in (examples.scalar_2d_materials, examples.scalar_3d_materials, examples.from_file_materials,
              examples.linelas_3d_materials):
        for name in f():
            out.append((f, name))
    return out


@pytest.mark.parametrize('mf,name', _all_materials(), ids=['%s-%s' % (m.__name__, n) for m, n in _all_materials()])
@pytest.mark.parametrize('pd', ['primal', 'dual'])
def test_gani_coefficient_matches_oracle(mf, name, pd):
    conf = mf()[name]
    dim = np.size(conf['Y'])
    for n in (5, 8):
        if 'fun' in conf and n != 5:
            continue
        N = (n,) * dim
        A = Material(mf()[name]).get_A_GaNi(N, pd).val
        Ao = OMaterial(mf()[name]).get_A_GaNi(N, pd).val
        assert np.abs(A - Ao).max() < 1e-12 * max(1., np.abs(Ao).max())


def test_classify_coefficient():
    rng = np.random.default_rng(0)
    a = 1. + rng.random((5, 6))
    kind, v = classify_coefficient(np.einsum('ij,...->ij...', np.eye(2), a), 'scalar')
    assert kind == 'iso' and np.array_equal(v, a)
    lam, mu = 1. + rng.random((4, 4, 4)), 1. + rng.random((4, 4, 4))
    m = np.array([1., 1., 1., 0, 0, 0])
    C = np.einsum('i,j,...->ij...', m, m, lam) + 2 * np.einsum('ij,...->ij...', np.eye(6), mu)
    kind, v = classify_coefficient(C, 'elasticity')
    assert kind == 'lame' and np.allclose(v[0], lam) and np.allclose(v[1], mu)
    C2 = C.copy()
    C2[0, 3] += 0.1
    kind, v = classify_coefficient(C2, 'elasticity')
    assert kind == 'full' and v.shape == (36, 4, 4, 4)
    # the elastic constants of the reference's linelas example are Lame-isotropic
    E = ElasticTensor(bulk=10, mu=5).mandel
    kind, v = classify_coefficient(np.einsum('ij,...->ij...', E, np.ones((3, 3, 3))), 'elasticity')
    assert kind == 'lame'


def test_pointwise_inverse():
    rng = np.random.default_rng(1)
    A = rng.random((3, 3, 4, 5)) + 3 * np.eye(3)[:, :, None, None]
    inv = pointwise_inverse(A)
    assert np.abs(np.einsum('ij...,jk...->ik...', A, inv) - np.eye(3)[:, :, None, None]).max() < 1e-13


def test_problem_parsing():
    mats = examples.scalar_2d_materials()
    conf = type('conf', (), {'materials': mats})
    pb = Problem({'name': 'p', 'physics': 'scalar', 'material': 'square',
                  'solve': {'kind': 'GaNi', 'N': np.array([5, 5]), 'primaldual': ['primal']},
                  'postprocess': [{'kind': 'GaNi'}], 'solver': {'kind': 'CG', 'tol': 1e-6, 'maxiter': 1e3}}, conf)
    assert pb.dim == 2 and pb.shape == (2,)
    bad = dict(mats['square'])
    bad['inclusions'] = ['otherwise', 'square']
    with pytest.raises(ValueError):
        Problem.parse_material(bad)


def test_ga_coefficient_add_mean_layout():
    """GaCoefficient.add_mean: an inclusion coefficient gains an 'all' inclusion carrying the
    constant, a nodal one the constant on every node; both become the 'full' layout."""
    from ffthompy_amd.materials import Material
    d = 2
    K = np.array([[5.5, -4.5], [-4.5, 5.5]])
    conf = {'inclusions': ['square', 'otherwise'], 'positions': [np.zeros(d), ''], 'params': [0.6 * np.ones(d), ''],
            'vals': [10 * np.eye(d), 1. * np.eye(d)], 'Y': np.ones(d), 'P': (5, 5), 'order': None}
    A = Material(dict(conf)).get_A_Ga((9, 9), 'primal', order=None).add_mean(K)
    assert A.kind == 'full' and list(A.inclusions[0]) == ['square', 'otherwise', 'all']
    v = A.vals.reshape(d * d, 3)
    assert np.allclose(v[:, 0], (10 * np.eye(d)).ravel()) and np.allclose(v[:, 2], K.ravel())
    G = Material(dict(conf)).get_A_Ga((9, 9), 'primal', order=0, P=(5, 5))
    base = G.vals.copy()
    G.add_mean(K)
    assert G.kind == 'full' and G.vals.shape == (d * d, 5, 5)
    assert np.allclose(G.vals[1], K[0, 1]) and np.allclose(G.vals[0], base + K[0, 0])
