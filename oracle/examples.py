"""The reference's example problems, restated as data.

Material and problem definitions of examples/scalar/scalar_2d.py:15-161,
examples/scalar/scalar_3d.py:14-127, examples/scalar/from_file.py:12-66 and
examples/elasticity/linelas_3d.py:13-62; these produced the golden results in
test_results/python3/. The solver tolerance of each problem is the one the
golden run used (1e-6, 1e-8 for scalar_2d prob5), which the example files
themselves have since drifted from. Test infrastructure only.
"""
import numpy as np

from .matcoef import ElasticTensor

# examples/scalar/topology.txt (5x5 image: 0 = matrix, 1 = inclusion)
TOPOLOGY_5x5 = np.array([[0, 0, 0, 0, 0],
                         [0, 1, 1, 1, 0],
                         [0, 1, 1, 1, 0],
                         [0, 1, 1, 1, 0],
                         [0, 0, 0, 0, 0]], dtype=np.float64)


def _two_phase(kind, param, vals, Y, dim, **extra):
    m = {'inclusions': [kind, 'otherwise'], 'positions': [np.zeros(dim), ''],
         'params': [param, ''], 'vals': vals, 'Y': Y}
    m.update(extra)
    return m


def scalar_2d_materials():
    dim = 2
    N = np.array([5, 5])
    I = np.eye(dim)
    return {
        'square': _two_phase('square', 0.6 * np.ones(dim), [11 * I, 1. * I], np.ones(dim), dim, order=None),
        'square_Ga': _two_phase('square', 0.6 * np.ones(dim), [11 * I, 1. * I], np.ones(dim), dim,
                                order=0, P=N),
        'ball': _two_phase('ball', 1., [11 * I, 1. * I], np.ones(dim), dim, order=None),
        'pyramid': {'inclusions': ['pyramid', 'all'], 'positions': [np.zeros(dim), ''],
                    'params': [1. * np.ones(dim), ''], 'vals': [10. * I, 1. * I],
                    'Y': np.ones(dim), 'order': None},
        'laminate': _two_phase('square', np.array([1., 0.5]), [11 * I, 1. * I], np.ones(dim), dim,
                               order=None),
    }


def scalar_3d_materials():
    dim = 3
    N = np.array([5, 5, 5])
    I = np.eye(dim)
    return {
        'cube': _two_phase('cube', 0.7 * np.ones(dim), [11 * I, 1. * I], np.ones(dim), dim, order=None),
        'cube_Ga': _two_phase('cube', 1.4 * np.ones(dim), [11 * I, 1. * I], 2. * np.ones(dim), dim,
                              order=0, P=2 * N),
    }


def from_file_materials():
    dim = 2

    def get_mat(coord=None):
        topo = TOPOLOGY_5x5
        if coord is not None and topo.shape != coord.shape[1:]:
            raise ValueError()
        vals = np.multiply.outer(np.eye(dim), topo == 0).astype(np.float64)
        vals += np.multiply.outer(11. * np.eye(dim), topo == 1)
        return vals

    return {'file': {'fun': get_mat, 'Y': np.ones(dim), 'order': 0, 'P': np.array([5, 5])}}


def linelas_3d_materials():
    dim = 3
    M = ElasticTensor(bulk=1, mu=1)
    I = ElasticTensor(bulk=10, mu=5)
    return {'square': _two_phase('square', 0.6 * np.ones(dim), [I.mandel, M.mandel], np.ones(dim), dim,
                                 order=None, P=np.array([5, 5, 5]))}


def _cg(tol):
    return {'kind': 'CG', 'tol': tol, 'maxiter': 1e3}


def problems():
    """(file, problem name, problem dict, materials dict) for every golden problem."""
    N2 = np.array([5, 5])
    N3 = np.array([5, 5, 5])
    pd = ['primal', 'dual']
    out = []
    m2 = scalar_2d_materials
    out.append(('scalar_2d', 'prob1', {
        'physics': 'scalar', 'material': 'square', 'solve': {'kind': 'GaNi', 'N': N2, 'primaldual': pd},
        'postprocess': [{'kind': 'GaNi'}, {'kind': 'Ga', 'order': None}, {'kind': 'Ga', 'order': 0, 'P': N2},
                        {'kind': 'Ga', 'order': 1, 'P': 27 * N2}], 'solver': _cg(1e-6)}, m2))
    out.append(('scalar_2d', 'prob2', {
        'physics': 'scalar', 'material': 'square', 'solve': {'kind': 'Ga', 'N': N2, 'primaldual': pd},
        'postprocess': [{'kind': 'Ga'}], 'solver': _cg(1e-6)}, m2))
    out.append(('scalar_2d', 'prob3', {
        'physics': 'scalar', 'material': 'square_Ga', 'solve': {'kind': 'Ga', 'N': N2, 'primaldual': pd},
        'postprocess': [{'kind': 'Ga'}], 'solver': _cg(1e-6)}, m2))
    out.append(('scalar_2d', 'prob4', {
        'physics': 'scalar', 'material': 'pyramid', 'solve': {'kind': 'Ga', 'N': N2, 'primaldual': ['primal']},
        'postprocess': [{'kind': 'Ga', 'order': None}, {'kind': 'Ga', 'order': 0, 'P': N2},
                        {'kind': 'Ga', 'order': 1, 'P': N2}], 'solver': _cg(1e-6)}, m2))
    out.append(('scalar_2d', 'prob5', {
        'physics': 'scalar', 'material': 'ball', 'solve': {'kind': 'Ga', 'N': N2, 'primaldual': pd},
        'postprocess': [{'kind': 'Ga', 'order': None}, {'kind': 'Ga', 'order': 0, 'P': N2}],
        'solver': _cg(1e-8)}, m2))
    m3 = scalar_3d_materials
    out.append(('scalar_3d', 'prob1', {
        'physics': 'scalar', 'material': 'cube', 'solve': {'kind': 'GaNi', 'N': N3, 'primaldual': pd},
        'postprocess': [{'kind': 'GaNi'}, {'kind': 'Ga', 'order': None}, {'kind': 'Ga', 'order': 0, 'P': 3 * N3},
                        {'kind': 'Ga', 'order': 1, 'P': 3 * N3}], 'solver': _cg(1e-6)}, m3))
    out.append(('scalar_3d', 'prob2', {
        'physics': 'scalar', 'material': 'cube', 'solve': {'kind': 'Ga', 'N': N3, 'primaldual': pd},
        'postprocess': [{'kind': 'Ga'}], 'solver': _cg(1e-6)}, m3))
    out.append(('scalar_3d', 'prob3', {
        'physics': 'scalar', 'material': 'cube_Ga', 'solve': {'kind': 'Ga', 'N': N3, 'primaldual': pd},
        'postprocess': [{'kind': 'Ga'}], 'solver': _cg(1e-6)}, m3))
    mf = from_file_materials
    out.append(('from_file', 'prob1', {
        'physics': 'scalar', 'material': 'file', 'solve': {'kind': 'GaNi', 'N': N2, 'primaldual': pd},
        'postprocess': [{'kind': 'GaNi'}, {'kind': 'Ga', 'order': 0, 'P': N2}, {'kind': 'Ga', 'order': 1, 'P': N2}],
        'solver': _cg(1e-6)}, mf))
    out.append(('from_file', 'prob2', {
        'physics': 'scalar', 'material': 'file', 'solve': {'kind': 'Ga', 'N': N2, 'primaldual': pd},
        'postprocess': [{'kind': 'Ga'}], 'solver': _cg(1e-6)}, mf))
    me = linelas_3d_materials
    out.append(('linelas_3d', 'prob1', {
        'physics': 'elasticity', 'material': 'square', 'solve': {'kind': 'GaNi', 'N': N3, 'primaldual': pd},
        'postprocess': [{'kind': 'GaNi'}, {'kind': 'Ga', 'order': None}, {'kind': 'Ga', 'order': 0, 'P': N3},
                        {'kind': 'Ga', 'order': 1, 'P': N3}], 'solver': _cg(1e-6)}, me))
    out.append(('linelas_3d', 'prob2', {
        'physics': 'elasticity', 'material': 'square', 'solve': {'kind': 'Ga', 'N': N3, 'primaldual': pd},
        'postprocess': [{'kind': 'Ga', 'order': None}], 'solver': _cg(1e-6)}, me))
    return out
