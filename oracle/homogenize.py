"""Homogenization driver: primal/dual solves over unit loads and the bound matrices.

Restates ffthompy/applications.py:11-176, ffthompy/postprocess.py:6-86 and the
Problem container of ffthompy/problem.py:9-83. Test infrastructure only.
"""
import itertools
import numpy as np

from . import projections as proj
from .fields import DFT, Field, Operator
from .materials import Material
from .solver import CallBack, linear_solver


def add_macro2minimizer(X, E):
    """Return X + E for a zero-mean corrector (postprocess.py:73-86)."""
    if np.allclose(X.mean(), E):
        return X
    if np.allclose(X.mean(), np.zeros_like(E)):
        EN = X.zeros_like()
        EN.set_mean(E)
        return X + EN
    raise ValueError("Field is neither zero-mean nor E-mean.")


def assembly_matrix(A, solutions):
    """AH[i,j] = <A e_i, e_j>, projecting fields to A's grid first (postprocess.py:53-70)."""
    dim = len(solutions)
    if not np.allclose(A.N, solutions[0].N):
        sols = [s.project(A.N) for s in solutions]
    else:
        sols = solutions
    AH = np.zeros((dim, dim))
    for i, j in itertools.product(range(dim), repeat=2):
        AH[i, j] = A(sols[i]) * sols[j]
    return AH


def postprocess(pb, A, mat, solutions, primaldual):
    """Named bound matrices for every requested post-processing (postprocess.py:6-51)."""
    out = {}
    for pp in pb.postprocess:
        order_name = Nname = ''
        if pp['kind'] in ('GaNi', 'gani'):
            if A.name != 'A_GaNi':
                A = mat.get_A_GaNi(pb.solve['N'], primaldual)
        elif pp['kind'] in ('Ga', 'ga'):
            if 'order' in pp:
                Nbarpp = tuple(2 * np.asarray(pb.solve['N']) - 1)
                if pp['order'] is None:
                    A = mat.get_A_Ga(Nbar=Nbarpp, primaldual=primaldual, order=None)
                else:
                    order_name = '_o' + str(pp['order'])
                    Nname = '_P%d' % np.mean(pp['P'])
                    A = mat.get_A_Ga(Nbar=Nbarpp, primaldual=primaldual,
                                     order=pp['order'], P=pp['P'])
        name = 'AH_%s%s%s_%s' % (pp['kind'], order_name, Nname, primaldual)
        AH = assembly_matrix(A, solutions)
        out[name] = AH if primaldual == 'primal' else np.linalg.inv(AH)
    return out


class Problem:
    """Parsed problem definition (problem.py:9-41). ``material`` is a dict."""

    def __init__(self, conf_problem, materials=None):
        self.__dict__.update(conf_problem)
        if isinstance(self.material, str):
            self.material = materials[self.material]
        self.Y = np.asarray(self.material['Y'], dtype=np.float64)
        self.dim = self.Y.size
        if self.physics == 'scalar':
            self.shape = (self.dim,)
        else:
            self.shape = (self.dim * (self.dim + 1) // 2,)
        self.output = {}

    def calculate(self):
        """problem.py:70-83 -> applications.scalar / elasticity."""
        run(self)


def run(pb):
    """One driver for both physics (applications.py:11-90 scalar, 93-176 elasticity)."""
    N = tuple(int(n) for n in pb.solve['N'])
    if pb.physics == 'scalar':
        _, hG1, hG2 = proj.scalar(N, pb.Y, NyqNul=True)
        nload = pb.dim
    elif pb.physics == 'elasticity':
        _, G1h, G1s, G2h, G2s = proj.elasticity(N, pb.Y, NyqNul=True)
        hG1, hG2 = G1h + G1s, G2h + G2s
        nload = pb.dim * (pb.dim + 1) // 2
    else:
        raise NotImplementedError(pb.physics)
    if pb.solve['kind'] == 'GaNi':
        Nbar = N
    elif pb.solve['kind'] == 'Ga':
        Nbar = tuple(2 * n - 1 for n in N)
        hG1 = hG1.enlarge(Nbar)
        hG2 = hG2.enlarge(Nbar)
    FN = DFT(Nbar, inverse=False)
    FiN = DFT(Nbar, inverse=True)
    G1N = Operator([[FiN, hG1, FN]], name='G1')
    G2N = Operator([[FiN, hG2, FN]], name='G2')
    pb.operators = {'hG1': hG1, 'hG2': hG2}
    for primaldual in pb.solve['primaldual']:
        mat = Material(pb.material)
        if pb.solve['kind'] == 'GaNi':
            A = mat.get_A_GaNi(N, primaldual)
        else:
            A = mat.get_A_Ga(Nbar=Nbar, primaldual=primaldual)
        GN = G1N if primaldual == 'primal' else G2N
        Afun = Operator([[GN, A]], name='FiGFA')
        solutions, results = [], []
        for iL in range(nload):
            E = np.zeros(nload)
            E[iL] = 1.
            EN = Field(shape=(nload,), N=Nbar, Y=pb.Y)
            EN.set_mean(E)
            x0 = Field(shape=(nload,), N=Nbar, Y=pb.Y)
            B = Afun(-EN)
            cb = CallBack(A=Afun, B=B)
            X, info = linear_solver(Afun, B, x0=x0, par=dict(pb.solver), solver=pb.solver['kind'],
                                    callback=cb)
            solutions.append(add_macro2minimizer(X, E))
            results.append({'info': info, 'res_norm': cb.res_norm, 'B': B, 'X': X})
        pb.output['A_' + primaldual] = A
        pb.output['sol_' + primaldual] = solutions
        pb.output['res_' + primaldual] = results
        pb.output['mat_' + primaldual] = postprocess(pb, A, mat, solutions, primaldual)
    return pb.output


def homogenize_field(A, physics, N, Y, primaldual, tol, maxiter=1000):
    """GaNi solve of every unit load for an explicit coefficient field A ((k, k)+N, origin 0; for the
    dual problem already the pointwise inverse): the loop of ffthompy/applications.py:58-90 (scalar)
    and :143-176 (elasticity) with the driver's materialised projection, B = Afun(-E), CG from
    x0 = 0 (ffthompy/general/solver.py:80-139) and assembly_matrix (ffthompy/postprocess.py:53-70;
    the dual matrix inverted, postprocess.py:43-46). Returns (AH, kits, norm_res)."""
    N = tuple(int(n) for n in N)
    Y = np.asarray(Y, dtype=np.float64)
    if physics == 'scalar':
        _, hG1, hG2 = proj.scalar(N, Y, NyqNul=True)
    else:
        _, G1h, G1s, G2h, G2s = proj.elasticity(N, Y, NyqNul=True)
        hG1, hG2 = G1h + G1s, G2h + G2s
    nload = A.shape[0]
    GN = Operator([[DFT(N, inverse=True), hG1 if primaldual == 'primal' else hG2, DFT(N)]], name='G')
    Af = Field(val=A, order=2, multype=21, name='A_GaNi', Y=Y)
    Afun = Operator([[GN, Af]], name='FiGFA')
    sols, kits, norms = [], [], []
    for iL in range(nload):
        E = np.zeros(nload)
        E[iL] = 1.
        EN = Field(shape=(nload,), N=N, Y=Y)
        EN.set_mean(E)
        x0 = Field(shape=(nload,), N=N, Y=Y)
        B = Afun(-EN)
        X, info = linear_solver(Afun, B, x0=x0, par={'tol': tol, 'maxiter': maxiter}, solver='CG')
        kits.append(int(info['kit']))
        norms.append(float(info['norm_res']))
        sols.append(add_macro2minimizer(X, E))
    AH = assembly_matrix(Af, sols)
    if primaldual == 'dual':
        AH = np.linalg.inv(AH)
    return AH, kits, norms
