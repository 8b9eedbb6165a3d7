"""Homogenization drivers (GaNi and Ga): primal and dual solves over the unit loads.

Mirrors ffthompy/applications.py:11-90 (scalar) and 93-176 (elasticity). The operator
Afun = F^-1 G F A is one GpuOperator (HIP library); CG runs on the device. A Ga solve
(exactly integrated coefficient, projection of the grid N enlarged to 2N-1) runs on a
quadrature grid of at least 2N-1 points (GpuOperator.ga).
"""
import numpy as np

from .materials import Material
from .postprocess import add_macro2minimizer, postprocess
from .solver import CallBack, linear_solver
from .tensors import GpuOperator, Tensor


def _homogenize(pb, physics):
    kind = pb.solve['kind']
    if kind not in ('GaNi', 'Ga'):
        raise NotImplementedError("solve kind %r" % kind)
    N = tuple(int(n) for n in pb.solve['N'])
    nload = pb.dim if physics == 'scalar' else pb.dim * (pb.dim + 1) // 2
    for primaldual in pb.solve['primaldual']:
        mat = Material(pb.material)
        if kind == 'GaNi' and 'inclusions' in mat.conf:
            Afun = GpuOperator(N, pb.Y, physics, primaldual, _defer=True)
            mat.set_A_GaNi_device(Afun.plan, primaldual)
        elif kind == 'GaNi':
            A = mat.get_A_GaNi(N, primaldual)
            Afun = GpuOperator(N, pb.Y, physics, primaldual, A=A)
        else:
            A = mat.get_A_Ga(Nbar=tuple(2 * n - 1 for n in N), primaldual=primaldual)
            Afun = GpuOperator.ga(N, pb.Y, physics, primaldual, A)
        Ng = Afun.N  # grid of the unknowns (a Ga solve runs on its quadrature grid)
        solutions, results = [], []
        for iL in range(nload):
            E = np.zeros(nload)
            E[iL] = 1.
            EN = Tensor(name='EN', N=Ng, shape=(nload,), Y=pb.Y)
            EN.set_mean(E)
            x0 = Tensor(name='x0', N=Ng, shape=(nload,), Y=pb.Y)
            B = Afun(-EN)
            cb = CallBack(A=Afun, B=B)
            X, info = linear_solver(Afun=Afun, B=B, x0=x0, par=dict(pb.solver), solver=pb.solver['kind'],
                                    callback=cb)
            solutions.append(add_macro2minimizer(X, E))
            results.append({'cb': cb, 'info': info})
        postprocess(pb, Afun, mat, solutions, results, primaldual)


def scalar(problem):
    """Scalar elliptic homogenization (applications.py:11-90)."""
    _homogenize(problem, 'scalar')


def elasticity(problem):
    """Small-strain linear elasticity homogenization, Mandel notation (applications.py:93-176)."""
    _homogenize(problem, 'elasticity')
