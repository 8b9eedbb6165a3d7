"""Material coefficients on the grid (host side): the GaNi nodal coefficient A(x).

Mirrors ffthompy/materials.py: materials are defined by inclusions ('square'/'cube',
'ball'/'circle', 'pyramid', 'otherwise', 'all') with per-phase coefficient matrices, or
by a user function 'fun' of the nodal coordinates (materials.py:171-210). Evaluation
follows the reference's rules: nodes x = Y*k/N on the centred index set
(ffthompy/trigpol.py:55-71), periodic copies over the neighbouring cells
(materials.py:226-275), inclusion membership > -a/2 and <= a/2 per axis for boxes and
< r for balls; the dual problem uses the pointwise inverse (materials.py:121-122) and the
result is shifted to origin 0 (objects.py:169-186). Exact-integration ('Ga')
coefficients are not provided on this path.
"""
import itertools

import numpy as np

from .tensors import Tensor

INCLUSION_KEYS = {'ball': ('ball', 'circle'), 'cube': ('cube', 'square'),
                  'pyramid': ('bilinear_pyramid', 'pyramid')}


def node_coordinates(N, Y):
    """Nodal coordinates (d,)+N in centred order (ffthompy/trigpol.py:55-71)."""
    N = tuple(int(n) for n in N)
    d = len(N)
    out = np.zeros((d,) + N)
    for a in range(d):
        k = np.arange(np.fix(-N[a] / 2.), np.fix(N[a] / 2. + 0.5))
        shp = [1] * d
        shp[a] = N[a]
        out[a] = np.broadcast_to((Y[a] * k / N[a]).reshape(shp), N)
    return out


class Material:
    """Material definition of an input file (ffthompy/materials.py:15-52)."""

    def __init__(self, conf):
        if 'Y' not in conf:
            raise ValueError("The definition of PUC size (Y) is missing!")
        self.conf = conf
        self.Y = np.asarray(conf['Y'], dtype=np.float64)
        if 'fun' in conf:
            return
        if 'inclusions' not in conf:
            raise NotImplementedError("Improper material definition!")
        n_incl = len(conf['inclusions'])
        for key in ('inclusions', 'positions', 'params', 'vals'):
            if key in conf and len(conf[key]) != n_incl:
                raise ValueError("Improper no. of values in material for (%s)!" % key)
        for ii, incl in enumerate(conf['inclusions']):
            if incl in ('all', 'otherwise'):
                continue
            if np.any(np.greater(conf['params'][ii], self.Y)):
                raise ValueError("Improper parameters of inclusion!")
            conf['positions'][ii] = np.asarray(conf['positions'][ii]) % self.Y

    def get_A_GaNi(self, N, primaldual='primal'):
        """Nodal coefficient Tensor (k,k)+N, origin 0 (materials.py:115-123)."""
        A = self.evaluate(node_coordinates(N, self.Y))
        if primaldual == 'dual':
            A = pointwise_inverse(A)
        A = np.fft.ifftshift(A, axes=tuple(range(2, A.ndim)))
        return Tensor(name='A_GaNi' if primaldual == 'primal' else 'inv(A_GaNi)', val=A, order=2,
                      Y=self.Y, multype=21)

    def evaluate(self, coord):
        """Coefficient values at the nodes, (k,k)+N (materials.py:171-210)."""
        Ng = coord.shape[1:]
        if 'fun' in self.conf:
            return np.asarray(self.conf['fun'](coord), dtype=np.float64)
        shp = np.asarray(self.conf['vals'][0]).shape
        A = np.zeros(shp + Ng)
        for ii, topo in enumerate(self.topologies(coord)):
            A += np.multiply.outer(np.asarray(self.conf['vals'][ii], dtype=np.float64), topo)
        return A

    def topologies(self, coord):
        """Characteristic function of every inclusion at the nodes, summed over the
        periodic copies of the neighbouring cells (materials.py:212-301)."""
        conf = self.conf
        d = coord.shape[0]
        Ng = coord.shape[1:]
        shifts = list(itertools.product((-1, 0, 1), repeat=d))
        out = []
        for ii, kind in enumerate(conf['inclusions']):
            if kind in INCLUSION_KEYS['cube']:
                a = np.asarray(conf['params'][ii], dtype=np.float64)
                pos = np.asarray(conf['positions'][ii], dtype=np.float64)
                t = np.zeros(Ng)
                for s in shifts:
                    m = np.ones(Ng, dtype=bool)
                    for ax in range(d):
                        z = coord[ax] - pos[ax] + self.Y[ax] * s[ax]
                        m &= (z > -a[ax] / 2) & (z <= a[ax] / 2)
                    t += m
                out.append(t)
            elif kind in INCLUSION_KEYS['ball']:
                pos = np.asarray(conf['positions'][ii], dtype=np.float64)
                t = np.zeros(Ng)
                for s in shifts:
                    r2 = np.zeros(Ng)
                    for ax in range(d):
                        r2 += (coord[ax] - pos[ax] - self.Y[ax] * s[ax]) ** 2
                    t += np.sqrt(r2) < conf['params'][ii] / 2
                out.append(t)
            elif kind in INCLUSION_KEYS['pyramid']:
                a = np.asarray(conf['params'][ii], dtype=np.float64)
                pos = np.asarray(conf['positions'][ii], dtype=np.float64)
                t = np.zeros(Ng)
                for s in shifts:
                    v = np.ones(Ng)
                    for ax in range(d):
                        z = coord[ax] - pos[ax] + 2 * self.Y[ax] * s[ax]
                        v *= np.maximum(1. - np.abs(z) / (a[ax] / 2.), 0.)
                    t += v
                out.append(t)
            elif kind == 'otherwise':
                t = np.ones(Ng) - sum(out) if out else np.ones(Ng)
                if not (t >= 0).all():
                    raise NotImplementedError("Overlapping inclusions!")
                out.append(t)
            elif kind == 'all':
                out.append(np.ones(Ng))
            else:
                raise NotImplementedError("Inclusion (%s) is not implemented." % kind)
        return out

    def get_A_Ga(self, *args, **kwargs):
        raise NotImplementedError('exact-integration (Ga) coefficients are not available on the GPU path')


def pointwise_inverse(A):
    """Inverse of the (k,k) matrix at every node (the reference's Gauss-Jordan sweep,
    ffthompy/trigpol.py:120-159, evaluated by LAPACK per node)."""
    k = A.shape[0]
    M = np.moveaxis(A.reshape((k, k, -1)), -1, 0)
    inv = np.linalg.inv(M)
    return np.moveaxis(inv, 0, -1).reshape(A.shape)


class ElasticTensor:
    """Isotropic stiffness C = 3K vol + 2 mu dev in Mandel form
    (ffthompy/mechanics/matcoef.py:12-71, 151-207), 3-D."""

    def __init__(self, bulk=None, mu=None, stiffness=True):
        self.bulk, self.mu = float(bulk), float(mu)
        lam = self.bulk - 2. * self.mu / 3.
        m = np.array([1., 1., 1., 0., 0., 0.])
        C = lam * np.outer(m, m) + 2. * self.mu * np.eye(6)
        self.mandel = C if stiffness else np.linalg.inv(C)

    @staticmethod
    def from_young(E, nu):
        return ElasticTensor(bulk=E / (3 * (1 - 2 * nu)), mu=E / (2 * (1 + nu)))
