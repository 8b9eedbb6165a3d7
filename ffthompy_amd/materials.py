"""Material coefficients on the grid (host side): the GaNi nodal coefficient A(x).

Mirrors ffthompy/materials.py: materials are defined by inclusions ('square'/'cube',
'ball'/'circle', 'pyramid', 'otherwise', 'all') with per-phase coefficient matrices, or
by a user function 'fun' of the nodal coordinates (materials.py:171-210). Evaluation
follows the reference's rules: nodes x = Y*k/N on the centred index set
(ffthompy/trigpol.py:55-71), periodic copies over the neighbouring cells
(materials.py:226-275), inclusion membership > -a/2 and <= a/2 per axis for boxes and
< r for balls; the dual problem uses the pointwise inverse (materials.py:121-122) and the
result is shifted to origin 0 (objects.py:169-186). Exact-integration ('Ga')
coefficients (materials.py:54-113) are described here and evaluated on the device
(GaCoefficient, ffthompy_amd._native.Ga).
"""
import itertools

import numpy as np

from .tensors import Tensor, classify_coefficient

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

    def _inclusion_arrays(self):
        """Inclusion kinds, (nincl, d) parameters (a ball's diameter in column 0) and positions,
        and the phase coefficient matrices."""
        d = len(self.Y)
        kinds = list(self.conf['inclusions'])
        params = np.zeros((len(kinds), d))
        positions = np.zeros((len(kinds), d))
        mats = []
        for ii, kind in enumerate(kinds):
            if kind not in GA_KINDS:
                raise NotImplementedError('The inclusion (%s) is not supported!' % kind)
            if kind in INCLUSION_KEYS['ball']:
                params[ii, 0] = float(np.asarray(self.conf['params'][ii]).reshape(-1)[0])
            elif kind not in ('all', 'otherwise'):
                params[ii] = np.asarray(self.conf['params'][ii], dtype=np.float64)
            if kind not in ('all', 'otherwise'):
                positions[ii] = np.asarray(self.conf['positions'][ii], dtype=np.float64)
            mats.append(np.asarray(self.conf['vals'][ii], dtype=np.float64))
        return kinds, params, positions, mats

    def set_A_GaNi_device(self, plan, primaldual='primal'):
        """get_A_GaNi evaluated on the device straight into `plan` (materials defined by
        inclusions): same nodes, membership rules and pointwise dual inverse, no host arrays of
        the grid's size (fh_set_coeff_inclusions)."""
        kinds, params, positions, mats = self._inclusion_arrays()
        kind, vals = classify_coefficient(np.stack(mats, axis=-1), _physics_of(mats[0].shape[0], len(self.Y)))
        plan.set_coeff_inclusions(kinds, params, positions, kind, vals.reshape(-1, len(kinds)).T,
                                  invert=primaldual == 'dual')
        return plan

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

    def get_A_Ga(self, Nbar, primaldual='primal', order=-1, P=None):
        """Exactly integrated coefficient on the grid Nbar (materials.py:54-113), as a
        GaCoefficient evaluated on the device. order None/'exact': closed-form integrals of the
        inclusions' shape functions; order 0 / 1: box / tent interpolation of the nodal values
        on the grid P (default conf['P']). The dual problem inverts each phase (order None) or
        the nodal values (order 0/1), as materials.py:73, 84."""
        if order == -1:
            if 'order' not in self.conf:
                raise ValueError('The material order is undefined!')
            order = self.conf['order']
        elif order not in (None, 'exact', 0, 1):
            raise ValueError('Wrong material order (%s)!' % str(order))
        Nbar = tuple(int(n) for n in np.atleast_1d(Nbar))
        d = len(self.Y)
        if order in (None, 'exact'):
            if 'inclusions' not in self.conf:
                raise NotImplementedError('exact integration needs a material defined by inclusions')
            kinds, params, positions, mats = self._inclusion_arrays()
            if primaldual == 'dual':
                mats = [np.linalg.inv(Ai) for Ai in mats]
            kind, vals = classify_coefficient(np.stack(mats, axis=-1), _physics_of(mats[0].shape[0], d))
            return GaCoefficient(Nbar, self.Y, 'A_Ga', kind, vals, inclusions=(kinds, params, positions))
        if P is None and 'P' in self.conf:
            P = self.conf['P']
        P = tuple(int(p) for p in np.broadcast_to(np.asarray(P, dtype=np.int64), (d,)))
        A = self.evaluate(node_coordinates(P, self.Y))
        if primaldual == 'dual':
            A = pointwise_inverse(A)
        A = np.fft.ifftshift(A, axes=tuple(range(2, A.ndim)))
        kind, vals = classify_coefficient(A, _physics_of(A.shape[0], d))
        return GaCoefficient(Nbar, self.Y, 'A_Ga_o%s_P%d' % (order, max(P)), kind, vals, order=int(order), P=P)


GA_KINDS = INCLUSION_KEYS['ball'] + INCLUSION_KEYS['cube'] + INCLUSION_KEYS['pyramid'] + ('all', 'otherwise')


def _physics_of(k, d):
    return 'elasticity' if k == d * (d + 1) // 2 and k != d else 'scalar'


class GaCoefficient:
    """Exactly integrated coefficient A_Ga of materials.py:54-113 on the grid Nbar, held as its
    description (inclusions, or nodal values on P with interpolation order 0/1) and evaluated
    on the device by an ffthompy_amd._native.Ga, whose quadrature grid may be finer than
    Nbar (the energies are the same, see fh_ga_create). kind/vals: compact phase or nodal
    coefficients ('iso', 'lame' or 'full' layout)."""

    def __init__(self, Nbar, Y, name, kind, vals, inclusions=None, order=None, P=None):
        self.N = tuple(Nbar)
        self.Y = np.asarray(Y, dtype=np.float64)
        self.name = name
        self.kind = kind
        self.vals = np.asarray(vals, dtype=np.float64)
        self.inclusions = inclusions
        self.order = order
        self.P = P

    def add_mean(self, mean):
        """Add a constant matrix (the anisotropic part of the reference's low-rank examples,
        Tensor.add_mean on A_Ga): an inclusion 'all' with that value, or the value added to every
        nodal value (box and tent interpolation reproduce constants). The layout becomes 'full'."""
        mean = np.asarray(mean, dtype=np.float64)
        k = mean.shape[0]
        if self.kind == 'iso':
            full = np.einsum('ij,...->ij...', np.eye(k), self.vals).reshape((k * k,) + self.vals.shape)
        elif self.kind == 'full':
            full = self.vals.reshape((k * k,) + self.vals.shape[1:] if self.inclusions is None
                                     else (k * k, -1))
        else:
            raise NotImplementedError('add_mean of a %s coefficient' % self.kind)
        if self.inclusions is not None:
            kinds, params, positions = self.inclusions
            d = len(self.Y)
            params = np.vstack([np.asarray(params, dtype=np.float64).reshape(-1, d), np.zeros((1, d))])
            positions = np.vstack([np.asarray(positions, dtype=np.float64).reshape(-1, d), np.zeros((1, d))])
            self.inclusions = (list(kinds) + ['all'], params, positions)
            full = np.concatenate([full.reshape(k * k, -1), mean.reshape(k * k, 1)], axis=1)
        else:
            full = full + mean.reshape((k * k,) + (1,) * (full.ndim - 1))
        self.kind, self.vals = 'full', full
        return self

    def load(self, ga):
        """Evaluate this coefficient on the quadrature grid of `ga`."""
        if self.inclusions is not None:
            kinds, params, positions = self.inclusions
            nincl = len(kinds)
            vals = self.vals.reshape(-1, nincl).T
            ga.set_inclusions(kinds, params, positions, self.kind, vals)
        else:
            ga.set_grid_coeff(self.order, self.P, self.kind, self.vals)
        return ga

    def __repr__(self):
        return 'GaCoefficient(%s, Nbar=%s, %s)' % (self.name, self.N, self.kind)


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


def lame_compliance(lam, mu):
    """Lame-form parameters (lambda', mu') of the compliance of C = lambda m m^T + 2 mu I (Mandel,
    m = (1,1,1,0,0,0)): C^-1 = lambda' m m^T + 2 mu' I with mu' = 1/(4 mu) and
    lambda' = -lambda / (2 mu (3 lambda + 2 mu)) (Sherman-Morrison on the rank-one term). The dual
    problem's coefficient is this pointwise inverse (ffthompy/materials.py:121-122,
    ffthompy/trigpol.py:120-159); works on arrays."""
    lam = np.asarray(lam, dtype=np.float64)
    mu = np.asarray(mu, dtype=np.float64)
    return -lam / (2. * mu * (3. * lam + 2. * mu)), 1. / (4. * mu)


# Mandel -> Voigt for a stiffness-like 6x6 matrix (ffthompy/mechanics/matcoef.py:151-207 builds
# the Mandel form with sqrt(2) shear weights): AH_voigt = T AH_mandel T, T = diag(1,1,1,1/r2,1/r2,1/r2).
VOIGT_FROM_MANDEL = np.diag([1., 1., 1., 2 ** -0.5, 2 ** -0.5, 2 ** -0.5])


def mandel_to_voigt(AH):
    """Stiffness-like homogenized matrix in Voigt notation (engineering shear strains)."""
    return VOIGT_FROM_MANDEL @ np.asarray(AH) @ VOIGT_FROM_MANDEL
