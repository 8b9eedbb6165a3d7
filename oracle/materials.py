"""Material coefficients: nodal (GaNi) and exactly integrated (Ga) on grids.

Restates ffthompy/materials.py. Test infrastructure only (see oracle/__init__.py).
"""
import itertools
import numpy as np
import scipy.special as sps

from . import grid
from .fields import Field, cfftnc, icfftnc

INCLUSION_KEYS = {'ball': ('ball', 'circle'),
                  'cube': ('cube', 'square'),
                  'pyramid': ('bilinear_pyramid', 'pyramid')}   # materials.py:9-12


class Material:
    """Material defined by inclusions or by a user function 'fun' (materials.py:15-52)."""

    def __init__(self, conf):
        if 'Y' not in conf:
            raise ValueError("The definition of PUC size (Y) is missing!")
        self.conf = conf
        self.Y = np.asarray(conf['Y'], dtype=np.float64)
        if 'fun' in conf:
            return
        if 'inclusions' not in conf:
            raise NotImplementedError("Improper material definition!")
        n = len(conf['inclusions'])
        for key in ('inclusions', 'positions', 'params', 'vals'):
            if key in conf and len(conf[key]) != n:
                raise ValueError("Improper no. of values in material for (%s)!" % key)
        for ii, incl in enumerate(conf['inclusions']):
            if incl in ('all', 'otherwise'):
                continue
            if np.any(np.greater(conf['params'][ii], self.Y)):
                raise ValueError("Improper parameters of inclusion!")
            conf['positions'][ii] = np.asarray(conf['positions'][ii]) % self.Y

    # ---- nodal evaluation (GaNi): materials.py:115-123, 171-301
    def get_A_GaNi(self, N, primaldual='primal'):
        A = self.evaluate(grid.coordinates(N, self.Y))
        if primaldual == 'dual':
            A = A.inv()
        return A.shift()

    def evaluate(self, coord):
        Ng = coord.shape[1:]
        if 'fun' in self.conf:
            vals = np.asarray(self.conf['fun'](coord), dtype=np.float64)
            shp = vals.shape[:vals.ndim - len(Ng)]
        else:
            shp = np.asarray(self.conf['vals'][0]).shape
            vals = np.zeros(shp + Ng)
            for ii, topo in enumerate(self.topologies(coord)):
                vals += np.multiply.outer(np.asarray(self.conf['vals'][ii]), topo)
        multype = {2: 21, 4: 42}[len(shp)]
        return Field(val=vals, order=len(shp), Y=self.Y, multype=multype, name='A_GaNi',
                     origin='c')

    def topologies(self, coord):
        """Characteristic functions of the inclusions at the nodes, periodic over the
        27 (or 9) neighbouring cells (materials.py:212-301)."""
        conf = self.conf
        dim = coord.shape[0]
        Ng = coord.shape[1:]
        shifts = list(itertools.product((-1, 0, 1), repeat=dim))
        out = []
        for ii, kind in enumerate(conf['inclusions']):
            if kind in INCLUSION_KEYS['cube']:
                p = np.asarray(conf['params'][ii], dtype=np.float64)
                pos = np.asarray(conf['positions'][ii], dtype=np.float64)
                t = np.zeros(Ng)
                for s in shifts:
                    loc = np.ones(Ng)
                    for dd in range(dim):
                        r = coord[dd] - pos[dd] + self.Y[dd] * s[dd]
                        loc *= (r > -p[dd] / 2) * (r <= p[dd] / 2)
                    t += loc
                out.append(t)
            elif kind in INCLUSION_KEYS['ball']:
                pos = np.asarray(conf['positions'][ii], dtype=np.float64)
                t = np.zeros(Ng)
                for s in shifts:
                    r2 = 0.
                    for dd in range(dim):
                        r2 = r2 + (coord[dd] - pos[dd] - self.Y[dd] * s[dd]) ** 2
                    t += (r2 ** 0.5 < conf['params'][ii] / 2)
                out.append(t)
            elif kind in INCLUSION_KEYS['pyramid']:
                p = np.asarray(conf['params'][ii], dtype=np.float64)
                pos = np.asarray(conf['positions'][ii], dtype=np.float64)
                t = np.zeros(Ng)
                for s in shifts:
                    loc = np.ones(Ng)
                    for dd in range(dim):
                        r = coord[dd] - pos[dd] + 2 * self.Y[dd] * s[dd]
                        loc *= np.maximum(1 - np.abs(r) / (p[dd] / 2.), 0.)
                    t += loc
                out.append(t)
            elif kind == 'otherwise':
                t = np.ones(Ng)
                for prev in out:
                    t = t - prev
                if not (t >= 0).all():
                    raise NotImplementedError("Overlapping inclusions!")
                out.append(t)
            elif kind == 'all':
                out.append(np.ones(Ng))
            else:
                raise NotImplementedError("Inclusion (%s) is not implemented." % kind)
        return out

    # ---- exact integration (Ga): materials.py:54-113, 125-169
    def get_A_Ga(self, Nbar, primaldual='primal', order=-1, P=None):
        Nbar = tuple(int(n) for n in Nbar)
        if order == -1:
            if 'order' not in self.conf:
                raise ValueError('The material order is undefined!')
            order = self.conf['order']
        elif order not in (None, 'exact', 0, 1):
            raise ValueError('Wrong material order (%s)!' % str(order))
        if order in (None, 'exact'):
            chars = self.shape_functions(Nbar)
            shp = np.asarray(self.conf['vals'][0]).shape
            val = np.zeros(shp + Nbar)
            for ii in range(len(self.conf['inclusions'])):
                Ai = np.asarray(self.conf['vals'][ii], dtype=np.float64)
                if primaldual == 'dual':
                    Ai = np.linalg.inv(Ai)
                val += np.multiply.outer(Ai, chars[ii])
            name = 'A_Ga'
        else:
            if P is None and 'P' in self.conf:
                P = self.conf['P']
            P = np.asarray(P, dtype=np.int64)
            vals = self.evaluate(grid.coordinates(P, self.Y))
            if primaldual == 'dual':
                vals = vals.inv()
            h = self.Y / P
            W = weights_con(h, Nbar, self.Y) if order in (0, 'constant') else \
                weights_lin(h, Nbar, self.Y)
            val = np.zeros(vals.shape + Nbar)
            Pt = tuple(int(p) for p in P)
            for m, n in itertools.product(*(range(s) for s in vals.shape)):
                hA0 = np.prod(P) * cfftnc(vals.val[m, n], Pt)
                if np.allclose(P, Nbar):
                    hA = hA0
                elif np.all(np.greater_equal(P, Nbar)):
                    hA = grid.decrease_array(hA0, Nbar)
                elif np.all(np.less(P, Nbar)):
                    factor = np.ceil(np.asarray(Nbar, dtype=np.float64) / 2 / P)
                    hA = grid.decrease_array(np.tile(hA0, 2 * factor.astype(np.int64) + 1), Nbar)
                else:
                    raise NotImplementedError("combination of Nbar and P")
                val[m, n] = icfftnc(W * hA, Nbar)
            name = 'A_Ga_o{0}_P{1}'.format(order, np.asarray(P).max())
        return Field(val=val, order=2, Y=self.Y, multype=21, name=name, origin='c').shift()

    def shape_functions(self, N2):
        conf = self.conf
        N2 = tuple(int(n) for n in N2)
        chars = []
        for ii, incl in enumerate(conf['inclusions']):
            pos = conf['positions'][ii]
            if isinstance(pos, np.ndarray):
                SS = shift_phase(N2, pos, self.Y)
            if incl in INCLUSION_KEYS['cube']:
                W = weights_con(np.asarray(conf['params'][ii], dtype=np.float64), N2, self.Y)
                chars.append(icfftnc(SS * W, N2))
            elif incl in INCLUSION_KEYS['ball']:
                r = conf['params'][ii] / 2
                W = np.zeros(N2) if r == 0 else weights_ball(r, N2, self.Y)
                chars.append(icfftnc(SS * W, N2))
            elif incl in INCLUSION_KEYS['pyramid']:
                W = weights_lin(np.asarray(conf['params'][ii], dtype=np.float64) / 2., N2, self.Y)
                chars.append(icfftnc(SS * W, N2))
            elif incl == 'all':
                chars.append(np.ones(N2))
            elif incl == 'otherwise':
                c = np.ones(N2)
                for jj in range(len(conf['inclusions']) - 1):
                    c = c - chars[jj]
                chars.append(c)
            else:
                raise NotImplementedError('The inclusion (%s) is not supported!' % incl)
        return chars


def shift_phase(N, h, Y):
    """exp(-2 pi i h.k/Y) over the ifftshift-ordered index set (materials.py:304-317)."""
    N = tuple(int(n) for n in N)
    ks = grid.index_set(N)
    SS = np.ones(N, dtype=np.complex128)
    for ii in range(len(N)):
        shp = [1] * len(N)
        shp[ii] = N[ii]
        SS = SS * np.exp(-2j * np.pi * h[ii] * ks[ii] / Y[ii]).reshape(shp)
    return SS


def _sep_weights(h, Nbar, Y, power):
    d = np.size(Y)
    ks = grid.index_set(Nbar, fft_form='c')
    W = np.ones(Nbar) / np.prod(Y)
    for ii in range(d):
        shp = [1] * d
        shp[ii] = Nbar[ii]
        W = W * (h[ii] * np.sinc(h[ii] * ks[ii] / Y[ii]) ** power).reshape(shp)
    return W


def weights_con(h, Nbar, Y):
    """Box of side h: prod_i h_i sinc(h_i k_i / Y_i) / |Y| (materials.py:320-349)."""
    return _sep_weights(h, tuple(int(n) for n in Nbar), Y, 1)


def weights_lin(h, Nbar, Y):
    """Tent of half-width h: prod_i h_i sinc^2(h_i k_i / Y_i) / |Y| (materials.py:352-378)."""
    return _sep_weights(h, tuple(int(n) for n in Nbar), Y, 2)


def weights_ball(r, Nbar, Y):
    """Disk of radius r: r^2 J1(2 pi |k| r)/(|k| r) / |Y|, pi r^2 at k=0
    (materials.py:381-425). The reference applies this 2-D formula in 3-D too."""
    Nbar = tuple(int(n) for n in Nbar)
    d = np.size(Y)
    ks = grid.index_set(Nbar, fft_form='c')
    circ = np.zeros(Nbar)
    for m in range(d):
        shp = [1] * d
        shp[m] = Nbar[m]
        circ = circ + ((ks[m] / Y[m]) ** 2).reshape(shp)
    circ = circ ** 0.5
    ind = grid.mean_index(Nbar, fft_form='c')
    circ[ind] = 1.
    W = r ** 2 * sps.jv(1, 2 * np.pi * circ * r) / (circ * r)
    W[ind] = np.pi * r ** 2
    return W / np.prod(Y)
