"""Isotropic elastic tensors in Mandel notation.

Restates ffthompy/mechanics/matcoef.py:4-207. Test infrastructure only.
"""
import numpy as np


def decomposition():
    """Symmetric identity, volumetric and deviatoric 4th-order projections
    (matcoef.py:126-149)."""
    I = np.eye(3)
    vol = np.einsum('ij,kl->ijkl', I, I) / 3.
    idsym = 0.5 * (np.einsum('ik,jl->ijkl', I, I) + np.einsum('il,jk->ijkl', I, I))
    return idsym, vol, idsym - vol


def mandel(T):
    """Mandel form of a 3x3x3x3 minor-symmetric tensor or a symmetric 3x3 matrix
    (matcoef.py:151-207): shear order (23, 13, 12), weights sqrt(2) and 2."""
    if T.ndim == 4:
        pairs = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
        w = [1., 1., 1., 2 ** 0.5, 2 ** 0.5, 2 ** 0.5]
        M = np.zeros((6, 6))
        for a, (i, j) in enumerate(pairs):
            for b, (k, l) in enumerate(pairs):
                M[a, b] = w[a] * w[b] * T[i, j, k, l]
        return M
    r2 = 2 ** 0.5
    return np.array([T[0, 0], T[1, 1], T[2, 2], r2 * T[1, 2], r2 * T[0, 2], r2 * T[0, 1]])


class ElasticTensor:
    """C = 3K vol + 2 mu dev (stiffness) or its inverse (matcoef.py:12-71, 3-D only)."""

    def __init__(self, bulk, mu, stiffness=True):
        self.bulk = float(bulk)
        self.mu = float(mu)
        _, vol, dev = decomposition()
        if stiffness:
            self.val = 3. * bulk * vol + 2. * mu * dev
        else:
            self.val = vol / (3 * bulk) + dev / (2 * mu)
        self.mandel = mandel(self.val)

    @staticmethod
    def from_young(E, nu):
        """Bulk and shear moduli from Young's modulus and Poisson's ratio."""
        return ElasticTensor(bulk=E / (3 * (1 - 2 * nu)), mu=E / (2 * (1 + nu)))
