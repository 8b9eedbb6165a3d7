"""Fourier projections G0, G1, G2 (scalar) and G1h, G1s, G2h, G2s (elasticity, Mandel).

Restates ffthompy/projections.py (materialised, float64) and the closed forms of
ffthompy/tensors/projection.py used to pin it. Test infrastructure only.
"""
import itertools
import numpy as np

from . import grid
from .fields import Field


def _outer_freq(xi, Nred, m, n=None):
    """Broadcast xi_m (or xi_m*xi_n) to the full grid Nred."""
    d = len(Nred)
    sm = [1] * d
    sm[m] = Nred[m]
    a = np.broadcast_to(xi[m].reshape(sm), tuple(Nred))
    if n is None:
        return np.array(a)
    sn = [1] * d
    sn[n] = Nred[n]
    return a * np.broadcast_to(xi[n].reshape(sn), tuple(Nred))


def _to_r_form(tensors, N, NyqNul):
    """Nyquist-null through the odd grid, then convert to the 'r' half spectrum
    (projections.py:101-111 and 256-265)."""
    out = []
    for t in tensors:
        if NyqNul:
            t = t.enlarge(N)
        t.set_fft_form('r')
        t.val = t.val / np.prod(t.N)
        out.append(t)
    return out


def scalar(N, Y, NyqNul=True, fft_form='r'):
    """G0 (mean), G1 (curl-free zero-mean), G2 (divergence-free zero-mean) for the
    scalar problem (projections.py:9-112)."""
    want_r = fft_form == 'r'
    form = 0 if want_r else fft_form
    d = np.size(N)
    N = np.asarray(N, dtype=np.int64)
    Nred = grid.odd_grid(N) if NyqNul else N
    xi = grid.frequencies(Nred, Y, fft_form=form)
    shp = (d, d) + tuple(Nred)
    G0, G1, G2 = np.zeros(shp), np.zeros(shp), np.zeros(shp)
    num = np.zeros(shp)
    denom = np.zeros(tuple(Nred))
    ic = grid.mean_index(Nred, fft_form=form)
    for m in range(d):
        num[m, m] = _outer_freq(xi, Nred, m) ** 2
        denom = denom + num[m, m]
        G0[(m, m) + ic] = 1.
    for m in range(d):
        for n in range(m + 1, d):
            num[m, n] = _outer_freq(xi, Nred, m, n)
    denom[ic] = 1.
    for m in range(d):
        for n in range(m, d):
            G1[m, n] = num[m, n] / denom
            G2[m, n] = float(m == n) - G1[m, n]
            G2[(m, n) + ic] = 0.
    for m in range(1, d):
        for n in range(m):
            G1[m, n] = G1[n, m]
            G2[m, n] = G2[n, m]
    ts = [Field(val=v, order=2, N=N, multype=21, Fourier=True, fft_form=form, name=nm)
          for v, nm in ((G0, 'hG0'), (G1, 'hG1'), (G2, 'hG2'))]
    if want_r:
        return tuple(_to_r_form(ts, tuple(N), NyqNul))
    return tuple(t.enlarge(tuple(N)) if NyqNul else t for t in ts)


def _others(i, d):
    return [k for k in range(d) if k != i]


def elasticity(N, Y, NyqNul=True, fft_form='r'):
    """Mandel-notation projections G0, G1h, G1s, G2h, G2s for small-strain elasticity
    (projections.py:114-267). G1h+G1s projects onto compatible strains, G2h+G2s onto
    zero-mean equilibrated stresses."""
    want_r = fft_form == 'r'
    form = 0 if want_r else fft_form
    N = np.asarray(N, dtype=np.int64)
    xi = grid.frequencies(N, Y, fft_form=form)
    d = N.size
    D = d * (d + 1) // 2
    Nred = grid.odd_grid(N) if NyqNul else N
    if NyqNul and not np.array_equal(Nred, N):
        xi = grid.frequencies(Nred, Y, fft_form=form)
    Ng = tuple(Nred)
    num = np.zeros((d, d) + Ng)
    n2 = np.zeros(Ng)
    for m in range(d):
        num[m, m] = _outer_freq(xi, Nred, m) ** 2
        n2 += num[m, m]
    n4 = n2 ** 2
    ic = grid.mean_index(Nred, fft_form=form)
    n2[ic] = 1.
    n4[ic] = 1.
    for m in range(d):
        for n in range(m + 1, d):
            num[m, n] = _outer_freq(xi, Nred, m, n)
    z = lambda: np.zeros((D, D) + Ng)
    G1h, G1s, IS0, mean, Lamh, S, W, WT = z(), z(), z(), z(), z(), z(), z(), z()
    for m in range(d):
        S[m, m] = 2 * num[m, m] / n2
        for n in range(d):
            G1h[m, n] = num[m, m] * num[n, n] / n4
            Lamh[m, n] = 1. / d
            Lamh[(m, n) + ic] = 0.
    for m in range(D):
        IS0[m, m] = 1.
        IS0[(m, m) + ic] = 0.
        mean[(m, m) + ic] = 1.
    r2 = 2 ** 0.5
    if d == 2:
        S[0, 2] = r2 * num[0, 1] / n2
        S[1, 2] = r2 * num[0, 1] / n2
        S[2, 2] = 1.
        S[(2, 2) + ic] = 0.
        G1h[0, 2] = r2 * num[0, 0] * num[0, 1] / n4
        G1h[1, 2] = r2 * num[0, 1] * num[1, 1] / n4
        G1h[2, 2] = 2 * num[0, 0] * num[1, 1] / n4
        for m in range(d):
            for n in range(d):
                W[m, n] = num[m, m] / n2
            W[2, m] = r2 * num[0, 1] / n2
    elif d == 3:
        for m in range(d):
            S[m + 3, m + 3] = 1 - num[m, m] / n2
            S[(m + 3, m + 3) + ic] = 0.
        for m in range(d):
            for n in range(m + 1, d):
                S[m + 3, n + 3] = num[m, n] / n2
                G1h[m + 3, n + 3] = num[m, m] * num[n, n] / n4
        for m in range(d):
            for n in range(d):
                i0, i1 = _others(n, d)
                S[m, n + 3] = float(m != n) * r2 * num[i0, i1] / n2
                G1h[m, n + 3] = r2 * num[m, m] * num[i0, i1] / n4
                W[m, n] = num[m, m] / n2
                W[n + 3, m] = r2 * num[i0, i1] / n2
        for m in range(d):
            for n in range(d):
                a0, a1 = _others(m, d)
                b0, b1 = _others(n, d)
                G1h[m + 3, n + 3] = 2 * num[a0, a1] * num[b0, b1] / n4
    for n in range(D):
        for m in range(n + 1, D):
            S[m, n] = S[n, m]
            G1h[m, n] = G1h[n, m]
    for m in range(D):
        for n in range(D):
            G1s[m, n] = S[m, n] - 2 * G1h[m, n]
            WT[m, n] = W[n, m]
    G2h = (d * Lamh + G1h - W - WT) / (d - 1)
    G2s = IS0 - G1h - G1s - G2h
    ts = [Field(val=v, order=2, N=N, multype=21, Fourier=True, fft_form=form, name=nm)
          for v, nm in ((mean, 'hG0'), (G1h, 'hG1h'), (G1s, 'hG1s'), (G2h, 'hG2h'), (G2s, 'hG2s'))]
    if want_r:
        return tuple(_to_r_form(ts, tuple(N), NyqNul))
    return tuple(t.enlarge(tuple(N)) if NyqNul else t for t in ts)


# ---- closed forms (ffthompy/tensors/projection.py) used to pin the above ----

def _freq_grid(N, Y, fft_form):
    xi = grid.frequencies(N, Y, fft_form=fft_form)
    return np.stack(np.meshgrid(*xi, indexing='ij'))


def scalar_closed_form(N, Y, fft_form='r'):
    """G1 = xi xi^T / |xi|^2, G2 = I - G1 - G0 (tensors/projection.py:6-31)."""
    N = np.asarray(N)
    q = _freq_grid(N, Y, fft_form)
    kok = np.einsum('i...,j...->ij...', q, q)
    k2 = np.einsum('i...,i...', q, q)
    ic = grid.mean_index(N, fft_form)
    k2[ic] = 1.
    d = N.size
    G0 = np.zeros_like(kok)
    I = np.zeros_like(kok)
    for i in range(d):
        G0[(i, i) + ic] = 1.
        I[i, i] = 1.
    G1 = kok / k2
    return G0, G1, I - G1 - G0


def gamma_small_strain(N, Y, fft_form='r'):
    """4th-order Gamma = sym(delta xi xi)/|xi|^2 - xi xi xi xi/|xi|^4
    (tensors/projection.py:33-51), zero at xi = 0."""
    N = np.asarray(N)
    q = _freq_grid(N, Y, fft_form)
    d = N.size
    qq = np.einsum('i...,i...', q, q)
    nz = qq != 0
    qs = np.where(nz, qq, 1.)
    delta = np.eye(d)
    G = np.zeros((d,) * 4 + q.shape[1:])
    for i, j, k, l in itertools.product(range(d), repeat=4):
        G[i, j, k, l] = (-q[i] * q[j] * q[k] * q[l] / qs ** 2
                         + 0.5 * (delta[i, k] * q[j] * q[l] + delta[i, l] * q[j] * q[k]
                                  + delta[j, k] * q[i] * q[l] + delta[j, l] * q[i] * q[k]) / qs)
        G[i, j, k, l][~nz] = 0.
    return G


def mandel_matrix_of(T4):
    """Mandel 6x6 (or 3x3) matrix of a minor-symmetric 4th-order grid tensor
    (mechanics/matcoef.py:151-207, ndim=4)."""
    d = T4.shape[0]
    D = d * (d + 1) // 2
    pairs = [(i, i) for i in range(d)]
    if d == 3:
        pairs += [(1, 2), (0, 2), (0, 1)]
    else:
        pairs += [(0, 1)]
    w = [1.] * d + [2 ** 0.5] * (D - d)
    out = np.zeros((D, D) + T4.shape[4:], dtype=T4.dtype)
    for a, (i, j) in enumerate(pairs):
        for b, (k, l) in enumerate(pairs):
            out[a, b] = w[a] * w[b] * T4[i, j, k, l]
    return out
