"""Grid maths: index sets, frequencies, coordinates, spectral pad/crop, pointwise inverse.

Restates ffthompy/trigpol.py. Test infrastructure only (see oracle/__init__.py).
"""
import numpy as np

FFT_FORM_DEFAULT = 'r'   # ffthompy/trigpol.py:6


def index_set(N, fft_form=FFT_FORM_DEFAULT):
    """Integer frequencies k with -N/2 <= k < N/2 per axis (ffthompy/trigpol.py:11-24).

    For fft_form 'r' or 0 the sets are returned in ifftshift order (0 first),
    for 'c' in centred (ascending) order.
    """
    N = np.atleast_1d(np.asarray(N, dtype=np.int64))
    out = []
    for n in N:
        k = np.arange(np.fix(-n / 2.), np.fix(n / 2. + 0.5), dtype=np.int64)
        out.append(np.fft.ifftshift(k) if fft_form in ('r', 0) else k)
    return out


def frequencies(N, Y, fft_form=FFT_FORM_DEFAULT):
    """Frequencies xi_m = k_m / Y_m (ffthompy/trigpol.py:26-40).

    The 'r' form keeps only the first N[-1]//2+1 entries of the last axis.
    """
    N = np.atleast_1d(np.asarray(N))
    xis = []
    for m in range(N.size):
        xis.append(np.arange(np.fix(-N[m] / 2.), np.fix(N[m] / 2. + 0.5)) / Y[m])
    if fft_form == 'r':
        xis = [np.fft.ifftshift(x) for x in xis]
        xis[-1] = xis[-1][:int(np.fix(N[-1] / 2) + 1)]
    elif fft_form == 0:
        xis = [np.fft.ifftshift(x) for x in xis]
    return xis


def coordinates(N, Y):
    """Nodal coordinates x = Y*k/N with centred k (ffthompy/trigpol.py:55-71)."""
    N = np.atleast_1d(np.asarray(N, dtype=np.int64))
    d = N.size
    ks = index_set(N, fft_form='c')
    coord = np.zeros((d,) + tuple(N))
    for ii in range(d):
        shape = [1] * d
        shape[ii] = N[ii]
        coord[ii] = np.broadcast_to((Y[ii] * ks[ii] / N[ii]).reshape(shape), tuple(N))
    return coord


def pointwise_inverse(A):
    """Per-voxel inverse of a (d,d)+N array by Gauss-Jordan elimination without pivoting.

    Same elimination order as ffthompy/trigpol.py:120-159 (forward sweep that
    normalises row m and clears rows below, then a backward sweep).
    """
    B = np.array(A, dtype=np.float64, copy=True)
    d = B.shape[0]
    if B.shape[0] != B.shape[1]:
        raise NotImplementedError("Non-square matrix!")
    grid = B.shape[2:]
    inv = np.zeros_like(B)
    for m in range(d):
        inv[m, m] = 1.
    for m in range(d):
        piv = B[m, m].copy()
        B[m, m] = 1.
        for n in range(m + 1, d):
            B[m, n] = B[m, n] / piv
        for n in range(d):
            inv[m, n] = inv[m, n] / piv
        for k in range(m + 1, d):
            f = B[k, m].copy()
            for l in range(d):
                B[k, l] = B[k, l] - B[m, l] * f
                inv[k, l] = inv[k, l] - inv[m, l] * f
    for m in range(d - 1, -1, -1):
        for k in range(m - 1, -1, -1):
            f = B[k, m].copy()
            for l in range(d):
                B[k, l] = B[k, l] - B[m, l] * f
                inv[k, l] = inv[k, l] - inv[m, l] * f
    assert inv.shape[2:] == grid
    return inv


def enlarge_array(xN, M):
    """Centre-embed an array of (centred) Fourier coefficients into shape M by zeros
    (ffthompy/trigpol.py:162-189)."""
    M = np.asarray(M, dtype=np.float64)
    N = np.asarray(xN.shape, dtype=np.float64)
    if np.allclose(M, N):
        return xN
    ibeg = np.ceil((M - N) / 2).astype(np.int64)
    iend = np.ceil((M + N) / 2).astype(np.int64)
    xM = np.zeros(M.astype(np.int64), dtype=xN.dtype)
    xM[tuple(slice(b, e) for b, e in zip(ibeg, iend))] = xN
    return xM


def decrease_array(xN, M):
    """Crop the centred block of shape M (ffthompy/trigpol.py:191-214)."""
    M = np.asarray(M, dtype=np.float64)
    N = np.asarray(xN.shape, dtype=np.float64)
    ibeg = np.fix((N - M + (M % 2)) / 2).astype(np.int64)
    iend = np.fix((N + M + (M % 2)) / 2).astype(np.int64)
    return xN[tuple(slice(b, e) for b, e in zip(ibeg, iend))]


def odd_grid(N):
    """Largest odd grid not above N (ffthompy/trigpol.py:216-218)."""
    N = np.asarray(N)
    return N - ((N + 1) % 2)


def mean_index(N, fft_form=FFT_FORM_DEFAULT):
    """Index of the zero frequency (ffthompy/trigpol.py:220-224)."""
    N = np.atleast_1d(np.asarray(N))
    if fft_form in (0, 'r'):
        return tuple(0 for _ in N)
    return tuple(int(v) for v in np.fix(N / 2))
