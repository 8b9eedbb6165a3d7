"""Synthetic microstructures for the benchmark configurations (host side, seeded).

Coefficients are returned on the grid in origin-0 voxel order, the order the
reference reaches after Tensor.shift() (ffthompy/materials.py:119-123,
ffthompy/tensors/objects.py:169-186). Voxel i on an axis of N points sits at
x = Y*k/N with k = i (i <= (N-1)/2) or i - N, the centred node set of
ffthompy/trigpol.py:55-71; distances are periodic minimum-image, and membership is
strict (distance < radius) as in ffthompy/materials.py:270.
"""
import numpy as np

BASE_SEED = 20261202


def _axis_coords(n, Y):
    i = np.arange(n)
    k = np.where(2 * i <= n - 1, i, i - n)
    return Y * k / n


def _wrap(d, Y):
    return d - Y * np.round(d / Y)


def random_spheres(count=30, rmin=0.08, rmax=0.15, seed=BASE_SEED, max_tries=200000):
    """Non-overlapping spheres in the unit periodic cube by rejection sampling:
    centres U[0,1)^3, radii U[rmin,rmax], periodic centre distance >= r_i + r_j."""
    rng = np.random.default_rng(seed)
    centres, radii = [], []
    tries = 0
    while len(centres) < count:
        tries += 1
        if tries > max_tries:
            raise RuntimeError('could not place %d spheres' % count)
        c = rng.random(3)
        r = rng.uniform(rmin, rmax)
        ok = True
        for c2, r2 in zip(centres, radii):
            d = _wrap(c - c2, 1.)
            if np.sqrt(np.sum(d * d)) < r + r2:
                ok = False
                break
        if ok:
            centres.append(c)
            radii.append(r)
    return np.array(centres), np.array(radii)


def sphere_indicator(N, centres, radii, Y=(1., 1., 1.)):
    """uint8 indicator of the union of balls on an N grid (origin-0 order)."""
    N = tuple(int(n) for n in N)
    ind = np.zeros(N, dtype=np.uint8)
    coords = [_axis_coords(N[a], Y[a]) for a in range(3)]
    for c, r in zip(centres, radii):
        sl = []
        d2 = []
        for a in range(3):
            d = _wrap(coords[a] - c[a], Y[a])
            idx = np.nonzero(np.abs(d) < r)[0]
            sl.append(idx)
            d2.append(d[idx] ** 2)
        if any(len(s) == 0 for s in sl):
            continue
        inside = (d2[0][:, None, None] + d2[1][None, :, None] + d2[2][None, None, :]) < r * r
        sub = ind[np.ix_(*sl)]
        sub[inside] = 1
        ind[np.ix_(*sl)] = sub
    return ind


def spheres_conductivity(N, count=30, contrast=100., seed=BASE_SEED):
    """a(x) = contrast inside the spheres, 1 outside; also returns the sphere set."""
    centres, radii = random_spheres(count=count, seed=seed)
    ind = sphere_indicator(N, centres, radii)
    a = np.where(ind == 1, contrast, 1.)
    return a, centres, radii


def random_rotation(rng):
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def ellipsoid_indicator(N, target_fraction=0.25, amin=0.06, amax=0.16, seed=BASE_SEED, max_count=10000):
    """Union of seeded random ellipsoids (semi-axes U[amin,amax], uniform random
    rotation, centres U[0,1)^3) added until the voxel volume fraction first reaches
    target_fraction."""
    N = tuple(int(n) for n in N)
    rng = np.random.default_rng(seed)
    ind = np.zeros(N, dtype=np.uint8)
    coords = [_axis_coords(N[a], 1.) for a in range(3)]
    count = 0
    while ind.mean() < target_fraction:
        count += 1
        if count > max_count:
            raise RuntimeError('volume fraction not reached')
        c = rng.random(3)
        ax = rng.uniform(amin, amax, 3)
        R = random_rotation(rng)
        M = R @ np.diag(1. / ax ** 2) @ R.T
        rb = ax.max()
        sl, ds = [], []
        for a in range(3):
            d = _wrap(coords[a] - c[a], 1.)
            idx = np.nonzero(np.abs(d) < rb)[0]
            sl.append(idx)
            ds.append(d[idx])
        X = ds[0][:, None, None]
        Yc = ds[1][None, :, None]
        Z = ds[2][None, None, :]
        q = (M[0, 0] * X * X + M[1, 1] * Yc * Yc + M[2, 2] * Z * Z
             + 2 * (M[0, 1] * X * Yc + M[0, 2] * X * Z + M[1, 2] * Yc * Z))
        sub = ind[np.ix_(*sl)]
        sub[q < 1.] = 1
        ind[np.ix_(*sl)] = sub
    return ind, count


def lame_from_young(E, nu):
    """(lambda, mu) of an isotropic phase from Young's modulus and Poisson's ratio
    (K = E/(3(1-2nu)), mu = E/(2(1+nu)), lambda = K - 2mu/3 as in
    ffthompy/mechanics/matcoef.py:12-71)."""
    K = E / (3. * (1. - 2. * nu))
    mu = E / (2. * (1. + nu))
    return K - 2. * mu / 3., mu


def gaussian_random_field(N, corr_len, std, seed, Y=None):
    """Periodic Gaussian random field by spectral synthesis, squared-exponential
    covariance C(r) = std^2 exp(-r^2 / (2 corr_len^2)): white noise filtered by
    exp(-pi^2 corr_len^2 |xi|^2) (the square root of the covariance's spectral
    density), then scaled so that the pointwise variance is std^2."""
    N = tuple(int(n) for n in N)
    Y = np.ones(len(N)) if Y is None else np.asarray(Y, dtype=np.float64)
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(N)
    xi2 = np.zeros(N)
    for a, n in enumerate(N):
        shp = [1] * len(N)
        shp[a] = n
        xi2 = xi2 + (np.fft.fftfreq(n, d=1. / n) / Y[a]).reshape(shp) ** 2
    g = np.exp(-np.pi ** 2 * corr_len ** 2 * xi2)
    g *= std / np.sqrt(np.mean(g ** 2))
    return np.real(np.fft.ifftn(np.fft.fftn(w) * g))


def uq_samples(count=512, lengths=(63, 81, 105, 135), corr_len=0.1, std=1.0, seed=BASE_SEED):
    """Batched uncertainty-quantification workload: `count` 2-D samples whose grid axes are
    drawn uniformly from `lengths`; each conductivity is exp of a periodic Gaussian random
    field (squared-exponential covariance, correlation length corr_len, log-std std) on a
    unit cell. Returns (dims, fields): dims (count, 2) int, fields a list of (n0, n1) arrays."""
    rng = np.random.default_rng(seed)
    dims = rng.choice(np.asarray(lengths), size=(count, 2))
    fields = [np.exp(gaussian_random_field(tuple(d), corr_len, std, seed + 1 + s)) for s, d in enumerate(dims)]
    return dims, fields


def gaussian_random_field_device(N, corr_len, std, seed, Y=None, exponentiate=False):
    """gaussian_random_field with the spectral synthesis on the device (fh_grf): the same white
    noise (numpy default_rng(seed)), filtered and transformed by the HIP FFT passes; returns a
    device array of shape N (exp of the field if exponentiate)."""
    from . import _native as nat
    N = tuple(int(n) for n in N)
    Y = np.ones(len(N)) if Y is None else np.asarray(Y, dtype=np.float64)
    w = np.random.default_rng(seed).standard_normal(N)
    plan = nat.Plan(N, Y)
    return plan.grf(nat.DeviceArray.from_host(w), corr_len, std, exponentiate=exponentiate)
