"""Hand-written backward of the homogenized matrix (fh_energy_gradient behind
torch.autograd.Function) against central finite differences and a float64 PyTorch
reference that differentiates through its own CG iterations."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')
nat = pytest.importorskip('ffthompy_amd._native')
from ffthompy_amd.autograd import upper_bound  # noqa: E402
from ffthompy_amd.microstructure import gaussian_random_field, BASE_SEED  # noqa: E402
from torch_reference import upper_bound_torch  # noqa: E402


def loss_of(AH):
    return torch.sum((AH - 2. * torch.eye(AH.shape[0], dtype=AH.dtype)) ** 2)


def test_gradient_matches_torch_reference_15():
    """15^3, GRF theta (correlation 0.15, std 0.8): loss ||AH - 2I||_F^2."""
    print('case autograd torch-reference 15^3', flush=True)
    th = gaussian_random_field((15, 15, 15), 0.15, 0.8, BASE_SEED + 5)
    t1 = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    L1 = loss_of(upper_bound(t1, tol=1e-12))
    L1.backward()
    t2 = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    L2 = loss_of(upper_bound_torch(t2, tol=1e-13))
    L2.backward()
    assert abs(L1.item() - L2.item()) < 1e-12 * abs(L2.item())
    g1, g2 = t1.grad.numpy(), t2.grad.numpy()
    assert np.abs(g1 - g2).max() < 1e-8 * np.abs(g2).max(), np.abs(g1 - g2).max() / np.abs(g2).max()


def test_gradient_matches_central_differences_99():
    """99^3, GRF theta (correlation 0.15, std 0.8): 20 seeded voxels, step 1e-6, fp64."""
    print('case autograd FD 99^3', flush=True)
    N = (99, 99, 99)
    th = gaussian_random_field(N, 0.15, 0.8, BASE_SEED + 5)
    t = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    L = loss_of(upper_bound(t, tol=1e-11))
    L.backward()
    g = t.grad.numpy()
    rng = np.random.default_rng(BASE_SEED)
    vox = [tuple(rng.integers(0, 99, 3)) for _ in range(20)]
    h = 1e-6
    L0 = L.item()
    # Rounding floor of a central difference of FP64 loss values: one evaluation of L
    # (fields from a CG to 1e-11, then 9 AH entries) reproduces to some tens of ulps of
    # |L| (printed below: L at tolerance 1e-11 against 1e-12), divided by 2h. At h = 1e-6
    # and |grad| ~ 5e-6 that is ~1e-3 of the gradient: the FD check resolves no better.
    with torch.no_grad():
        L12 = loss_of(upper_bound(torch.tensor(th), tol=1e-12)).item()
    print('L %.17g, L(tol 1e-12) - L = %.2e (%.1f ulp)' % (L0, L12 - L0, (L12 - L0) / np.spacing(L0)))
    floor = 64 * np.finfo(np.float64).eps * abs(L0) / (2 * h)
    rows = []
    for v in vox:
        vals = []
        for s in (+1, -1):
            tt = th.copy()
            tt[v] += s * h
            with torch.no_grad():
                vals.append(loss_of(upper_bound(torch.tensor(tt), tol=1e-11)).item())
        fd = (vals[0] - vals[1]) / (2 * h)
        rows.append((v, g[v], fd, abs(fd - g[v])))
    for v, gv, fd, err in rows:
        print('voxel %s grad %.6e fd %.6e |diff| %.2e (floor %.2e)' % (v, gv, fd, err, floor), flush=True)
    rel = [err / abs(gv) for _, gv, _, err in rows]
    print('FD relative errors: max %.2e median %.2e' % (max(rel), float(np.median(rel))), flush=True)
    assert all(err <= 1e-5 * abs(gv) + floor for _, gv, _, err in rows)
    assert float(np.median(rel)) < 1e-3


def test_directional_derivative_richardson_99():
    """99^3: the directional derivative <grad L, phi> along a random field phi (GRF, correlation
    0.15, std 1) against central differences of L(theta + t phi) with Richardson extrapolation
    over steps h = 2e-3 and h/2 (truncation O(h^4)). Summed over all 970299 voxels the derivative
    is large against the rounding floor of the loss differences (~2e-11 of it), which bounds the
    single-voxel check above to ~1e-3; measured agreement 4e-13 relative, D(h) and D(h/2) off by
    1.2e-6 and 3.1e-7 (second order)."""
    print('case autograd directional FD 99^3', flush=True)
    N = (99, 99, 99)
    th = gaussian_random_field(N, 0.15, 0.8, BASE_SEED + 5)
    phi = gaussian_random_field(N, 0.15, 1.0, BASE_SEED + 13)
    t = torch.tensor(th, dtype=torch.float64, requires_grad=True)
    L = loss_of(upper_bound(t, tol=1e-12))
    L.backward()
    D = float(np.sum(t.grad.numpy() * phi))

    def Lat(s):
        with torch.no_grad():
            return loss_of(upper_bound(torch.tensor(th + s * phi), tol=1e-12)).item()

    h = 2e-3
    Dh = (Lat(h) - Lat(-h)) / (2 * h)
    Dh2 = (Lat(h / 2) - Lat(-h / 2)) / h
    DR = (4 * Dh2 - Dh) / 3
    # rounding floor of DR: some tens of ulps of |L| per loss value, through the two quotients
    floor = 64 * np.finfo(np.float64).eps * abs(L.item()) * (4 / h + 1 / (2 * h)) * 2 / 3
    print('L %.17g  <g,phi> %.15e  D(h) %.15e  D(h/2) %.15e  Richardson %.15e  rel %.2e  floor rel %.2e'
          % (L.item(), D, Dh, Dh2, DR, abs(DR - D) / abs(D), floor / abs(D)), flush=True)
    # the central differences converge at second order towards the gradient's value
    assert abs(Dh2 - D) < 0.5 * abs(Dh - D) + floor
    assert abs(DR - D) <= 1e-10 * abs(D) + floor


def test_local_field_gradient_adjoint_15():
    """Loss on the solution fields themselves (needs the adjoint CG): L = sum w (e_0)_0^2 +
    sum u . e_2 on 15^3, gradient against the float64 torch reference (autograd through its
    CG) and central differences at 5 voxels."""
    from ffthompy_amd.autograd import solution_fields
    from torch_reference import fields_torch
    print('case autograd local fields 15^3', flush=True)
    N = (15, 15, 15)
    th = gaussian_random_field(N, 0.15, 0.8, BASE_SEED + 9)
    rng = np.random.default_rng(9)
    w = torch.tensor(rng.random(N))
    u = torch.tensor(rng.standard_normal((3,) + N))

    def loss(e):
        return torch.sum(w * e[0][0] ** 2) + torch.sum(u * e[2])

    t1 = torch.tensor(th, requires_grad=True)
    L1 = loss(solution_fields(t1, tol=1e-12))
    L1.backward()
    t2 = torch.tensor(th, requires_grad=True)
    L2 = loss(fields_torch(t2, tol=1e-13))
    L2.backward()
    assert abs(L1.item() - L2.item()) < 1e-10 * abs(L2.item())
    g1, g2 = t1.grad.numpy(), t2.grad.numpy()
    assert np.abs(g1 - g2).max() < 1e-8 * np.abs(g2).max(), np.abs(g1 - g2).max() / np.abs(g2).max()
    h = 1e-6
    for v in [(0, 0, 0), (3, 7, 11), (14, 2, 5), (8, 8, 8), (1, 13, 6)]:
        vals = []
        for s in (1, -1):
            tt = th.copy()
            tt[v] += s * h
            with torch.no_grad():
                vals.append(loss(solution_fields(torch.tensor(tt), tol=1e-12)).item())
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g1[v]) <= 1e-5 * abs(g1[v]) + 1e-7 * np.abs(g1).max(), (v, fd, g1[v])
