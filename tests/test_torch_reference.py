"""The float64 PyTorch reference used by the autograd tests, pinned on the CPU: its
homogenized matrix equals the oracle's GaNi upper bound, and its autograd gradient
(through the CG iterations) equals central finite differences of itself."""
import numpy as np
import pytest

torch = pytest.importorskip('torch')
from torch_reference import upper_bound_torch  # noqa: E402
from oracle.homogenize import Problem  # noqa: E402


def oracle_upper(a):
    d = a.ndim
    conf = {'physics': 'scalar',
            'material': {'fun': lambda c: np.einsum('ij,...->ij...', np.eye(d), np.fft.fftshift(a)),
                         'Y': np.ones(d)},
            'solve': {'kind': 'GaNi', 'N': np.array(a.shape), 'primaldual': ['primal']},
            'postprocess': [{'kind': 'GaNi'}],
            'solver': {'kind': 'CG', 'tol': 1e-13, 'maxiter': 1e3}}
    pb = Problem(conf)
    pb.calculate()
    return pb.output['mat_primal']['AH_GaNi_primal']


@pytest.mark.parametrize('N', [(5, 5, 5), (7, 9)])
def test_matches_oracle(N):
    th = np.random.default_rng(3).standard_normal(N) * 0.5
    AH = upper_bound_torch(torch.tensor(th), tol=1e-13).numpy()
    assert np.abs(AH - oracle_upper(np.exp(th))).max() < 1e-13


def test_autograd_matches_finite_differences():
    th = np.random.default_rng(4).standard_normal((5, 5, 5)) * 0.5
    t = torch.tensor(th, requires_grad=True)
    L = torch.sum((upper_bound_torch(t, tol=1e-14) - 2. * torch.eye(3, dtype=torch.float64)) ** 2)
    L.backward()
    g = t.grad.numpy()
    h = 1e-5
    for v in [(0, 0, 0), (1, 2, 3), (4, 4, 1)]:
        vals = []
        for s in (1, -1):
            tt = th.copy()
            tt[v] += s * h
            AH = upper_bound_torch(torch.tensor(tt), tol=1e-14)
            vals.append(float(torch.sum((AH - 2. * torch.eye(3, dtype=torch.float64)) ** 2)))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g[v]) < 1e-6 * abs(g[v]), (v, fd, g[v])


def test_fields_match_oracle_solutions():
    """The reference fields (CG through torch) equal the oracle's GaNi primal solutions."""
    th = np.random.default_rng(6).standard_normal((5, 5, 5)) * 0.5
    from torch_reference import fields_torch
    e = [f.numpy() for f in fields_torch(torch.tensor(th), tol=1e-13)]
    a = np.exp(th)
    d = 3
    conf = {'physics': 'scalar',
            'material': {'fun': lambda c: np.einsum('ij,...->ij...', np.eye(d), np.fft.fftshift(a)),
                         'Y': np.ones(d)},
            'solve': {'kind': 'GaNi', 'N': np.array(a.shape), 'primaldual': ['primal']},
            'postprocess': [{'kind': 'GaNi'}],
            'solver': {'kind': 'CG', 'tol': 1e-13, 'maxiter': 1e3}}
    pb = Problem(conf)
    pb.calculate()
    sols = pb.output['sol_primal']
    for i in range(d):
        assert np.abs(e[i] - sols[i].val).max() < 1e-12
