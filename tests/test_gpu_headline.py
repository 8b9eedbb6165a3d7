"""The headline workload at its full size against the float64 oracle: 3-D scalar 243^3, 30 seeded
spheres of conductivity 100 in a matrix of 1, unit loads e1-e3, primal and dual GaNi solves to the
absolute tolerance 1e-8. The oracle's results (tests/golden/oracle_scalar243.json) were computed on
the CPU by tests/golden/make_oracle_scalar243.py with the oracle's own operators.

* The microstructure is the one the fixture was made from (coefficient sum, volume fraction).
* fp64: CG iteration counts equal the oracle's for every load; the bound matrices agree to 1e-10.
* mixed (fp32 CG + fp64 reliable updates, each fp32 correction projected in fp64): the counts exceed
  the oracle's by at most 5 (the iteration restarts its search direction after a projected update),
  bounds within 1e-10.
* lower <= upper holds within the rounding slack 64 eps max|AH| in both modes.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
nat = pytest.importorskip('ffthompy_amd._native')
from ffthompy_amd import microstructure as ms  # noqa: E402

FIX = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'oracle_scalar243.json')))
EPS = np.finfo(np.float64).eps


def test_scalar243_spheres_matches_fp64_oracle():
    N = tuple(FIX['N'])
    print('case headline 243', flush=True)
    a, _, _ = ms.spheres_conductivity(N, count=FIX['spheres'], contrast=FIX['contrast'], seed=FIX['seed'])
    assert float(a.sum()) == FIX['coef_sum'] and float((a > 1).mean()) == FIX['volume_fraction']
    plans = {}
    for pd, coef in (('primal', a), ('dual', 1. / a)):
        pl = nat.Plan(N, np.ones(3), 'scalar', pd)
        pl.set_coeff('iso', coef)
        plans[pd] = pl
    EN = nat.DeviceArray((3,) + N)
    B = nat.DeviceArray((3,) + N)
    sols = [nat.DeviceArray((3,) + N) for _ in range(3)]
    for mode, tol in (('fp64', 1e-10), ('mixed', 1e-10)):
        AH = {}
        for pd, pl in plans.items():
            kits = []
            for iL in range(3):
                E = np.zeros(3)
                E[iL] = 1.
                nat.call('fh_memset_zero', EN.ptr, EN.nbytes)
                nat.add_constant(EN, -E)
                pl.matvec(EN, B)
                x = sols[iL]
                nat.call('fh_memset_zero', x.ptr, x.nbytes)
                info = pl.cg_solve(B, x, tol=FIX['tol'], maxiter=2000, mode=mode, delta=1e-4)
                assert info['status'] == nat.FH_CG_OK and info['norm_res'] <= FIX['tol']
                kits.append(info['kit'])
                nat.add_constant(x, E)
            M = pl.energy_matrix(sols)
            AH[pd] = M if pd == 'primal' else np.linalg.inv(M)
            ref = np.asarray(FIX['AH_' + pd])
            rel = np.abs(AH[pd] - ref).max() / np.abs(ref).max()
            print('%s %s kits %s (oracle %s) rel diff %.2e' % (mode, pd, kits, FIX['kit_' + pd], rel), flush=True)
            if mode == 'fp64':
                assert kits == FIX['kit_' + pd], (mode, pd)
            else:
                assert all(0 <= k - r <= 5 for k, r in zip(kits, FIX['kit_' + pd])), (mode, pd)
            assert rel < tol, (mode, pd, rel)
        up, lo = AH['primal'], AH['dual']
        slack = 64 * EPS * np.abs(up).max()
        assert np.linalg.eigvalsh(0.5 * (up - lo + (up - lo).T)).min() >= -slack, mode
