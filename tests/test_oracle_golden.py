"""The CPU oracle reproduces every golden result of the reference's example problems.

Goldens: tests/golden/reference_results.json (from test_results/python3/*), compared
at the reference's own tolerance of 1e-9 absolute (run_unittests.py:57-64), plus the
golden CG iteration counts per load.
"""
import json
import os

import numpy as np
import pytest

from oracle import examples
from oracle.homogenize import Problem

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), 'golden', 'reference_results.json')))
TOL = GOLD['tolerance_abs']
CASES = [(f, n, c, m) for f, n, c, m in examples.problems()]


@pytest.mark.parametrize('case', CASES, ids=['%s_%s' % (c[0], c[1]) for c in CASES])
def test_example_problem_matches_golden(case):
    f, name, conf, mats = case
    gold = GOLD['problems']['%s_%s' % (f, name)]
    pb = Problem(dict(conf), mats())
    pb.calculate()
    mats_out = {}
    for pd in pb.solve['primaldual']:
        mats_out.update(pb.output['mat_' + pd])
        kits = [r['info']['kit'] for r in pb.output['res_' + pd]]
        assert kits == gold['kit'][pd], (pd, kits)
    for key, val in gold.get('diag', {}).items():
        AH = mats_out[key]
        assert np.abs(AH - val * np.eye(AH.shape[0])).max() < TOL, key
    for key, entries in gold.get('entries', {}).items():
        AH = mats_out[key]
        for ij, val in entries.items():
            i, j = map(int, ij.split(','))
            assert abs(AH[i, j] - val) < TOL, (key, ij)
        # cubic symmetry of the cube inclusion: the normal block is c11 on the
        # diagonal and c12 off it
        assert np.abs(AH[:3, :3] - (AH[0, 1] * np.ones((3, 3)) + (AH[0, 0] - AH[0, 1]) * np.eye(3))).max() < 1e-12


def test_planted_sign_error_is_caught(monkeypatch):
    """A plausible mistake in the oracle (dual projection G2 = I + G1 instead of I - G1)
    must fail the golden comparison: the goldens pin the projection."""
    from oracle import projections
    orig = projections.scalar

    def broken(N, Y, NyqNul=True, fft_form='r'):
        G0, G1, G2 = orig(N, Y, NyqNul, fft_form)
        G2.val = G2.val + 2 * G1.val
        return G0, G1, G2

    monkeypatch.setattr(projections, 'scalar', broken)
    f, name, conf, mats = CASES[0]
    pb = Problem(dict(conf), mats())
    with np.errstate(all='ignore'):
        try:
            pb.calculate()
            AH = pb.output['mat_dual']['AH_GaNi_dual']
            ok = abs(AH[0, 0] - GOLD['problems']['scalar_2d_prob1']['diag']['AH_GaNi_dual']) < TOL
        except (np.linalg.LinAlgError, ValueError, FloatingPointError):
            ok = False
    assert not ok
