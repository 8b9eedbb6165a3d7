#!/usr/bin/env python3
"""Writes tests/golden/oracle_scalar243.json: the oracle's float64 GaNi solve of the headline
workload, to pin the GPU solve at its full size.

Workload: 3-D scalar conductivity on a 243^3 grid, 30 non-overlapping spheres (seed 20261202,
radii U[0.08, 0.15]) of conductivity 100 in a matrix of 1, unit loads e1, e2, e3, primal (G1, a)
and dual (G2, 1/a) problems, CG to the absolute tolerance 1e-8 on the grid-mean residual norm.
The solve is the oracle's restatement of ffthompy/applications.py:24-90 step by step: the
materialised projections of oracle.projections.scalar (ffthompy/projections.py:9-112), the
(d, d)+N coefficient applied by einsum, Operator([[FiN, hG, FN], A]), the oracle's CG
(ffthompy/general/solver.py:80-139) from x0 = 0 with B = Afun(-E), and assembly_matrix
(ffthompy/postprocess.py:53-70; the dual matrix inverted, postprocess.py:43-46). The FFTs run
through scipy.fft with every core of this machine (pocketfft, as numpy.fft; same values to
rounding); the reference's CallBack matvec is left out (it does not change the iterates).

CPU only, about 15 minutes on 8 cores and 12 GB of memory. Run from the repository root:
    python tests/golden/make_oracle_scalar243.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle.fields as fields  # noqa: E402
from oracle import projections as proj  # noqa: E402
from oracle.fields import DFT, Field, Operator  # noqa: E402
from oracle.homogenize import add_macro2minimizer, assembly_matrix  # noqa: E402
from oracle.solver import linear_solver  # noqa: E402
from ffthompy_amd import microstructure as ms  # noqa: E402  (host-side input generator)

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'oracle_scalar243.json')


def log(*a):
    print('[oracle243]', *a, flush=True)


def main(n=243, tol=1e-8):
    fields.FFT_WORKERS = len(os.sched_getaffinity(0))
    N = (n, n, n)
    Y = np.ones(3)
    t0 = time.perf_counter()
    a, centres, radii = ms.spheres_conductivity(N, count=30, contrast=100., seed=ms.BASE_SEED)
    log('microstructure: volume fraction %.6f' % float((a > 1).mean()))
    _, hG1, hG2 = proj.scalar(N, Y, NyqNul=True)
    log('projections built (%.0f s)' % (time.perf_counter() - t0))
    FN = DFT(N, inverse=False)
    FiN = DFT(N, inverse=True)
    out = {
        '_about': ('float64 oracle solve of the 243^3 sphere workload (tests/golden/make_oracle_scalar243.py): '
                   'AH upper (primal) and lower (dual, inverted) bound matrices of the GaNi solve, CG iterations '
                   'and final residual norms per unit load'),
        'N': list(N), 'seed': ms.BASE_SEED, 'spheres': 30, 'contrast': 100.0, 'tol': tol,
        'volume_fraction': float((a > 1).mean()), 'coef_sum': float(a.sum()),
        'fft': 'scipy.fft (pocketfft), workers=%d' % fields.FFT_WORKERS,
    }
    for pd, hG, coef in (('primal', hG1, a), ('dual', hG2, 1. / a)):
        A = Field(val=np.einsum('ij,...->ij...', np.eye(3), coef), order=2, multype=21, name='A_GaNi')
        GN = Operator([[FiN, hG, FN]], name='G')
        Afun = Operator([[GN, A]], name='FiGFA')
        sols, kits, norms = [], [], []
        for iL in range(3):
            E = np.zeros(3)
            E[iL] = 1.
            EN = Field(shape=(3,), N=N, Y=Y)
            EN.set_mean(E)
            x0 = Field(shape=(3,), N=N, Y=Y)
            B = Afun(-EN)
            t = time.perf_counter()
            X, info = linear_solver(Afun, B, x0=x0, par={'tol': tol, 'maxiter': 1000}, solver='CG')
            kits.append(int(info['kit']))
            norms.append(float(info['norm_res']))
            log('%s load %d: kit %d, norm_res %.3e (%.0f s)' % (pd, iL, info['kit'], info['norm_res'],
                                                                time.perf_counter() - t))
            sols.append(add_macro2minimizer(X, E))
            del B, X
        AH = assembly_matrix(A, sols)
        if pd == 'dual':
            AH = np.linalg.inv(AH)
        out['AH_' + pd] = AH.tolist()
        out['kit_' + pd] = kits
        out['norm_res_' + pd] = norms
        log('AH_%s diag %s' % (pd, np.diag(AH)))
        del sols, A, Afun
    out['cpu_seconds'] = round(time.perf_counter() - t0, 1)
    with open(OUT, 'w') as f:
        json.dump(out, f, indent=1)
    log('wrote', OUT)


if __name__ == '__main__':
    main()
