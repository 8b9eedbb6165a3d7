"""Batched homogenization of many 2-D scalar samples on uneven grids, in one GPU launch.

The reference solves one Problem at a time (ffthompy/problem.py:77-117 calling
ffthompy/applications.py:24-90). Here every sample is one such scalar GaNi problem with an
isotropic voxel conductivity; all samples (primal and dual) are handed to the HIP library
together (fh_batch_*), solved by one persistent kernel with per-sample CG state on the
device, and the guaranteed bounds of ffthompy/postprocess.py:53-70 come back per sample.
"""
import numpy as np

from . import _native as nat


def homogenize_batch(fields, Y=None, precision='mixed', tol=1e-8, maxiter=10000, delta=1e-4, schedule=None):
    """Upper and lower bounds of the homogenized conductivity of every sample.

    fields: sequence of 2-D arrays a_s (positive, origin-0 voxel values, axis lengths 27, 45,
    63, 81, 105 or 135). Y: cell sizes, (2,) or (S, 2); unit cells by default.
    precision: 'mixed' (fp32 CG, fp64 refinement and energies) or 'fp64'.
    Returns a dict: 'upper', 'lower' (S, 2, 2); 'kit' (S, 2, 2) iterations
    [primal|dual][load]; 'status' (S,); 'updates' (S, 2, 2) fp64 residual evaluations;
    'time' device seconds of the batch; 'schedule' (variant, task order) of the solve kernel and
    'tune_ms' the autotuner's timings (schedule=None: tuned on this batch; or a fixed pair).
    """
    nat.solve_mode(precision)  # 'fp64' or 'mixed'; anything else raises before any device call
    dims = [np.shape(a) for a in fields]
    if any(len(d) != 2 for d in dims):
        raise ValueError('every sample must be a 2-D array')
    b = nat.Batch(dims, Y)
    b.set_coeff(fields)
    if schedule is None:
        tune, var, order = b.autotune(mode=precision, delta=delta)
    else:
        var, order = schedule
        tune = None
        b.set_schedule(var, order)
    up, lo, kit, st = b.solve(tol=tol, maxiter=maxiter, mode=precision, delta=delta)
    nupd, _, ms = b.stats()
    return {'upper': up, 'lower': lo, 'kit': kit.reshape(-1, 2, 2), 'status': st,
            'updates': nupd.reshape(-1, 2, 2), 'time': ms * 1e-3, 'schedule': (var, order), 'tune_ms': tune}
