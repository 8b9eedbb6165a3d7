"""Post-processing: homogenized (bound) matrices from the solution fields.

Mirrors ffthompy/postprocess.py:6-86. The energy forms AH_ij = <A e_i, e_j> run on the
GPU in fp64 with double-double accumulation (fh_energy_matrix); the dual matrix is
the inverse of the dual energy matrix. GaNi post-processing is provided; the 'Ga'
kinds (exact-integration coefficient on the 2N-1 grid) are not on this path.
"""
import numpy as np

from . import _native as nat
from .tensors import GpuOperator, Tensor, classify_coefficient


def assembly_matrix(A, solutions):
    """AH[i,j] = <A e_i, e_j> (postprocess.py:53-70). A: GpuOperator or coefficient Tensor."""
    fields = [s.to_device() for s in solutions]
    if isinstance(A, GpuOperator):
        plan = A.plan
    else:
        N = fields[0].N
        physics = 'scalar' if len(fields) == len(N) else 'elasticity'
        plan = nat.Plan(N, A.Y if isinstance(A, Tensor) else np.ones(len(N)), physics)
        kind, vals = classify_coefficient(A.val if isinstance(A, Tensor) else A, physics)
        plan.set_coeff(kind, vals)
    return plan.energy_matrix([f.val for f in fields])


def add_macro2minimizer(X, E):
    """Return the corrector plus the macroscopic value E (postprocess.py:73-86)."""
    m = X.mean()
    if np.allclose(m, E):
        return X
    if np.allclose(m, np.zeros_like(E)):
        out = X.copy(name='EN+' + X.name[:10])
        out.add_mean(np.asarray(E, dtype=np.float64))
        return out
    raise ValueError("Field is neither zero-mean nor E-mean.")


def postprocess(pb, A, mat, solutions, results, primaldual):
    """Store 'mat_<pd>' (named AH matrices), 'sol_<pd>' and 'res_<pd>' in pb.output
    (postprocess.py:6-51)."""
    matrices = {}
    for pp in pb.postprocess:
        if pp['kind'] in ('GaNi', 'gani'):
            name = 'AH_%s_%s' % (pp['kind'], primaldual)
            AH = assembly_matrix(A, solutions)
            matrices[name] = AH if primaldual == 'primal' else np.linalg.inv(AH)
        elif pp['kind'] in ('Ga', 'ga'):
            raise NotImplementedError("'Ga' post-processing is not available on the GPU path")
        else:
            raise ValueError(pp['kind'])
    pb.output.update({'sol_' + primaldual: solutions, 'res_' + primaldual: results,
                      'mat_' + primaldual: matrices})
