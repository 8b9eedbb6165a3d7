"""Post-processing: homogenized (bound) matrices from the solution fields.

Mirrors ffthompy/postprocess.py:6-86. The energy forms AH_ij = <A e_i, e_j> run on the
GPU in fp64 with double-double accumulation (fh_energy_matrix); the dual matrix is
the inverse of the dual energy matrix. The 'Ga' kinds (exactly integrated coefficient on
the 2N-1 grid, postprocess.py:20-41) run on the device as well (fh_ga_*): coefficient
spectra from closed forms, solutions interpolated by zero-padded spectra.
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


def ga_assembly_matrix(A_ga, Afun, solutions):
    """AH[i,j] = <A_Ga P(e_i), P(e_j)>: assembly_matrix with the solutions projected onto the
    coefficient's grid (postprocess.py:53-70, Tensor.project objects.py:488-511), on the
    device. A_ga: materials.GaCoefficient; Afun: the GpuOperator of the solve (its plan
    holds the solutions' grid)."""
    fields = [s.to_device() for s in solutions]
    if isinstance(Afun, GpuOperator) and Afun.Nsolve is not None:
        # solutions of a Ga solve: already on the quadrature grid of the solve grid
        ga = A_ga.load(nat.Ga(Afun.coarse_plan, grid=Afun.N))
        return ga.energy_matrix([f.val for f in fields], on_grid=True)
    if isinstance(Afun, GpuOperator):
        plan = Afun.plan
    else:
        N = fields[0].N
        physics = 'scalar' if len(fields) == len(N) else 'elasticity'
        plan = nat.Plan(N, A_ga.Y, physics)
    ga = A_ga.load(nat.Ga(plan))
    return ga.energy_matrix([f.val for f in fields])


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
            if isinstance(A, GpuOperator) and A.Nsolve is not None:
                raise NotImplementedError("'GaNi' post-processing of a Ga solve")
            name = 'AH_%s_%s' % (pp['kind'], primaldual)
            AH = assembly_matrix(A, solutions)
            matrices[name] = AH if primaldual == 'primal' else np.linalg.inv(AH)
        elif pp['kind'] in ('Ga', 'ga'):
            order_name = Nname = ''
            if 'order' in pp:
                Nbar = tuple(2 * np.asarray(pb.solve['N']) - 1)
                if pp['order'] is None:
                    A_ga = mat.get_A_Ga(Nbar=Nbar, primaldual=primaldual, order=None)
                else:
                    order_name = '_o' + str(pp['order'])
                    Nname = '_P%d' % np.mean(pp['P'])
                    A_ga = mat.get_A_Ga(Nbar=Nbar, primaldual=primaldual, order=pp['order'], P=pp['P'])
                AH = ga_assembly_matrix(A_ga, A, solutions)
            else:
                AH = assembly_matrix(A, solutions)
            name = 'AH_%s%s%s_%s' % (pp['kind'], order_name, Nname, primaldual)
            matrices[name] = AH if primaldual == 'primal' else np.linalg.inv(AH)
        else:
            raise ValueError(pp['kind'])
    pb.output.update({'sol_' + primaldual: solutions, 'res_' + primaldual: results,
                      'mat_' + primaldual: matrices})
