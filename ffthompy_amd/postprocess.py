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


def gani_of_ga_solve(Afun, mat, solutions, primaldual):
    """'GaNi' energies of a Ga solve's minimisers: the nodal coefficient on the solve grid N
    (postprocess.py:11-15) and the solutions restricted from the quadrature grid to N by
    cropping their spectra (Tensor.project inside assembly_matrix, objects.py:469-511)."""
    ga = nat.Ga(Afun.coarse_plan, grid=Afun.N)
    fields = [ga.restrict(s.to_device().val) for s in solutions]
    plan = nat.Plan(Afun.Nsolve, Afun.Y, Afun.physics, primaldual)
    if 'inclusions' in mat.conf:
        mat.set_A_GaNi_device(plan, primaldual)
    else:
        kind, vals = classify_coefficient(mat.get_A_GaNi(Afun.Nsolve, primaldual).val, Afun.physics)
        plan.set_coeff(kind, vals)
    return plan.energy_matrix(fields)


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
    (postprocess.py:6-51). As in the reference, the coefficient carries over from one entry
    to the next: 'GaNi' switches to the nodal coefficient (unless it is already in use),
    'Ga' with 'order' to that exactly integrated coefficient, and 'Ga' without 'order' uses
    whichever coefficient is current (the solve's at first)."""
    matrices = {}
    ga_solve = isinstance(A, GpuOperator) and A.Nsolve is not None
    current = ('solve', None)   # the solve's coefficient: nodal (GaNi solve) or A_Ga (Ga solve)
    for pp in pb.postprocess:
        order_name = Nname = ''
        if pp['kind'] in ('GaNi', 'gani'):
            if ga_solve or current[0] == 'ga':
                current = ('gani', None)
        elif pp['kind'] in ('Ga', 'ga'):
            if 'order' in pp:
                Nbar = tuple(2 * np.asarray(pb.solve['N']) - 1)
                if pp['order'] is None:
                    current = ('ga', mat.get_A_Ga(Nbar=Nbar, primaldual=primaldual, order=None))
                else:
                    order_name = '_o' + str(pp['order'])
                    Nname = '_P%d' % np.mean(pp['P'])
                    current = ('ga', mat.get_A_Ga(Nbar=Nbar, primaldual=primaldual, order=pp['order'], P=pp['P']))
        else:
            raise ValueError(pp['kind'])
        if current[0] == 'ga':
            AH = ga_assembly_matrix(current[1], A, solutions)
        elif current[0] == 'gani' and ga_solve:
            AH = gani_of_ga_solve(A, mat, solutions, primaldual)
        else:  # the solve's coefficient, or the nodal one of a GaNi solve (the same)
            AH = assembly_matrix(A, solutions)
        name = 'AH_%s%s%s_%s' % (pp['kind'], order_name, Nname, primaldual)
        matrices[name] = AH if primaldual == 'primal' else np.linalg.inv(AH)
    pb.output.update({'sol_' + primaldual: solutions, 'res_' + primaldual: results,
                      'mat_' + primaldual: matrices})
