#!/usr/bin/env python3
"""Benchmark of the GaNi homogenization solve on one MI355X.

Workload: 3-D scalar conductivity on a 243^3 grid, 30 non-overlapping spheres
(seeded centres, radii U[0.08, 0.15]) of conductivity 100 in a matrix of 1, unit
loads e1, e2, e3. One step = the complete primal + dual homogenization: 6 CG solves
to the reference's absolute tolerance 1e-8 (grid-mean residual norm) plus the
energy forms that give the GaNi primal and dual matrices (equal analytically; reported
with their ordering slack and against the float64 oracle fixture). The guaranteed,
separated bounds (Ga post-processing of the same solutions) are computed after the
timed steps.
The headline runs the fp32 CG with fp64 reliable updates; the plain fp64 solve is
timed beside it. Inputs are resident in HBM before the timed region.

Prints one JSON line (rank 0). Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'time-to-tol 1e-8 (s) + matvec HBM GB/s, 3D scalar 243^3; FP32-refined vs FP64'
HBM_PEAK_GBS = 8000.0


def log(*a):
    print('[bench]', *a, file=sys.stderr, flush=True)


class Homogenizer:
    """Primal and dual plans plus resident fields for the 3 unit loads."""

    def __init__(self, nat, a, Y=(1., 1., 1.)):
        self.nat = nat
        N = a.shape
        self.N = N
        self.plans = {}
        for pd, coef in (('primal', a), ('dual', 1. / a)):
            pl = nat.Plan(N, Y, 'scalar', pd)
            pl.set_coeff('iso', coef)
            self.plans[pd] = pl
        self.B = nat.DeviceArray((3,) + N)
        self.EN = nat.DeviceArray((3,) + N)
        self.sols = {pd: [nat.DeviceArray((3,) + N) for _ in range(3)] for pd in ('primal', 'dual')}

    def step(self, mode, tol=1e-8, maxiter=2000, delta=1e-4):
        nat = self.nat
        out = {'kit': {}, 'solve_ms': {}, 'updates': {}}
        for pd, pl in self.plans.items():
            kits, times, ups = [], [], []
            for iL in range(3):
                E = np.zeros(3)
                E[iL] = 1.
                nat.call('fh_memset_zero', self.EN.ptr, self.EN.nbytes)
                nat.add_constant(self.EN, -E)
                pl.matvec(self.EN, self.B)
                x = self.sols[pd][iL]
                nat.call('fh_memset_zero', x.ptr, x.nbytes)
                info = pl.cg_solve(self.B, x, tol=tol, maxiter=maxiter, mode=mode, delta=delta)
                if info['status'] != nat.FH_CG_OK:
                    raise RuntimeError('CG breakdown')
                kits.append(info['kit'])
                times.append(pl.last_timing()[1])
                ups.append(pl.last_updates())
                nat.add_constant(x, E)
            AH = pl.energy_matrix(self.sols[pd])
            out['AH_' + pd] = AH if pd == 'primal' else np.linalg.inv(AH)
            out['kit'][pd] = kits
            out['solve_ms'][pd] = times
            out['updates'][pd] = ups
        return out


def bounds_report(r_mixed, r_fp64, n):
    """GaNi bound matrices of both arithmetics, the ordering check and the comparison with the float64
    oracle fixture of this workload (tests/golden/oracle_scalar243.json, written on the CPU by
    tests/golden/make_oracle_scalar243.py). GaNi primal and dual agree analytically; the CG iterates
    only overestimate both energies, so lower <= upper holds up to the rounding of the energies: the
    slack stated is 64 eps max|AH|."""
    eps = np.finfo(np.float64).eps
    out = {'note': 'GaNi upper = primal, lower = inverted dual; equal analytically, ordered within the slack'}
    for name, r in (('fp64', r_fp64), ('mixed', r_mixed)):
        up, lo = r['AH_primal'], r['AH_dual']
        slack = 64 * eps * float(np.abs(up).max())
        gap = float(np.linalg.eigvalsh(0.5 * ((up - lo) + (up - lo).T)).min())
        out[name] = {'AH_GaNi_primal': up.tolist(), 'AH_GaNi_dual': lo.tolist(), 'slack': slack,
                     'min_eig_upper_minus_lower': gap, 'lower_le_upper_within_slack': bool(gap >= -slack)}
    path = os.path.join(ROOT, 'tests', 'golden', 'oracle_scalar243.json')
    if n == 243 and os.path.exists(path):
        fx = json.load(open(path))
        ref = {pd: np.asarray(fx['AH_' + pd]) for pd in ('primal', 'dual')}
        rel = {}
        for name, r in (('fp64', r_fp64), ('mixed', r_mixed)):
            rel[name] = [float(np.abs(r['AH_' + pd] - ref[pd]).max() / np.abs(ref[pd]).max())
                         for pd in ('primal', 'dual')]
        out['oracle'] = {'source': 'tests/golden/oracle_scalar243.json (float64 CPU oracle, same microstructure)',
                         'rel_diff': rel,
                         'kit_equal_fp64': r_fp64['kit'] == {pd: fx['kit_' + pd] for pd in ('primal', 'dual')},
                         'kit_equal_mixed': r_mixed['kit'] == {pd: fx['kit_' + pd] for pd in ('primal', 'dual')}}
    return out


def measured_traffic(key):
    """HBM bytes per launch of a kernel, from the committed rocprofv3 counter summary."""
    path = os.path.join(ROOT, 'profiles', 'matvec_243_hbm_traffic.md')
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == 'traffic_bytes' and parts[1] == key:
            return int(parts[2])
    return None


def timed(fn, steps, warmup, sync):
    for _ in range(warmup):
        fn()
    sync()
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = fn()
    sync()
    return (time.perf_counter() - t0) / steps, res


def bench_uq(args, nat, ms, dev):
    """512 2-D samples, grids from {63,81,105,135}^2, a = exp(GRF(corr 0.1, log-std 1)); primal + dual
    bounds of every sample in one persistent graph-launched kernel (fp32 CG + fp64 refinement, tol 1e-8)."""
    from ffthompy_amd.batch import homogenize_batch
    dims, fields = ms.uq_samples()
    log('uq: %d samples, %d voxels' % (len(fields), sum(a.size for a in fields)))
    b = nat.Batch(dims)
    b.set_coeff(fields)
    tune, var, order = b.autotune(mode='mixed', delta=args.delta)
    log('uq: autotune ms %s -> variant %d, order %d' % (np.round(tune, 2).tolist(), var, order))
    for _ in range(max(args.warmup, 1)):
        b.solve(tol=1e-8, mode='mixed', delta=args.delta)
    ts = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        up, lo, kit, st = b.solve(tol=1e-8, mode='mixed', delta=args.delta)
        ts.append(time.perf_counter() - t0)
    _, _, dev_ms = b.stats()
    t64 = time.perf_counter()
    up64, lo64, kit64, _ = b.solve(tol=1e-8, mode='fp64')
    t64 = time.perf_counter() - t64
    b.solve(tol=1e-8, mode='mixed', delta=args.delta)
    gap = min(float(np.linalg.eigvalsh(u - l).min()) for u, l in zip(up, lo))
    it_vox = float(sum(int(k.sum()) * a.size for k, a in zip(kit, fields)))
    cpu = None
    if not args.no_cpu:
        from oracle.baseline import time_cg_iterations
        pick = [int(i) for i in np.random.default_rng(3).choice(len(fields), 4, replace=False)]
        sec_vox = []
        for i in pick:
            s_it = time_cg_iterations(fields[i], np.ones(2), iters=args.cpu_iters * 5, callback=True)
            sec_vox.append(s_it / fields[i].size)
        per = float(np.mean(sec_vox))
        cpu = {'value': per * it_vox, 'unit': 's', 'cores': 1, 'kind': 'port',
               'sample': '%d CG iterations each on 4 samples %s of the oracle (reference-style NumPy, CallBack '
                         'on), seconds per voxel-iteration extrapolated to the batch\'s %.3g voxel-iterations'
                         % (args.cpu_iters * 5, [tuple(int(v) for v in dims[i]) for i in pick], it_vox)}
    t = float(np.median(ts))
    out = {'metric': 'time to bounds of 512 2-D UQ samples (s)', 'value': round(t, 4), 'unit': 's',
           'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup, 'higher_is_better': False,
           'dtype': 'f32+f64', 'data': 'synthetic (GRF samples, seed %d)' % ms.BASE_SEED,
           'config': {'workload': 'batched UQ: 512 2-D scalar samples, grids {63,81,105,135}^2, a = exp(GRF '
                                  'corr 0.1, log-std 1.0), primal+dual, CG tol 1e-8, one persistent launch'},
           'device_ms': round(dev_ms, 2), 'samples_per_s': round(len(fields) / t, 1),
           'cg_iterations': int(kit.sum()), 'cg_iter_per_s': round(float(kit.sum()) / t, 0),
           'schedule': {'variant': var, 'order': order, 'tune_ms': np.round(tune, 3).tolist()},
           'fp64_time_s': round(t64, 4), 'bounds_rel_diff_mixed_vs_fp64': float(
               max(np.abs(up - up64).max() / np.abs(up64).max(), np.abs(lo - lo64).max() / np.abs(lo64).max())),
           'min_gap_upper_minus_lower': gap, 'cpu_baseline': cpu}
    print(json.dumps(out), flush=True)


def bench_elastic(args, nat, ms, dev):
    """81^3 small-strain elasticity: seeded random ellipsoids (~25 %), matrix (E 1, nu 0.3) vs inclusion
    (10, 0.2), 6 Mandel unit strains, primal + dual, fp32 CG + fp64 refinement vs plain fp64; plans and
    coefficients resident, timed: right-hand sides, 12 solves, energies."""
    from ffthompy_amd.materials import lame_compliance, mandel_to_voigt
    N = (81, 81, 81)
    ind, n_ell = ms.ellipsoid_indicator(N)
    ind = ind.astype(bool)
    lm0, mu0 = ms.lame_from_young(1.0, 0.3)
    lm1, mu1 = ms.lame_from_young(10.0, 0.2)
    lam = np.where(ind, lm1, lm0)
    mu = np.where(ind, mu1, mu0)
    log('elastic81: %d ellipsoids, inclusion fraction %.4f' % (n_ell, ind.mean()))
    m = np.array([1., 1., 1., 0., 0., 0.])

    # plans and coefficients are resident before the timed region (as in the headline workload)
    plans = {}
    for pd in ('primal', 'dual'):
        pl = nat.Plan(N, (1., 1., 1.), 'elasticity', pd)
        if pd == 'primal':
            pl.set_coeff('lame', np.stack([lam, mu]))
        else:
            # the dual coefficient: the pointwise compliance C^-1, again of Lame form
            pl.set_coeff('lame', np.stack(lame_compliance(lam, mu)))
        plans[pd] = pl
    schedule = {}
    if not args.no_autotune:
        for pd, pl in plans.items():
            for dt in (np.float32, np.float64):
                r = pl.autotune(dt, nrep=5)
                schedule['%s_%s' % (pd, np.dtype(dt).name)] = (r['proj_variant'], r['cols_variant'], r['rows_variant'])
        log('elastic81: schedules %s' % schedule)
    # right-hand sides, solutions and the macroscopic field are allocated before the timed region
    EN = nat.DeviceArray((6,) + N)
    B = nat.DeviceArray((6,) + N)
    X = {pd: [nat.DeviceArray((6,) + N) for _ in range(6)] for pd in plans}

    def run(mode):
        res = {}
        for pd in ('primal', 'dual'):
            pl = plans[pd]
            kits = []
            for iL in range(6):
                E = np.zeros(6)
                E[iL] = 1.
                nat.axpby(0., EN, 0., EN)
                nat.add_constant(EN, -E)
                pl.matvec(EN, B)
                x = X[pd][iL]
                nat.axpby(0., x, 0., x)
                info = pl.cg_solve(B, x, tol=1e-8, maxiter=5000, mode=mode, delta=args.delta)
                kits.append(info['kit'])
                nat.add_constant(x, E)
            AH = pl.energy_matrix(X[pd])
            res[pd] = (AH if pd == 'primal' else np.linalg.inv(AH), kits)
        return res

    run('mixed')
    run('fp64')
    nat.synchronize()
    t = time.perf_counter()
    rm = run('mixed')
    nat.synchronize()
    tm = time.perf_counter() - t
    t = time.perf_counter()
    r64 = run('fp64')
    nat.synchronize()
    t64 = time.perf_counter() - t
    up, lo = rm['primal'][0], rm['dual'][0]
    out = {'metric': 'time to primal+dual bounds, 3-D elasticity 81^3 (s)', 'value': round(tm, 4), 'unit': 's',
           'n_gpus': 1, 'higher_is_better': False, 'dtype': 'f32+f64',
           'data': 'synthetic (seeded ellipsoids, seed %d)' % ms.BASE_SEED,
           'config': {'workload': '3-D elasticity 81^3, ellipsoids %.3f, (E,nu) (1,0.3) vs (10,0.2), 6 Mandel '
                                  'loads, primal+dual, CG tol 1e-8' % ind.mean(),
                      'solve': 'fp32 CG (fused iteration) + fp64 reliable updates (delta %g), each fp32 '
                               'correction projected in fp64' % args.delta},
           'fp64_time_s': round(t64, 4), 'speedup_mixed_vs_fp64': round(t64 / tm, 3), 'schedule': schedule,
           'kit_mixed': {k: v[1] for k, v in rm.items()}, 'kit_fp64': {k: v[1] for k, v in r64.items()},
           'AH_upper_mandel_diag': np.round(np.diag(up), 6).tolist(),
           'AH_lower_mandel_diag': np.round(np.diag(lo), 6).tolist(),
           'AH_upper_voigt_diag': np.round(np.diag(mandel_to_voigt(up)), 6).tolist(),
           'AH_lower_voigt_diag': np.round(np.diag(mandel_to_voigt(lo)), 6).tolist(),
           'voigt_from_mandel': 'AH_voigt = T AH_mandel T, T = diag(1, 1, 1, 1/sqrt2, 1/sqrt2, 1/sqrt2)',
           'bounds_rel_diff_mixed_vs_fp64': float(max(
               np.abs(up - r64['primal'][0]).max() / np.abs(r64['primal'][0]).max(),
               np.abs(lo - r64['dual'][0]).max() / np.abs(r64['dual'][0]).max())),
           'min_gap_upper_minus_lower': float(np.linalg.eigvalsh(0.5 * ((up - lo) + (up - lo).T)).min())}
    print(json.dumps(out), flush=True)


def bench_grad(args, nat, ms, dev):
    """99^3 differentiable upper bound: theta = GRF(corr 0.15, std 0.8), loss ||AH - 2I||_F^2; forward
    (3 fp64 solves + energies) and the hand-written backward (one pass over the stored fields)."""
    import torch
    from ffthompy_amd.autograd import upper_bound
    N = (99, 99, 99)
    th = ms.gaussian_random_field(N, 0.15, 0.8, ms.BASE_SEED + 5)
    tf, tb = [], []
    for it in range(args.steps + 1):
        t = torch.tensor(th, dtype=torch.float64, requires_grad=True)
        t0 = time.perf_counter()
        AH = upper_bound(t, tol=1e-10)
        L = torch.sum((AH - 2. * torch.eye(3, dtype=torch.float64)) ** 2)
        t1 = time.perf_counter()
        L.backward()
        t2 = time.perf_counter()
        if it:
            tf.append(t1 - t0)
            tb.append(t2 - t1)
    out = {'metric': 'forward + backward time, differentiable 99^3 upper bound (s)',
           'value': round(float(np.median(tf) + np.median(tb)), 4), 'unit': 's', 'n_gpus': 1,
           'higher_is_better': False, 'dtype': 'f64', 'data': 'synthetic (GRF, seed %d)' % (ms.BASE_SEED + 5),
           'config': {'workload': 'a = exp(theta) on 99^3, theta GRF corr 0.15 std 0.8, loss ||AH-2I||_F^2, '
                                  'fp64 CG tol 1e-10'},
           'forward_s': round(float(np.median(tf)), 4), 'backward_s': round(float(np.median(tb)), 5),
           'loss': float(L.item()), 'grad_norm': float(t.grad.norm().item())}
    print(json.dumps(out), flush=True)


def bench_ga(args, nat, ms, dev):
    """243^3 spheres (contrast 100): GaNi primal+dual solves (mixed precision, tol 1e-8), then Ga
    post-processing of the voxel coefficient (order 0, P = N: each voxel a constant box,
    exactly integrated) on the 2N-1 = 485 grid, run on the 486^3 quadrature grid: guaranteed,
    separated bounds of the voxel microstructure."""
    N = (args.n,) * 3
    a, centres, radii = ms.spheres_conductivity(N, count=30, contrast=100., seed=ms.BASE_SEED)
    H = Homogenizer(nat, a)
    r = H.step('mixed', delta=args.delta)
    log('ga243: GaNi solves done, kit %s' % r['kit'])
    res, times = {}, []
    for rep in range(args.steps + 1):
        t0 = time.perf_counter()
        for pd, coef in (('primal', a), ('dual', 1. / a)):
            ga = nat.Ga(H.plans[pd])
            ga.set_grid_coeff(0, N, 'iso', coef)
            AH = ga.energy_matrix(H.sols[pd])
            res[pd] = AH if pd == 'primal' else np.linalg.inv(AH)
            M = ga.M
            del ga
        nat.synchronize()
        if rep:
            times.append(time.perf_counter() - t0)
        log('ga243: Ga post-processing %.3f s' % (time.perf_counter() - t0))
    up, lo = res['primal'], res['dual']
    out = {'metric': 'time of the Ga post-processing (primal + dual), 243^3 spheres (s)',
           'value': round(float(np.median(times)), 4), 'unit': 's', 'n_gpus': 1, 'higher_is_better': False,
           'dtype': 'f64', 'data': 'synthetic (seeded spheres, seed %d)' % ms.BASE_SEED,
           'config': {'workload': 'Ga post-processing (order 0, P = N) of the GaNi solutions of 3-D scalar '
                                  '%d^3, 30 spheres contrast 100, loads e1-e3; quadrature grid %s (2N-1 = %d)'
                                  % (args.n, list(M), 2 * args.n - 1)},
           'AH_Ga_upper_diag': np.round(np.diag(up), 8).tolist(),
           'AH_Ga_lower_diag': np.round(np.diag(lo), 8).tolist(),
           'AH_GaNi_upper_diag': np.round(np.diag(r['AH_primal']), 8).tolist(),
           'AH_GaNi_lower_diag': np.round(np.diag(r['AH_dual']), 8).tolist(),
           'min_gap_upper_minus_lower': float(np.linalg.eigvalsh(0.5 * ((up - lo) + (up - lo).T)).min()),
           'kit_gani_mixed': r['kit']}
    print(json.dumps(out), flush=True)


def bench_potential(args, nat, ms, dev):
    """243^3 spheres (contrast 100), load e1, primal, tol 1e-8, fp64: the potential formulation
    (one scalar spectrum per CG vector, fh_potential_solve) beside the gradient formulation's
    fp64 CG of the same load (both end in the same AH_11) and the potential solve in mixed
    precision (fp32 iterations, fp64 reliable updates)."""
    N = (args.n,) * 3
    a = ms.spheres_conductivity(N, count=30, contrast=100., seed=ms.BASE_SEED)[0]
    plan = nat.Plan(N, (1., 1., 1.), 'scalar', 'primal')
    plan.set_coeff('iso', a)
    E = np.eye(3)[0]
    EN = nat.DeviceArray.zeros((3,) + N)
    nat.add_constant(EN, -E)
    B = nat.DeviceArray((3,) + N)
    plan.matvec(EN, B)
    del EN
    X = nat.DeviceArray((3,) + N)
    Fu = nat.DeviceArray((args.n, args.n, args.n // 2 + 1), np.complex128)
    tp, tg, tm = [], [], []
    Xm = nat.DeviceArray((3,) + N)
    Fum = nat.DeviceArray((args.n, args.n, args.n // 2 + 1), np.complex128)
    for rep in range(args.warmup + args.steps):
        nat.synchronize()
        t0 = time.perf_counter()
        _, pinfo = plan.potential_solve(E, 1e-8, 3000, Fu=Fu, X=X)
        t1 = time.perf_counter()
        _, minfo = plan.potential_solve(E, 1e-8, 3000, Fu=Fum, X=Xm, mode='mixed', delta=args.delta)
        tm.append(time.perf_counter() - t1)
        x = nat.DeviceArray.zeros((3,) + N)
        nat.synchronize()
        t2 = time.perf_counter()
        ginfo = plan.cg_solve(B, x, tol=1e-8, maxiter=3000, mode='fp64')
        t3 = time.perf_counter()
        if rep >= args.warmup:
            tp.append(t1 - t0)
            tg.append(t3 - t2)
        log('potential243: potential %.4f s (kit %d), mixed %.4f s (kit %d), gradient fp64 %.4f s (kit %d)'
            % (t1 - t0, pinfo['kit'], tm[-1], minfo['kit'], t3 - t2, ginfo['kit']))
    nat.add_constant(X, E)
    nat.add_constant(x, E)
    nat.add_constant(Xm, E)
    ahm = plan.energy_matrix([Xm, Xm, Xm])[0, 0]
    ahp = plan.energy_matrix([X, X, X])[0, 0]
    ahg = plan.energy_matrix([x, x, x])[0, 0]
    vp, vg = float(np.median(tp)), float(np.median(tg))
    out = {'metric': 'time-to-tol 1e-8 of the potential formulation (load e1, primal, fp64), 243^3 spheres (s)',
           'value': round(vp, 4), 'unit': 's', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
           'higher_is_better': False, 'dtype': 'f64', 'data': 'synthetic (seeded spheres, seed %d)' % ms.BASE_SEED,
           'config': {'workload': 'potential formulation, 3-D scalar %d^3, 30 spheres contrast 100, load e1, '
                                  'primal, CG tol 1e-8 from zero' % args.n},
           'kit_potential': pinfo['kit'], 'ms_per_iteration_potential': round(1e3 * vp / pinfo['kit'], 4),
           'gradient_fp64_s': round(vg, 4), 'kit_gradient_fp64': ginfo['kit'],
           'ms_per_iteration_gradient_fp64': round(1e3 * vg / ginfo['kit'], 4),
           'AH11_potential': ahp, 'AH11_gradient': ahg, 'AH11_rel_diff': abs(ahp - ahg) / ahg,
           'mixed_s': round(float(np.median(tm[args.warmup:])), 4), 'kit_mixed': minfo['kit'],
           'AH11_mixed_rel_diff': abs(ahm - ahp) / ahp}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=3)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--n', type=int, default=243, help='grid points per axis')
    ap.add_argument('--fp64-steps', type=int, default=None, help='timed fp64 steps (default: --steps)')
    ap.add_argument('--cpu-iters', type=int, default=2, help='CG iterations in the CPU baseline sample')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline leg')
    ap.add_argument('--delta', type=float, default=None,
                    help='mixed precision: residual reduction that triggers an fp64 reliable update (default '
                         '3e-5 for elastic81, 1e-4 otherwise: the fastest settings of a sweep; each update '
                         'projects the fp32 correction and recomputes the residual in fp64)')
    ap.add_argument('--profile-only', action='store_true', help='matvec profile only (for rocprofv3 runs)')
    ap.add_argument('--no-autotune', action='store_true', help='keep the default kernel schedules')
    ap.add_argument('--no-ga', action='store_true', help='skip the Ga post-processing bounds after the timed steps')
    ap.add_argument('--profile-solve', action='store_true',
                    help='with --profile-only: also one mixed and one fp64 CG solve (load e1, primal)')
    ap.add_argument('--workload', default='scalar243', choices=['scalar243', 'uq512', 'elastic81', 'grad99', 'ga243', 'potential243'],
                    help='scalar243: the headline (default); the others print their own JSON line')
    args = ap.parse_args()
    if args.delta is None:
        args.delta = 3e-5 if args.workload == 'elastic81' else 1e-4
    world = int(os.environ.get('WORLD_SIZE', '1'))
    if args.gpus != 1 or world != 1:
        print('bench.py runs on one GPU only: the library has no multi-GPU path '
              '(--gpus %d, WORLD_SIZE %d requested)' % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    from ffthompy_amd import _native as nat
    from ffthompy_amd import microstructure as ms
    dev = nat.init()
    if args.workload != 'scalar243':
        {'uq512': bench_uq, 'elastic81': bench_elastic, 'grad99': bench_grad,
         'ga243': bench_ga, 'potential243': bench_potential}[args.workload](args, nat, ms, dev)
        return
    N = (args.n,) * 3
    n = int(np.prod(N))
    nh = args.n * args.n * (args.n // 2 + 1)
    log('device', dev, 'grid', N)
    t = time.perf_counter()
    a, centres, radii = ms.spheres_conductivity(N, count=30, contrast=100., seed=ms.BASE_SEED)
    log('microstructure: %d spheres, volume fraction %.4f (%.1fs)' % (len(radii), (a > 1).mean(),
                                                                        time.perf_counter() - t))
    H = Homogenizer(nat, a)

    # --- schedule autotuner: every variant of the projection and strided passes, per arithmetic ---
    pl = H.plans['primal']
    sched = {}
    if not args.no_autotune:
        for dt, name in ((np.float32, 'fp32'), (np.float64, 'fp64')):
            res = pl.autotune(dt, nrep=5)
            H.plans['dual'].set_schedule(dt, res['proj_variant'], res['cols_variant'], res['rows_variant'])
            sched[name] = res
            log('autotune %s: projection tile variants %s ms -> %d; strided %s ms -> %d; rows %s ms -> %d' % (
                name, np.round(res['proj_ms'], 4).tolist(), res['proj_variant'], np.round(res['cols_ms'], 4).tolist(),
                res['cols_variant'], np.round(res['rows_ms'], 4).tolist(), res['rows_variant']))

    # --- matvec profile (per kernel, HIP events on the plan stream) ---
    prof = {}
    for dt, name in ((np.float32, 'fp32'), (np.float64, 'fp64')):
        x = nat.DeviceArray.from_host(np.random.default_rng(1).standard_normal((3,) + N).astype(dt))
        y = nat.DeviceArray((3,) + N, dt)
        pl.matvec(x, y)
        ms_k, by_k = pl.matvec_profile(x, y, nrep=10)
        sr = 4 if dt == np.float32 else 8
        model = n * sr * (2 * 3 + 1) + 8 * 3 * nh * 2 * sr
        prof[name] = {'kernel_ms': ms_k.tolist(), 'kernel_bytes': by_k.tolist(),
                      'matvec_ms': float(ms_k.sum()), 'model_bytes': model,
                      'matvec_GBps': model / (ms_k.sum() * 1e-3) / 1e9}
        log('matvec %s: %.3f ms total, kernels %s ms, %.0f GB/s (model bytes)' % (
            name, ms_k.sum(), np.round(ms_k, 4).tolist(), prof[name]['matvec_GBps']))
        del x, y
    if args.profile_only:
        if args.profile_solve:
            for mode in ('mixed', 'fp64'):
                E = np.array([1., 0., 0.])
                nat.call('fh_memset_zero', H.EN.ptr, H.EN.nbytes)
                nat.add_constant(H.EN, -E)
                pl.matvec(H.EN, H.B)
                x = H.sols['primal'][0]
                nat.call('fh_memset_zero', x.ptr, x.nbytes)
                info = pl.cg_solve(H.B, x, tol=1e-8, maxiter=2000, mode=mode, delta=args.delta)
                log('solve %s: kit %d, %.1f ms' % (mode, info['kit'], pl.last_timing()[1]))
        # plain device copy of a known size: the reference point for HBM byte counters
        src = nat.DeviceArray((1 << 27,), np.float64)   # 1 GiB
        dst = nat.DeviceArray((1 << 27,), np.float64)
        for _ in range(3):
            dst.copy_from(src)
        nat.synchronize()
        log('copy probe: 3 x 1 GiB device-to-device copies')
        return
    k32 = prof['fp32']
    dom = int(np.argmax(k32['kernel_ms']))
    names = ['rows_fwd(A, R2C axis 2)', 'cols axis 1 fwd', 'cols axis 0 fwd+G+inv', 'cols axis 1 inv',
             'rows_inv(C2R axis 2)']
    ach = k32['kernel_bytes'][dom] / (k32['kernel_ms'][dom] * 1e-3) / 1e9
    traffic = measured_traffic(['rows_fwd', 'cols_fwd', 'cols_g', 'cols_inv', 'rows_inv'][dom] + '_fp32')

    # --- timed steps ---
    sync = nat.synchronize
    log('mixed-precision steps: warmup %d, timed %d' % (args.warmup, args.steps))
    t_mixed, r_mixed = timed(lambda: H.step('mixed', delta=args.delta), args.steps, args.warmup, sync)
    log('mixed step %.3f s, kit %s' % (t_mixed, r_mixed['kit']))
    fs = args.steps if args.fp64_steps is None else args.fp64_steps
    t_fp64, r_fp64 = timed(lambda: H.step('fp64'), fs, 1 if args.warmup > 0 else 0, sync)
    log('fp64 step %.3f s, kit %s' % (t_fp64, r_fp64['kit']))
    relp = float(np.abs(r_mixed['AH_primal'] - r_fp64['AH_primal']).max() / np.abs(r_fp64['AH_primal']).max())
    reld = float(np.abs(r_mixed['AH_dual'] - r_fp64['AH_dual']).max() / np.abs(r_fp64['AH_dual']).max())
    kit64 = sum(sum(v) for v in r_fp64['kit'].values())
    bounds = bounds_report(r_mixed, r_fp64, args.n)

    # --- guaranteed bounds: Ga post-processing of the last step's (fp64) solutions, untimed ---
    ga = None
    if not args.no_ga:
        t = time.perf_counter()
        gres = {}
        for pd, coef in (('primal', a), ('dual', 1. / a)):
            gpp = nat.Ga(H.plans[pd])
            gpp.set_grid_coeff(0, N, 'iso', coef)  # order 0, P = N: each voxel an exactly integrated box
            AHg = gpp.energy_matrix(H.sols[pd])
            gres[pd] = AHg if pd == 'primal' else np.linalg.inv(AHg)
            M = gpp.M
            del gpp
        nat.synchronize()
        up, lo = gres['primal'], gres['dual']
        ga = {'kind': 'Ga post-processing, order 0, P = N, of the fp64 solutions (guaranteed bounds of the '
                      'voxel microstructure, ffthompy/postprocess.py:20-46)',
              'quadrature_grid': list(M), 'AH_Ga_upper_diag': np.diag(up).tolist(),
              'AH_Ga_lower_diag': np.diag(lo).tolist(),
              'min_eig_upper_minus_lower': float(np.linalg.eigvalsh(0.5 * ((up - lo) + (up - lo).T)).min()),
              'seconds': round(time.perf_counter() - t, 3)}
        log('Ga bounds: upper %s lower %s (%.2f s)' % (np.round(np.diag(up), 5), np.round(np.diag(lo), 5),
                                                       ga['seconds']))

    # --- CPU baseline: reference-style NumPy path on this host, bounded sample ---
    cpu = None
    if not args.no_cpu and args.cpu_iters > 0:
        from oracle.baseline import time_cg_iterations
        log('cpu baseline: %d reference-style CG iterations at %s (fp64, CallBack on)' % (args.cpu_iters, N))
        s_it = time_cg_iterations(a, np.ones(3), iters=args.cpu_iters, callback=True)
        cpu = {'value': s_it * kit64, 'unit': 's', 'cores': 1, 'kind': 'port',
               'sample': ('%d CG iterations of the oracle (reference-style NumPy: materialised hG1 and '
                          'full A via einsum, numpy.fft, allocating CG, CallBack matvec) on the 243^3 '
                          'load e1: %.2f s/iteration, extrapolated to the %d fp64 CG iterations of one '
                          'step; os.cpu_count()=%s, affinity=%d, numpy.fft single-threaded'
                          % (args.cpu_iters, s_it, kit64, os.cpu_count(), len(os.sched_getaffinity(0)))),
               'sec_per_iter': s_it}
        log('cpu baseline: %.2f s/iter -> %.0f s per step (extrapolated)' % (s_it, cpu['value']))
        # the same iterations without the CallBack matvec, and a tuned CPU variant (scipy.fft on every
        # core of this process, no CallBack): reported beside the baseline, not used as it
        s_nc = time_cg_iterations(a, np.ones(3), iters=args.cpu_iters, callback=False)
        nw = len(os.sched_getaffinity(0))
        s_tuned = time_cg_iterations(a, np.ones(3), iters=args.cpu_iters, callback=False, workers=nw)
        cpu['sec_per_iter_no_callback'] = s_nc
        cpu['tuned_cpu'] = {'value': s_tuned * kit64, 'unit': 's', 'cores': nw, 'sec_per_iter': s_tuned,
                            'sample': 'same iterations with scipy.fft (workers=%d) and no CallBack, '
                                      'extrapolated to %d iterations' % (nw, kit64)}
        log('cpu baseline: %.2f s/iter without CallBack; tuned (scipy.fft, %d workers) %.2f s/iter'
            % (s_nc, nw, s_tuned))

    out = {
        'metric': METRIC,
        'value': round(t_mixed, 4),
        'unit': 's',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(t_mixed * 1e3, 2),
        'higher_is_better': False,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32+f64',
        'data': 'synthetic (seeded spheres, seed %d)' % ms.BASE_SEED,
        'config': {'workload': 'GaNi 3-D scalar %d^3, 30 spheres contrast 100, loads e1-e3, primal+dual, '
                               'CG tol 1e-8' % args.n,
                   'grid': list(N), 'volume_fraction': round(float((a > 1).mean()), 4),
                   'solve': 'fp32 CG (fused iteration) + fp64 reliable updates (delta %g), each fp32 correction '
                            'projected in fp64' % args.delta},
        'fp64_time_s': round(t_fp64, 4),
        'speedup_mixed_vs_fp64': round(t_fp64 / t_mixed, 3),
        'kit_mixed': r_mixed['kit'], 'kit_fp64': r_fp64['kit'],
        'reliable_updates': r_mixed['updates'],
        'solve_ms_mixed': r_mixed['solve_ms'], 'solve_ms_fp64': r_fp64['solve_ms'],
        'bounds': bounds,
        'bounds_rel_diff_mixed_vs_fp64': [relp, reld],
        'bounds_rel_diff_vs_oracle': bounds.get('oracle', {}).get('rel_diff'),
        'ga_bounds': ga,
        'matvec_GBps_fp32': round(prof['fp32']['matvec_GBps'], 1),
        'matvec_frac_of_hbm_peak_fp32': round(prof['fp32']['matvec_GBps'] / HBM_PEAK_GBS, 4),
        'matvec_GBps_fp64': round(prof['fp64']['matvec_GBps'], 1),
        'matvec_ms_fp32': round(prof['fp32']['matvec_ms'], 4),
        'matvec_ms_fp64': round(prof['fp64']['matvec_ms'], 4),
        'schedule': sched,
        'matvec_kernels_fp32': {'names': names, 'ms': [round(v, 4) for v in k32['kernel_ms']],
                                'bytes': k32['kernel_bytes']},
        'roofline': {'bound': 'hbm', 'kernel': names[dom], 'achieved': round(ach, 1), 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': round(ach / HBM_PEAK_GBS, 4), 'traffic': traffic,
                     'traffic_source': 'profiles/matvec_243_hbm_traffic.md (rocprofv3 FETCH_SIZE/WRITE_SIZE)'
                     if traffic else None},
        'cpu_baseline': cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
