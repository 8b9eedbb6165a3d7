/*
 * ffthom: C ABI of the MI355X (gfx950) Fourier-Galerkin homogenization operator and solver.
 *
 * Drop-in boundary for the GaNi solve of FFTHomPy (vondrejc/FFTHomPy). The reference
 * builds, in ffthompy/applications.py:24-58,
 *     Afun = Operator([[FiN, hG, FN], A])     i.e.  y = F^-1 G(xi) F (A(x) v(x))
 * and hands it to linear_solver('CG', Afun, B, x0, par, callback)
 * (ffthompy/general/solver.py:8-60). Every entry point below replaces one piece of
 * that call graph; its comment names the reference interface (file:line).
 *
 * Conventions
 *  - All functions return FH_OK (0) or an FH_ERR_* code; fh_last_error() gives the
 *    message of the calling thread's last failure. No entry point aborts the process.
 *  - Fields are device pointers (from fh_malloc) in the reference Tensor layout
 *    shape+N: component-major, C-order, last axis contiguous, origin at index 0.
 *  - Spectra in the reference 'r' form: (ncomp,) + N[:-1] + (N[-1]//2+1,), complex
 *    interleaved (re, im), unnormalised forward transform (numpy.fft.rfftn).
 *  - A plan owns one HIP stream (default priority). Calls on one plan are not
 *    re-entrant; use one plan per host thread. Every entry point has finished its
 *    GPU work when it returns.
 */
#ifndef FFTHOM_H
#define FFTHOM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FH_OK = 0,
  FH_ERR_ARG = 1,          /* invalid argument (shape, pointer, enum) */
  FH_ERR_HIP = 2,          /* HIP runtime failure */
  FH_ERR_NODEV = 3,        /* no usable GPU */
  FH_ERR_UNSUPPORTED = 4,  /* grid or tile that does not fit a workgroup, wrong GPU arch */
  FH_ERR_STATE = 5         /* call out of order (e.g. solve before coefficients) */
};

enum { FH_SCALAR = 0, FH_ELASTICITY = 1 };           /* ffthompy/problem.py:33-36 */
enum { FH_PRIMAL = 0, FH_DUAL = 1 };                 /* ffthompy/applications.py:53-56 */
enum { FH_F64 = 0, FH_F32 = 1 };                     /* element type of a device field */
enum { FH_SOLVE_FP64 = 0, FH_SOLVE_MIXED = 1 };      /* CG arithmetic */
enum { FH_COEF_ISO = 1, FH_COEF_LAME = 2, FH_COEF_FULL = 3 };
enum { FH_ITER_RICHARDSON = 1, FH_ITER_CHEBYSHEV = 2 };
enum { FH_CG_OK = 0, FH_CG_BREAKDOWN = 1 };

typedef struct fh_plan fh_plan;
typedef struct fh_batch fh_batch;
typedef struct fh_ga fh_ga;

/* Message of the calling thread's last error ("" if none). */
const char* fh_last_error(void);

/* Select device 0 and check it is a gfx950 GPU. Writes the device name.
 * No reference counterpart (the reference is host-only NumPy). */
int fh_init(char* name, int name_len);

/* Device memory. No reference counterpart: NumPy owns the reference's memory
 * (ffthompy/tensors/objects.py:117-121 allocates Tensor.val with np.zeros). */
int fh_malloc(void** dptr, size_t bytes);
int fh_free(void* dptr);
int fh_memcpy_htod(void* dst, const void* src, size_t bytes);
int fh_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int fh_memcpy_dtod(void* dst, const void* src, size_t bytes);
int fh_memset_zero(void* dst, size_t bytes);

/* Plan: grid, cell size, physics and which projection the operator uses.
 * Replaces proj.scalar / proj.elasticity (ffthompy/projections.py:9-112, 114-267)
 * and the DFT pair FN, FiN (ffthompy/applications.py:33-37, ffthompy/tensors/operators.py:14-64):
 * G(xi) is never materialised, it is evaluated from the frequency index inside the FFT pass.
 * dim = 2 or 3, N[dim] grid points per axis, Y[dim] cell size. Any length, as numpy.fft / pocketfft
 * take (ffthompy/tensors/fft.py:39-43): lengths whose prime factors are <= 13 use radix-2..13
 * stages, others Bluestein's chirp convolution on a 7-smooth length >= 2N-1 (inside the same
 * LDS kernels; FH_ERR_UNSUPPORTED only if a row of that length does not fit a workgroup's LDS). */
int fh_plan_create(int dim, const int* N, const double* Y, int physics, int primaldual,
                   fh_plan** out);
int fh_plan_destroy(fh_plan* plan);
/* Components of a field: dim (scalar) or dim*(dim+1)/2 (elasticity, Mandel). */
int fh_plan_ncomp(const fh_plan* plan);
/* Switch the projection between G1 (primal) and G2 (dual): applications.py:53-56. */
int fh_plan_set_primaldual(fh_plan* plan, int primaldual);
/* generic = 1 forces the runtime-radix kernels even where a compile-time specialised
 * schedule exists (lengths 5, 9, 15, 25, 27, 45, 63, 75, 81, 99, 105, 125, 135, 225, 243); used to test
 * one path against the other. No reference counterpart. */
int fh_plan_use_generic(fh_plan* plan, int generic);

/* Iteration variant of the mixed-precision CG (FH_SOLVE_MIXED): fused = 1 (default) runs
 * each single-precision iteration as rows-forward (direction update, pending solution step,
 * a p, partial <p, a p>), strided passes, rows-inverse (r -= alpha G(a p), partial <r, r>),
 * using <p, a p> = <p, G(a p)> for p in the range of G; it applies to isotropic
 * coefficients on the fixed-length kernels and falls back otherwise. fused = 0 keeps the
 * reference's operator-form denominator <p, Afun(p)> (ffthompy/general/solver.py:125-126).
 * fused = 1 (default): on grids of at most 12288 rows (n0 * n1) alpha and beta are formed inside
 * the passes (each forward-pass workgroup sums the inverse pass's partials in one fixed order,
 * block 0 of the first strided pass finalises the state) instead of by two one-workgroup kernels
 * per iteration; fused = 2 keeps those kernels. Plain FP64 solves always use the operator form. */
int fh_plan_set_cg_variant(fh_plan* plan, int fused);

/* Mixed-precision CG with the fused iteration: enable = 1 (default) replays batches of
 * iterations from a HIP graph captured on the plan's stream (kernel nodes only, rebuilt when a
 * work buffer moves); 0 launches every kernel directly. Same arithmetic either way.
 * fh_last_solve_graph_batches: batches of the last solve that ran from the graph.
 * No reference counterpart (launch mechanics). */
int fh_plan_use_graphs(fh_plan* plan, int enable);

/* Residual log of the fp64 CG: every >= 1 makes fh_cg_solve evaluate ||b - A x|| (grid-mean norm,
 * one extra fp64 operator application) at x0 and after every every-th iteration and the last one:
 * the reference's CallBack, which evaluates B - A(x) after each iteration (every = 1),
 * ffthompy/general/solver_pp.py:13-24, installed at ffthompy/applications.py:71-72. 0 (default): off.
 * Mixed-precision solves reject a plan with the log on. fh_last_residual_log: the (iteration, norm)
 * pairs of the last solve, at most cap of them, *count = how many there are. */
int fh_plan_set_residual_log(fh_plan* plan, int every);
int fh_last_residual_log(const fh_plan* plan, int* kits, double* norms, int cap, int* count);
int fh_last_solve_graph_batches(const fh_plan* plan, int* n);

/* Host callback on the iterate: the reference's callback(x) after every CG iteration
 * (ffthompy/general/solver.py:134-136) and after every Richardson / Chebyshev step
 * (solver.py:75-76, 277-278). With fn set and every >= 1, fh_cg_solve (fp64) and fh_iter_solve
 * call fn(user, kit, x) after every every-th iteration and after the last one, with x the current
 * device iterate (fp64, the solve's own x buffer: read it, do not free or modify it); the solve waits
 * for fn. A nonzero return stops the solve with FH_ERR_STATE. fn = NULL or every = 0: off. Mixed
 * precision solves reject a plan with a callback (their iterate is split over two precisions). */
typedef int (*fh_iterate_fn)(void* user, int kit, const void* x);
int fh_plan_set_iterate_callback(fh_plan* plan, fh_iterate_fn fn, void* user, int every);

/* Coefficient A(x) of the operator, host array in origin-0 voxel order.
 * Replaces Material.get_A_GaNi (ffthompy/materials.py:115-123) as the right factor of
 * Afun = [[GN, A]] (ffthompy/applications.py:58). For the dual problem pass the inverse
 * coefficient, as the reference does (materials.py:121-122).
 *  FH_COEF_ISO : n values, A = a(x) I
 *  FH_COEF_LAME: 2n values (lambda then mu), Mandel C = lambda m m^T + 2 mu I (elasticity)
 *  FH_COEF_FULL: ncomp*ncomp*n values, entry (i,j) at [(i*ncomp+j)*n + voxel] */
int fh_set_coeff(fh_plan* plan, int kind, const double* host_vals);

/* Restrict the projection to |k_a| <= band[a] (dim entries; NULL lifts the restriction): G
 * built on the solution grid N and zero-padded to a finer grid, hG.enlarge(Nbar) of the Ga
 * solve (ffthompy/applications.py:28-31, ffthompy/tensors/objects.py:428-467), with
 * band[a] = (N_a - 1) / 2 (even N: its Nyquist mode is already null). scale multiplies the
 * operator G A (fh_matvec and the CG), not the bare projection (fh_project): the reference's
 * enlarge converts the 'r'-form hG through set_fft_form (objects.py:135-167), which leaves it
 * multiplied by prod(Nbar)/prod(N); passing that factor reproduces its Afun, residual norms and
 * iteration counts (the solutions do not depend on it). */
int fh_plan_set_band(fh_plan* plan, const int* band, double scale);

/* Coefficient evaluated on the device at the grid nodes from a material defined by inclusions:
 * Material.get_A_GaNi (ffthompy/materials.py:115-123) with evaluate / get_topologies
 * (materials.py:171-301): nodes x = Y k / N on the centred index set, each inclusion summed over
 * its periodic copies in the neighbouring cells (box: -a/2 < z <= a/2 per axis; ball: |z| < d/2;
 * pyramid: prod max(1 - |z|/(a/2), 0) with copies shifted by 2Y as in the reference; 'all' = 1;
 * 'otherwise' = 1 - sum of the others, negative -> "Overlapping inclusions!"), A = sum_ii v_ii t_ii,
 * for the dual problem (invert = 1) the pointwise inverse (materials.py:121-122); stored in origin-0
 * voxel order as the plan's coefficient. Inclusions as fh_ga_set_inclusions (kinds, 3 params and 3
 * position values per inclusion, problem axes first), vals: per inclusion the phase coefficient in
 * layout kind. The nodal expressions are evaluated without fused multiply-adds, so membership
 * decisions equal the reference's. */
int fh_set_coeff_inclusions(fh_plan* plan, int nincl, const int* kinds, const double* params,
                            const double* positions, int kind, const double* vals, int invert);
/* fh_set_coeff from a device fp64 array of the same layout (e.g. a field made by fh_grf). */
int fh_set_coeff_device(fh_plan* plan, int kind, const void* dev_vals);
/* The plan's fp64 coefficient to host (layout of fh_set_coeff). No reference counterpart. */
int fh_coeff_to_host(const fh_plan* plan, double* host_out);
/* Periodic Gaussian random field by spectral synthesis on the plan grid: out = irfftn(rfftn(white)
 * s exp(-pi^2 l^2 |xi|^2)), xi_a = k_a / Y_a, s = std / sqrt(mean of the squared filter) (the
 * field's pointwise variance is std^2 for white noise of unit variance); exponentiate = 1 returns
 * exp of it (a lognormal conductivity). white, out: device fp64, prod(N) values. The reference
 * has no random fields; this generates the inputs of batched and differentiable workloads. */
int fh_grf(fh_plan* plan, const void* white, double corr_len, double std, int exponentiate, void* out);

/* y = F^-1 G(xi) F (A(x) x): the matvec Afun(x) (ffthompy/tensors/operators.py:136-144)
 * in fp64 or fp32 arithmetic (dtype of both fields). */
int fh_matvec(fh_plan* plan, int dtype, const void* x, void* y);

/* y = F^-1 G(xi) F x: the projection operator G1N / G2N (ffthompy/applications.py:36-37). */
int fh_project(fh_plan* plan, int dtype, const void* x, void* y);

/* Forward / inverse real FFT of an ncomp-component field on the plan grid:
 * DFT.__call__ with fft_form 'r' (ffthompy/tensors/operators.py:49-58,
 * ffthompy/tensors/fft.py:39-43). rfftn is unnormalised, irfftn normalised by 1/n. */
int fh_rfftn(fh_plan* plan, int dtype, int ncomp, const void* x, void* xhat);
int fh_irfftn(fh_plan* plan, int dtype, int ncomp, const void* xhat, void* x);

/* out_hat = G(xi) in_hat on 'r'-form spectra of plan ncomp components: the einsum
 * 'ij...,j...->i...' of the materialised hG with a spectrum
 * (ffthompy/tensors/objects.py:223-232, 599-604). */
int fh_apply_projection_hat(fh_plan* plan, int dtype, const void* in_hat, void* out_hat);

/* Conjugate gradients on Afun x = b: CG of ffthompy/general/solver.py:80-139 with the
 * reference's absolute stop sqrt(<r,r>) <= tol, <u,v> = sum(u*v)/prod(N)
 * (ffthompy/tensors/objects.py:634-635). b and x are fp64 fields; x holds x0 on entry.
 * mode FH_SOLVE_FP64: plain fp64 CG (the reference's recurrence and stop rule); with x0 = 0 the
 *   converged x is projected once, x <- G x, removing the component outside the range of G that
 *   the recurrence's rounding accumulates (it would enter the energies at first order and can put
 *   the GaNi lower bound above the upper one by ~1e-14); with another x0 x is returned as iterated.
 * mode FH_SOLVE_MIXED: fp32 CG with fp64 reliable updates: whenever the fp32 residual fell by
 *   `delta` or claims convergence, the fp32 correction c is projected in fp64 and added,
 *   x <- x + G c, the true residual b - A x is recomputed in fp64 and the search direction
 *   restarts from it (p = r); the stop test always uses the fp64 true residual, so norm_res is
 *   the residual of the returned x. The projection removes the part of c outside the range of G
 *   that fp32 rounding leaves (about 1e-5 of c); x0 keeps its own part, as in mode FH_SOLVE_FP64.
 * kit, norm_res: iterations and final residual norm, as info['kit'], info['norm_res'].
 * res_hist (host, nullable): sqrt(<r,r>) after every iteration (entry 0: initial),
 * hist_len entries at most. cg_status: FH_CG_OK or FH_CG_BREAKDOWN (<p,Ap> <= 0). */
int fh_cg_solve(fh_plan* plan, int mode, const void* b, void* x, double tol, int maxiter,
                double delta, int* kit, double* norm_res, double* res_hist, int hist_len,
                int* cg_status);

/* Stationary iterations on Afun x = b, fp64, on the device (x holds x0 on entry):
 *  FH_ITER_RICHARDSON (richardson, ffthompy/general/solver.py:63-77): coefs[0] = alpha,
 *    x <- x + (b - A x) / alpha; stops once sqrt(<r,r>) <= tol for the residual r = b - A x taken
 *    before that step, or after maxiter steps; *norm_res is that norm (1e15 if no step ran).
 *  FH_ITER_CHEBYSHEV (cheby2TERM, solver.py:206-285): coefs[0..1] = eigenvalue range [min, max] of
 *    Afun on the range of G; two-term recurrence with the reference's coefficients; returns at once
 *    (kit 0) if |b - A x0| / |b| < tol, else stops once |r| / |r0| <= tol or after maxiter
 *    iterations (the reference's par['maxit']); *norm_res is that relative norm.
 * res_hist (nullable): entry 0 the value before the first iteration, entry k after iteration k. */
int fh_iter_solve(fh_plan* plan, int method, const void* b, void* x, const double* coefs, double tol,
                  int maxiter, int* kit, double* norm_res, double* res_hist, int hist_len);

/* Potential formulation of the scalar problem: the solve of homog_GaNi_full_potential and, on a
 * banded plan (fh_plan_set_band) carrying the Ga coefficient (fh_ga_set_plan_coeff), of
 * homog_Ga_full_potential (ffthompy/tensorsLowRank/homogenisation.py:43-128; preconditioner
 * get_preconditioner, homogenisation.py:296-300). Unknown: the rfftn spectrum of a scalar
 * potential; CG (x0 = 0, the reference's stop rule on the inner product of 'r' spectra,
 * ffthompy/tensors/objects.py:618-631) on  x -> P (-div) F A F^-1 grad (P x)  with right-hand
 * side P div F(A E), P = 1/|2 pi xi| (1 at xi = 0). On a banded plan grad, div and P act on
 * |k_a| <= band[a] only (the reference's enlarge / project between N and 2N-1).
 * mode FH_SOLVE_FP64 (the reference's precision) or FH_SOLVE_MIXED (fp32 iterations, fp64 right-hand
 * side and iterate, reliable updates when the residual drops by delta, as fh_cg_solve).
 * E: ncomp macroscopic gradient values (host). Outputs (device, fp64): Fu = P x, the potential's
 * spectrum ((n0, n1, n2/2+1) complex, the plan's grid); X (nullable) = F^-1 grad Fu, ncomp real
 * components. Plan of scalar physics; odd grid unless banded. res_hist as fh_cg_solve. */
int fh_potential_solve(fh_plan* plan, int mode, const double* E, double tol, int maxiter, double delta, void* Fu,
                       void* X, int* kit, double* norm_res, double* res_hist, int hist_len, int* cg_status);

/* AH[i][j] = <A e_i, e_j> for nf = ncomp fields e_i (fp64, macroscopic part included),
 * double-double accumulated: assembly_matrix (ffthompy/postprocess.py:53-70). */
int fh_energy_matrix(fh_plan* plan, const void* const* fields, int nf, double* AH);

/* fh_energy_matrix without a plan and without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code inside a stream capture,
 * tight serving loops, hosts that forbid hipMalloc). nvox voxels; kind FH_COEF_ISO / LAME / FULL
 * with coef a device fp64 array of that layout (as fh_set_coeff); nf = 2, 3 or 6 fp64 device fields of
 * nf components each (nf > 6 is rejected: the scratch holds the partial sums of at most six fields);
 * AH host, nf x nf. Same kernels and result as fh_energy_matrix. Its partial sums live in one
 * static device array of 576.5 KiB per GPU (held until the process ends, located at fh_init);
 * every call on a GPU shares it, so overlapping calls would race: the library serialises its own
 * calls with a lock, and the call returns after its GPU work (null stream) is done. */
int fh_energy_matrix_noalloc(long long nvox, int kind, const void* coef, const void* const* fields, int nf,
                             double* AH);

/* Gradient of a loss L(AH) w.r.t. the coefficient, for the minimisers e_i (fp64 fields,
 * macroscopic part included) of the solve: with W = dL/dAH (host, nf x nf) and the fields
 * held fixed (envelope theorem, the e_i are stationary points of the energy),
 *   FH_COEF_ISO : grad[x] = chain[x] * sum_ij W_ij e_i(x).e_j(x) / n        (n values)
 *   FH_COEF_LAME: grad[x] = sum_ij W_ij tr e_i tr e_j / n,
 *                 grad[n+x] = sum_ij W_ij 2 e_i.e_j / n                        (2n values)
 * chain: optional device per-voxel multiplier (ISO only, e.g. da/dtheta). grad: device fp64.
 * No reference counterpart (the reference has no derivatives); it differentiates
 * assembly_matrix (ffthompy/postprocess.py:53-70) through the solve. */
int fh_energy_gradient(fh_plan* plan, const void* const* fields, int nf, const double* W, const void* chain,
                       void* grad);

/* Adjoint gradient of a loss on the solution fields, isotropic coefficient:
 * grad[x] = -chain[x] * sum_i adjoints_i(x) . fields_i(x)   (chain nullable: 1),
 * where fields_i = E_i + x_i are the unit-load solutions and adjoints_i solve
 * G(a adj_i) = G dL/dfields_i (the same CG, fh_cg_solve on fh_project(dL/dfields_i)).
 * nf fields and adjoints of ncomp components, fp64 device pointers; grad: n fp64.
 * No reference counterpart (the reference has no derivatives). */
int fh_field_gradient(fh_plan* plan, const void* const* fields, const void* const* adjoints, int nf,
                      const void* chain, void* grad);

/* Helpers on fp64/fp32 device vectors (Tensor arithmetic, ffthompy/tensors/objects.py:188-245):
 * y = a x + b y; out = sum(x*y) (caller divides by prod(N) for the grid inner product,
 * ffthompy/tensors/objects.py:618-636); y_c += E_c on an ncomp-component field
 * (Tensor.add_mean, objects.py:274-287). */
int fh_axpby(int dtype, size_t n, double a, const void* x, double b, void* y);
int fh_dot(int dtype, size_t n, const void* x, const void* y, double* out);
int fh_add_constant(int dtype, size_t nvox, int ncomp, const double* E, void* y);

/* Spectrum forms and grid changes of device fields (no plan needed; dim = 1, 2 or 3 and N the grid,
 * count fields of the given layout one after another; `in` and `out` must not overlap).
 * FH_FORM_R is the 'r' half spectrum (unnormalised rfftn), FH_FORM_FULL the reference's fft_form 0
 * (fftn / prod(N)), FH_FORM_CENTRED its 'c' (form 0 with fftshift). */
enum { FH_FORM_R = 0, FH_FORM_FULL = 1, FH_FORM_CENTRED = 2 };
/* full = scale * Hermitian completion of half, fftshifted when centred
 * (Tensor.set_fft_form from 'r', ffthompy/tensors/objects.py:142-154). */
int fh_spectrum_half_to_full(int dtype, int dim, const int* N, long long count, const void* half, void* full,
                             double scale, int centred);
/* half = scale * the first N[-1]//2+1 entries of the last axis of full (ifftshifted first when
 * centred; Tensor.set_fft_form to 'r', objects.py:155-163); hermitian = 1 takes the Hermitian part
 * (full[k] + conj(full[-k])) / 2, so that irfftn of half is the real part of ifftn of full
 * (the inverse transforms of forms 0 and 'c', ffthompy/tensors/fft.py:24-37). */
int fh_spectrum_full_to_half(int dtype, int dim, const int* N, long long count, const void* full, void* half,
                             double scale, int centred, int hermitian);
/* fftshift (inverse = 0) or ifftshift (inverse = 1) over the grid axes of count fields of
 * elem_bytes-byte elements (4, 8 or 16): forms 0 <-> 'c' of a spectrum and Tensor.shift of a real
 * field (objects.py:164-186). */
int fh_fftshift(int elem_bytes, int dim, const int* N, long long count, const void* in, void* out, int inverse);
/* Centred spectra (FH_FORM_CENTRED) from grid N to grid M: enlarge = 1 zero-pads, splitting the
 * Nyquist entries of even axes (Tensor.enlarge, objects.py:428-467; needs M >= N + [N even] on
 * every axis); enlarge = 0 crops (Tensor.decrease, objects.py:469-486; needs M <= N). */
int fh_spectrum_resize(int dtype, int dim, const int* N, const int* M, long long count, const void* in, void* out,
                       int enlarge);
/* out = G_which in * (which 1: compatible, 2: equilibrated fields) on ncomp component spectra of the
 * given form (scalar ncomp = dim, elasticity Mandel dim(dim+1)/2; dim 2 or 3): the projections of
 * ffthompy/projections.py:9-112 (scalar) and 114-267 (elasticity, the sums G1h + G1s, G2h + G2s)
 * for any fft_form, null_nyq = 1 for NyqNul=True (Nyquist modes of even axes set to zero),
 * 0 for NyqNul=False (G at frequency -N/2 there). */
int fh_project_spectrum(int dtype, int dim, const int* N, const double* Y, int physics, int which, int form,
                        int null_nyq, const void* in, void* out);
/* Pointwise inverse of a (k, k) + grid matrix field, 1 <= k <= 6, by Gauss-Jordan elimination
 * without pivoting (Tensor.inv / get_inverse, objects.py:247-251, ffthompy/trigpol.py:120-159). */
int fh_pointwise_inverse(int dtype, int k, long long nvox, const void* in, void* out);
/* out = grad in: the dim-component 'r' spectrum (dim,) + half grid of 2 pi i xi_j times the scalar 'r'
 * spectrum in, xi_j = k_j / Y_j with the Nyquist index of an even axis at k = -N/2 (the reference's
 * grad of a scalar Tensor, ffthompy/tensors/operators.py:227-260; frequencies
 * ffthompy/trigpol.py:27-40). dim 1 to 3. */
int fh_spectrum_grad(int dtype, int dim, const int* N, const double* Y, const void* in, void* out);
/* Pointwise products of device Tensors (Tensor.__mul__ for multype 21, 42, 00, 'grad' and 'div',
 * ffthompy/tensors/objects.py:223-245 and the einsums of objects.py:599-604), nvox grid points,
 * component-major, real (is_complex 0) or complex values of dtype's precision:
 *   mode 0: out[a][v] = sum_b X[a][b][v] * Y[b][v]   (X: A x B components, Y: B, out: A)
 *   mode 1: out[a][v] = X[a][v] * Y[a][v]             (A components each; B unused)
 * out must not overlap X or Y; 1 <= A, B <= 4096. */
int fh_pointwise_product(int dtype, int is_complex, int mode, int A, int B, long long nvox, const void* X,
                         const void* Y, void* out);

/* Device-side wall time of the last fh_matvec / fh_cg_solve on this plan, in ms
 * (HIP events on the plan stream), and the number of fp64 residual recomputations
 * of the last mixed-precision solve. Replaces the reference's Timer around a solve
 * (ffthompy/general/base.py:63-81, ffthompy/general/solver.py:13,58-59). */
int fh_last_timing(const fh_plan* plan, double* matvec_ms, double* solve_ms);
int fh_last_solve_updates(const fh_plan* plan, int* n_updates);

/* Timing harness for the matvec passes (rows forward, strided passes, rows inverse): nrep
 * matvecs on the plan stream, each kernel launched with its own start/stop HIP events
 * (hipExtLaunchKernel: the dispatch's timestamps); returns each pass's average duration in
 * ms and the algorithmic HBM bytes it moves per launch. Needs the fixed-length kernels
 * (FH_ERR_UNSUPPORTED otherwise). y holds the matvec result afterwards. At most max_kernels
 * entries; *nkernels receives the count. No reference counterpart (measurement only). */
int fh_matvec_profile(fh_plan* plan, int dtype, const void* x, void* y, int nrep, int max_kernels,
                      double* kernel_ms, double* kernel_bytes, int* nkernels);

/* Schedule of the fixed-length kernels, per arithmetic (dtype FH_F64 / FH_F32). proj_variant
 * (0..3) sets the frequency points per tile of the projection pass (forward FFT along the
 * outermost axis, G(xi), inverse FFT), which also sets the threads per workgroup; cols_variant (0..1) the columns per tile of the other strided
 * passes (3-D); rows_variant (0..2) the grid rows per workgroup of the contiguous-axis passes
 * (variant 1: 3/2 as many in single precision, 1/2 as many in double; variant 2: 3/4 as many;
 * scalar plans, isotropic or no coefficient). Every variant computes the same values (the same
 * transforms per pencil); only speed differs. Variant 0 is the default.
 * fh_plan_autotune times every variant on this plan (3 interleaved rounds of nrep matvecs of x
 * into y per variant, each pass timed by its own HIP events, as fh_matvec_profile; a variant's
 * time is its best round; in single precision, where the mixed solve runs the fused CG iteration,
 * the row counts are timed on that iteration's two row passes, on the plan's work vectors and a
 * placeholder CG state that the next solve replaces), keeps the fastest of each pass and
 * returns the times in ms (proj_ms[4], cols_ms[2], rows_ms[3]; any may be NULL) and the choice.
 * FH_ERR_UNSUPPORTED where the plan runs the runtime-radix kernels. No reference counterpart (the
 * reference calls NumPy's FFT). */
int fh_plan_set_schedule(fh_plan* plan, int dtype, int proj_variant, int cols_variant, int rows_variant);
int fh_plan_get_schedule(const fh_plan* plan, int dtype, int* proj_variant, int* cols_variant, int* rows_variant);
int fh_plan_autotune(fh_plan* plan, int dtype, const void* x, void* y, int nrep, double* proj_ms, double* cols_ms,
                     double* rows_ms, int* proj_variant, int* cols_variant, int* rows_variant);

/* ---- Batched 2-D scalar homogenization (many samples, uneven grids, one launch) ----------
 * Each sample s is the reference's scalar GaNi problem on its own grid
 * (ffthompy/applications.py:24-90 run once per sample): isotropic conductivity a_s(x) on a
 * grid dims[2s] x dims[2s+1] (axis lengths 27, 45, 63, 81, 105 or 135) of cell
 * Y[2s] x Y[2s+1] (NULL: unit cells), unit loads E = e1, e2, CG (ffthompy/general/solver.py:80-139)
 * on the primal (G1, a) and dual (G2, 1/a) problems, and the bounds of
 * ffthompy/postprocess.py:53-70: upper = <a e_i, e_j>, lower = inv(<a^-1 e_i, e_j>).
 * All samples are solved by one persistent kernel (one workgroup per sample and primal/dual
 * task, CG state on the device), launched from a captured HIP graph. */
int fh_batch_create(int nsamples, const int* dims, const double* Y, fh_batch** out);
int fh_batch_destroy(fh_batch* batch);
/* a: host fp64, the samples' voxel values concatenated, each in C order with origin at
 * index 0 (the layout of Material.get_A_GaNi, ffthompy/materials.py:115-123). */
int fh_batch_set_coeff(fh_batch* batch, const double* a);
/* mode FH_SOLVE_FP64 or FH_SOLVE_MIXED (fp32 inner CG, fp64 reliable updates as fh_cg_solve:
 * each fp32 correction projected in fp64 before it is added; fp64 energies). Host outputs (any may be NULL): upper[4s..4s+3], lower[4s..],
 * kit[4s..4s+3] = iterations of (primal e1, primal e2, dual e1, dual e2), status[s] =
 * FH_CG_OK or FH_CG_BREAKDOWN. Stop rule as fh_cg_solve. */
int fh_batch_solve(fh_batch* batch, int mode, double tol, int maxiter, double delta, double* upper,
                   double* lower, int* kit, int* status);
/* Schedule of the batched solve kernel: variant 0 or 1 (radix order of the row and column
 * transforms of every length: forward/reversed or reversed/forward) and the task order (0
 * largest grids first, 1 sample order). fh_batch_autotune times every (variant, order) on
 * the loaded batch with `iters` CG iterations per task (times_ms[order * 2 + variant], may be
 * NULL) and keeps the fastest; fh_batch_solve runs it on first use unless a schedule was set.
 * No reference counterpart (the reference has no batched or device solve). */
int fh_batch_set_schedule(fh_batch* batch, int variant, int order_mode);
int fh_batch_autotune(fh_batch* batch, int mode, double delta, int iters, double* times_ms, int* variant,
                      int* order_mode);
/* Of the last solve: fp64 residual evaluations and final sqrt(<r,r>) per (sample, primal/dual,
 * load) (4 per sample, host arrays), and the device time of the whole batch in ms. */
int fh_batch_stats(const fh_batch* batch, int* n_updates, double* norm_res, double* solve_ms);

/* ---- Ga post-processing: bounds from the exactly integrated coefficient ---------------------
 * The 'Ga' entries of ffthompy/postprocess.py:20-41: AH_ij = <A_Ga P(e_i), P(e_j)> on the grid
 * Nbar = 2N-1, with A_Ga the exactly integrated coefficient (Material.get_A_Ga,
 * ffthompy/materials.py:54-113) and P(e) the trigonometric interpolation of a GaNi solution
 * (Tensor.project, ffthompy/tensors/objects.py:488-511, inside assembly_matrix,
 * ffthompy/postprocess.py:53-70); the dual matrix is the inverse (postprocess.py:43-46).
 * The quadrature grid is Nbar where its prime factors are <= 13; for odd N with a larger prime
 * in 2N-1 it is the next such length (the integrand is a trigonometric polynomial of degree
 * <= 2N-2 per axis, so every grid of at least 2N-1 points integrates it exactly); for even N it is
 * always 2N-1 (transformed with Bluestein's algorithm where needed). grid (nullable, dim entries):
 * an explicit quadrature grid of at least 2N-1 points per axis, exact for fields with
 * frequencies below N/2 (the solutions of a Ga solve); for GaNi solutions of even N only
 * 2N-1 reproduces the reference. */
int fh_ga_create(fh_plan* coarse, const int* grid, fh_ga** out);
int fh_ga_destroy(fh_ga* ga);
/* Quadrature grid actually used, dim entries. */
int fh_ga_grid(const fh_ga* ga, int* M);
/* Coefficient from inclusions (get_A_Ga with order None, materials.py:54-73; shape functions
 * materials.py:125-169; phase and weights materials.py:317-425): kinds[ii] 1 box ('cube'/'square'),
 * 2 ball ('ball'/'circle'; params[3ii] = diameter), 3 pyramid, 4 'all', 5 'otherwise' (last);
 * params/positions: 3 doubles per inclusion (problem axes first); vals: the phase coefficient
 * per inclusion in layout kind (FH_COEF_ISO 1, FH_COEF_LAME 2, FH_COEF_FULL ncomp^2 values),
 * already inverted for the dual problem (materials.py:73). */
int fh_ga_set_inclusions(fh_ga* ga, int nincl, const int* kinds, const double* params, const double* positions,
                         int kind, const double* vals);
/* Coefficient from nodal values on a grid P (get_A_Ga with order 0 or 1 and P,
 * materials.py:74-113): vals host fp64, component-major, each a P-grid array in origin-0
 * voxel order (dual: the pointwise inverse), layout kind as above. */
int fh_ga_set_grid_coeff(fh_ga* ga, int order, const int* P, int kind, const double* vals);
/* Copy A_Ga into a plan on the quadrature grid (same physics) as its coefficient: the
 * operator of a Ga solve (Material.get_A_Ga(Nbar), ffthompy/applications.py:48-51). */
int fh_ga_set_plan_coeff(const fh_ga* ga, fh_plan* plan);
/* A_Ga on the quadrature grid to host (ncoef arrays of prod(M) values, origin 0). */
int fh_ga_coefficient(const fh_ga* ga, double* host_out);
/* P(x) on the quadrature grid: x a device fp64 field of the coarse plan (ncomp components),
 * out a device fp64 field of ncomp * prod(M) values (Tensor.project, objects.py:488-511). */
int fh_ga_interpolate(fh_ga* ga, const void* x, void* out);
/* The reverse: a field on the quadrature grid (ncomp * prod(M) values) restricted to the coarse
 * grid by cropping its spectrum (Tensor.project to a smaller grid = decrease,
 * ffthompy/tensors/objects.py:469-511); exact for fields with frequencies below N/2, such as the
 * solutions of a Ga solve. out: device fp64, ncomp * prod(N) values. */
int fh_ga_restrict(fh_ga* ga, const void* x, void* out);
/* AH[i][j] = <A_Ga P(e_i), P(e_j)> for nf = ncomp fp64 device fields of the coarse plan. */
int fh_ga_energy(fh_ga* ga, const void* const* fields, int nf, double* AH);
/* The same for fields already on the quadrature grid (solutions of a Ga solve on that grid):
 * assembly_matrix without projection (ffthompy/postprocess.py:53-70). */
int fh_ga_energy_on_grid(fh_ga* ga, const void* const* fields, int nf, double* AH);

/* Wait for all GPU work of this process. */
int fh_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* FFTHOM_H */
