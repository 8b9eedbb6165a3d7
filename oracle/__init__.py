"""CPU oracle: a float64 NumPy restatement of FFTHomPy's Fourier-Galerkin homogenization.

TEST INFRASTRUCTURE ONLY. This package is the checker for the HIP product path
in ``ffthompy_amd``. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only to compute the
expected answer (or, in ``bench.py``, to time the reference-style CPU path
beside the GPU). The product never imports, calls or falls back to it.

Each function cites the reference file:line it restates. What pins each part:

* ``grid``         -- closed forms of the frequency/coordinate sets; golden AH values.
* ``fields``       -- numpy.fft (library routine), DFT matrix closed form, round trips.
* ``projections``  -- closed forms (xi xi^T/|xi|^2, Mandel of the 4th-order Gamma),
                      idempotence and orthogonality, golden AH values.
* ``materials``    -- golden AH values (all inclusion kinds used by the examples),
                      exact-integration == grid-integration identities.
* ``solver``       -- textbook CG, dense brute-force solve on tiny grids, golden
                      iteration counts.
* ``homogenize``   -- golden AH matrices of every reference example problem (1e-9).
* ``potential``    -- golden GaNi / Ga values of the gradient formulation (same Galerkin
                      problems), the constant-coefficient multiplier closed form, symmetry and
                      positivity of the operator.

Parity unpinned (no fixture or invariant fixes the value, documented in the tests):
* ``materials.weights_ball`` in 3-D: the reference uses the 2-D disk formula
  for 3-D balls (ffthompy/materials.py:405-425); restated as-is, not pinned.
"""
