"""CPU oracle for the PIMC per-block estimator.

TEST INFRASTRUCTURE ONLY. Nothing in the product package (`pibronic_amd`) may import,
call or link anything under `oracle/`. Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` use it, and only as the checker or as the timed CPU
baseline, never as the thing measured or shipped.

Modules
-------
pimc_oracle   FP64 NumPy restatement of Pibronic's per-block estimator
              (pibronic/pimc/pimc.py:59-692, 1062-1213, 1341-1462; pibronic/constants.py).
stats_oracle  restatement of the block statistics and jackknife
              (pibronic/stats/stats.py:38-130; pibronic/stats/jackknife.py:14-105).
models        seeded synthetic vibronic models for the benchmark configurations and the
              reference's JSON model format (pibronic/vibronic/vibronic_model_io.py:42-73,
              517-649, 749-761).

Pinning (what checks each function; the tests under tests/ implement every item)
------------------------------------------------------------------------------------
* constants, temperature constants, folding, surface weights, Gaussian covariance,
  O-matrices, M-matrices, g, rho (unscaled): the reference's golden vectors in
  tests/golden/explicit_data (A=2, N=2, P=5, B=10).
* closed-form Z_rho / E_rho / Cv_rho: the Julia sum-over-states values held in
  tests/golden/test_models/*/sos_B*.json.
* S12 scaling and the +-dbeta path (s_gP, s_gM): invariance of g/rho under the scaling,
  and the harmonic special case (V = 0, M = I, g == rho per sample) whose finite-difference
  energy must reproduce E_harmonic of sos_B80.json within its statistical error.
* quadratic couplings in V at A > 2: brute-force loops over (j, k) on tiny inputs.
* eigh / M: M . expm(+tau V) = I (scipy.linalg.expm) and symmetry.
* A_rho != A sampling path (the qTempTensor copy): the oracle's rho against
  sum_a F_a^P exp(-1/2 sum_j x^T K_j x) with K_j the inverse of the reference's golden
  sampling covariance (deterministic); and Z_rho * <g/rho> of the reference's Jahn-Teller
  model with its three alternate sampling models (A_rho = 2, 4, 8) agreeing within
  3 sigma (tests/test_closed_form.py).
* log_z_ratio_surface_diagonal (exact Z_g/Z_rho without inter-surface couplings): a
  Gauss-Hermite quadrature of the oracle's own g and rho at P = 3, the identity g == rho
  at Q = 0, and the oracle's harmonic Monte Carlo estimate (tests/test_closed_form.py).
* stats formulas: the reference's known-answer tests (tests/stats/test_stats.py:46-117,
  tests/stats/test_jackknife.py:22-57) re-stated as tests.
"""
