"""Pin the CPU oracle before trusting it (CPU only).

Each check is chosen so that a plausible mistake in the restatement -- a dropped term, a
wrong sign, a swapped index or operand -- fails one of them:
  * the reference's golden vectors (tests/pimc/explicit_data): constants, folding,
    surface weights, Gaussian covariance, O-matrices, M, g, rho;
  * the Julia sum-over-states values in tests/test_models/*/sos_B80.json (closed-form
    Z_rho, E_rho, Cv_rho of the sampling model);
  * the uncoupled special case g == rho per sample, and its finite-difference energy;
  * brute-force loops for V (including quadratic couplings no fixture contains);
  * M . expm(+tau V) = I;
  * invariance of g/rho under the S12 scaling;
  * the reference's stats known-answer tests.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg

from conftest import EXPLICIT, TEST_MODELS, golden
from oracle import pimc_oracle as po
from oracle import stats_oracle as so


@pytest.fixture(scope="module")
def setup():
    A, N, E, w, L, Q = po.read_model_json(os.path.join(EXPLICIT, "coupled_model.json"))
    Ar, Nr, Er, wr, Lr, _ = po.read_model_json(os.path.join(EXPLICIT, "sampling_model.json"),
                                               diagonal=True)
    return po.Setup((E, w, L, Q), (Er, wr, Lr), 300.0, 5)


def test_beta():
    # tests/pimc/test_pimc_explicit_example.py:36
    assert po.beta_of_T(300.0) == 38.68174020133669


def test_folding_and_constants(setup):
    s = setup
    assert np.allclose(s.vib["delta"], golden("coupled_Edeltas.npy"), rtol=1e-13, atol=0)
    assert np.allclose(s.rho["delta"], golden("sampling_Edeltas.npy"), rtol=1e-13, atol=0)
    assert np.allclose(s.vib["shift"], golden("coupled_ds.npy"), rtol=1e-13, atol=0)
    assert np.allclose(s.rho["shift"], golden("sampling_ds.npy"), rtol=1e-13, atol=0)
    assert np.allclose(s.c_vib[0], golden("coupled_cosh.npy") / golden("coupled_sinh.npy"), rtol=1e-13)
    assert np.allclose(s.c_vib[1], 1 / golden("coupled_sinh.npy"), rtol=1e-13)
    assert np.allclose(s.c_rho[0], golden("sampling_cosh.npy") / golden("sampling_sinh.npy"), rtol=1e-13)
    w = golden("sampling_surface_weights.npy")
    assert np.allclose(s.rho["weight"], w / w.sum(), rtol=1e-13)
    # both models' means are their state shifts, identical over beads
    assert np.allclose(golden("sampling_means.npy")[..., 0], s.rho["shift"], rtol=1e-13)
    assert np.allclose(golden("coupled_means.npy")[..., 0], s.vib["shift"], rtol=1e-13)


def test_gaussian_covariance(setup):
    s = setup
    cov = np.linalg.inv(2 * s.c_rho[0][..., None, None] * np.eye(5)
                        - s.c_rho[1][..., None, None] * s.C[None, None])
    assert np.allclose(cov, golden("sampling_covariance.npy"), rtol=1e-11, atol=0)
    # the sampler's per-normal-mode variances are the eigenvalues of the same covariance
    U = s.circ_evecs
    for a in range(2):
        for j in range(2):
            d = U.T @ cov[a, j] @ U
            assert np.allclose(np.diag(d), s.sigma[a, j] ** 2, rtol=1e-11)
            assert np.allclose(d - np.diag(np.diag(d)), 0, atol=1e-12)


def test_explicit_kernels(setup):
    q = golden("samples.npy")[:, 0].copy()
    r = po.evaluate(setup, q, scale=False, keep=True)
    nz = golden("rho_oMat.npy") != 0
    assert np.max(np.abs(r["O_rho"][nz] - golden("rho_oMat.npy")[nz]) / np.abs(golden("rho_oMat.npy")[nz])) < 1e-11
    nz = golden("vib_oMat.npy") != 0
    assert np.max(np.abs(r["O_vib"][nz] - golden("vib_oMat.npy")[nz]) / np.abs(golden("vib_oMat.npy")[nz])) < 1e-11
    assert np.all(r["O_rho"][~(golden("rho_oMat.npy") != 0)] == 0)
    assert np.max(np.abs(r["M"] - golden("vib_mMat.npy"))) < 1e-14
    assert np.allclose(r["s_g"], golden("numerator(g).npy"), rtol=1e-12, atol=0)
    assert np.allclose(r["s_rho"], golden("denominator(rho).npy"), rtol=1e-12, atol=0)


def test_scaling_invariance_and_log_scale(setup):
    q = golden("samples.npy")[:, 0].copy()
    a = po.evaluate(setup, q, scale=False, pm=True)
    b = po.evaluate(setup, q, scale=True, pm=True)
    for k in ("s_g", "s_gP", "s_gM"):
        assert np.allclose(a[k] / a["s_rho"], b[k] / b["s_rho"], rtol=1e-13)
        assert np.allclose(b[k] * np.exp(b["log_scale"]), a[k], rtol=1e-12)


def brute_force_V(q_all, E, L, Q):
    B, A, N, P = q_all.shape
    V = np.zeros((B, P, A, A))
    for s in range(B):
        for p in range(P):
            for b in range(A):
                for c in range(A):
                    v = E[b, c]
                    for j in range(N):
                        v += L[j, b, c] * q_all[s, b, j, p]
                        for k in range(N):
                            v += 0.5 * Q[j, k, b, c] * q_all[s, b, j, p] * q_all[s, c, k, p]
                    V[s, p, b, c] = v
    return V


def test_V_brute_force_and_expm():
    rng = np.random.default_rng(0)
    A, N, P, B = 3, 3, 4, 2
    E = rng.normal(0, 0.1, (A, A)); E = E + E.T
    L = rng.normal(0, 0.1, (N, A, A)); L = L + L.transpose(0, 2, 1)
    Q = rng.normal(0, 0.05, (N, N, A, A)); Q = Q + Q.transpose(0, 1, 3, 2)
    q = np.empty((B, A, N, P)); q[:] = rng.normal(0, 1, (B, 1, N, P))
    V, lam, U = po.coupling_eig(q, E, L, Q)
    assert np.allclose(V, brute_force_V(q, E, L, Q), rtol=1e-13, atol=1e-14)
    tau = 0.7
    M = po.m_matrix(lam, U, tau)
    for s in range(B):
        for p in range(P):
            assert np.allclose(M[s, p] @ scipy.linalg.expm(tau * V[s, p]), np.eye(A), atol=1e-13)
            assert np.allclose(M[s, p], M[s, p].T, atol=1e-15)


def test_ring_trace_order():
    # a non-commuting pair: the product must be left to right, M then O per bead
    M = np.array([[[[1.0, 2.0], [0.0, 1.0]], [[1.0, 0.0], [3.0, 1.0]]]])
    O = np.array([[[[2.0, 0.0], [0.0, 1.0]], [[1.0, 0.0], [0.0, 5.0]]]])
    expect = np.trace(M[0, 0] @ O[0, 0] @ M[0, 1] @ O[0, 1])
    assert po.ring_trace(M, O)[0] == expect


@pytest.mark.parametrize("ds,rho", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_closed_form_sampling_partition_function(ds, rho):
    """Z_rho, E_rho, Cv_rho against the Julia values stored with the reference's models."""
    base = os.path.join(TEST_MODELS, f"data_set_{ds}", f"rho_{rho}", "parameters")
    with open(os.path.join(base, "sos_B80.json")) as fh:
        sos = json.load(fh)
    A, N, E, w, L, _ = po.read_model_json(os.path.join(base, "sampling_model.json"), diagonal=True)
    beta = sos["beta"][0]
    Z, Em, d2 = po.z_rho((E, w, L), beta)
    assert Z == pytest.approx(sos["Z_sampling"][0], rel=1e-12)
    assert Em == pytest.approx(sos["E_sampling"][0], rel=1e-12)
    # Julia's k_B differs from the reference's in the 9th digit
    assert d2 / (po.BOLTZMANN_EV_PER_K * 300.0 ** 2) == pytest.approx(sos["Cv_sampling"][0], rel=1e-8)
    Zp, _, _ = po.z_rho((E, w, L), beta + po.DELTA_BETA)
    Zm, _, _ = po.z_rho((E, w, L), beta - po.DELTA_BETA)
    assert Zp == pytest.approx(sos["Z_sampling+beta"][0], rel=1e-12)
    assert Zm == pytest.approx(sos["Z_sampling-beta"][0], rel=1e-12)


def test_uncoupled_model_energy():
    """data_set_0 / rho_0: g == rho per sample, and the finite-difference energy of the
    +- dbeta path reproduces E_harmonic of sos_B80.json (the O-matrix is the exact
    harmonic propagator, so there is no Trotter error) within its statistical error."""
    base = os.path.join(TEST_MODELS, "data_set_0")
    A, N, E, w, L, Q = po.read_model_json(os.path.join(base, "parameters", "coupled_model.json"))
    Ar, Nr, Er, wr, Lr, _ = po.read_model_json(os.path.join(base, "rho_0", "parameters", "sampling_model.json"), diagonal=True)
    with open(os.path.join(base, "rho_0", "parameters", "sos_B80.json")) as fh:
        sos = json.load(fh)
    s = po.Setup((E, w, L, Q), (Er, wr, Lr), 300.0, 12)
    q, _ = s.draw_q(np.random.default_rng(0), 20000)
    r = po.evaluate(s, q, pm=True)
    assert np.allclose(r["s_g"], r["s_rho"], rtol=1e-14)
    ratio, d1, d2 = so.basic_terms(po.DELTA_BETA, r["s_rho"], r["s_g"], r["s_gP"], r["s_gM"])
    n = len(ratio)
    jk = [so.jackknife_term(n, t) for t in (ratio, d1, d2)]
    basic = so.estimate_basic(n, 300.0, ratio, d1, d2)
    j = so.estimate_jackknife(n, 300.0, po.DELTA_BETA, basic, *jk)
    assert basic["Z"] == pytest.approx(1.0, rel=1e-13)
    assert abs(j["E"] - sos["E_harmonic"][0]) < 4 * j["E error"], (j, sos["E_harmonic"])


# ---- the reference's stats known-answer tests (tests/stats/test_stats.py:46-117,
#      tests/stats/test_jackknife.py:22-57), re-stated
def test_stats_basic_terms():
    X = 100
    r = so.basic_terms(1.0, np.ones(X), np.ones(X), np.ones(X) + 1, np.ones(X))
    assert np.all(r[0] == 1.0) and np.all(r[1] == 0.5) and np.all(r[2] == 1.0)


def test_stats_alpha_terms():
    X = 100
    r = so.alpha_terms(1.0, np.ones(X), np.ones(X), np.ones(X) + 1, np.ones(X), 2.0, 2.0)
    assert np.all(r[0] == 1.0) and r[1] == 1.0 and r[2] == 4.0


def test_stats_estimate_basic():
    X, T = 100, 300.0
    ret = so.estimate_basic(X, T, np.ones(X), np.ones(X), np.ones(X) + 1)
    kBT = po.BOLTZMANN_EV_PER_K * T ** 2
    assert ret["Z"] == 1.0 and ret["Z error"] == 0.0 and ret["E"] == -1.0
    assert ret["Cv"] == 1.0 / kBT


def test_stats_jackknife():
    X, T = 100, 300.0
    kBT = po.BOLTZMANN_EV_PER_K * T ** 2
    ret = so.estimate_jackknife(X, T, 1.0, {"E": -1.0, "Cv": 1.0 / kBT}, np.ones(X), np.ones(X), np.ones(X) + 1)
    assert ret["E"] == -1.0 and ret["E error"] == 0.0
    assert ret["Cv"] == pytest.approx(1.0 / kBT)
    assert ret["Cv error"] == pytest.approx(2.7755575615628914e-16)


def test_jackknife_of_mean_is_mean():
    x = np.random.default_rng(1).normal(size=50)
    jk = so.jackknife_term(50, x)
    assert np.mean(jk) == pytest.approx(np.mean(x), rel=1e-14)
