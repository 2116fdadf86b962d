"""Pins the oracle's estimator mean and its A_rho != A sampling path against mathematics.

* log_z_ratio_surface_diagonal (the exact Z_g / Z_rho of a coupled model without
  inter-surface couplings) against a Gauss-Hermite quadrature of the oracle's own g and rho at
  P = 3 (where the oracle is pinned by the reference's golden vectors), against the identity
  g == rho when the model has no quadratic terms, and against the oracle's harmonic Monte
  Carlo estimate.
* The oracle's denominator for a sampling model with another surface count than the coupled
  model (the reference's qTempTensor copy, pibronic/pimc/pimc.py:625-626, 643-644, 1094-1096)
  against sum_a F_a^P exp(-1/2 sum_j x_aj^T K_j x_aj), K_j the inverse of the reference's
  golden sampling covariance (tests/pimc/explicit_data/sampling_covariance.npy).
* Z_rho * <g/rho> -- the coupled model's partition function, whatever rho is -- from the
  reference's Jahn-Teller model and its three sampling models (A_rho = 2, 4, 8;
  examples/paper_1.5025058/alternate_rhos/jahnteller_D2_R{0,1,2}.json) must agree.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden
from oracle import pimc_oracle as po
from pibronic_amd import synthetic
from pibronic_amd.vibronic_model_keys import VibronicModelKeys as VMK

ALT = os.path.join(GOLDEN, "alternate_rhos")


def _setup(m, T, P):
    d = synthetic.diagonal_of(m)
    return po.Setup((m[VMK.E], m[VMK.w], m[VMK.G1], m[VMK.G2]),
                    (d[VMK.E], d[VMK.w], d[VMK.G1]), T, P)


def _two_surface_model(q0, q1):
    E = np.array([[0.05, 0.0], [0.0, 0.12]])
    w = np.array([0.15])
    L = np.zeros((1, 2, 2))
    L[0, 0, 0], L[0, 1, 1] = 0.04, -0.03
    Q = np.zeros((1, 1, 2, 2))
    Q[0, 0, 0, 0], Q[0, 0, 1, 1] = q0, q1
    return {VMK.N: 1, VMK.A: 2, VMK.E: E, VMK.w: w, VMK.G1: L, VMK.G2: Q}


def test_closed_form_matches_quadrature():
    """A=2, N=1, P=3: Z_g and Z_rho by a 60^3-point Gauss-Hermite rule over R in R^3 of the
    oracle's unscaled g and rho; the ratio must equal the closed form to rounding."""
    s = _setup(_two_surface_model(-0.05, 0.08), 300.0, 3)
    t, wt = np.polynomial.hermite.hermgauss(60)
    grid = np.stack(np.meshgrid(t, t, t, indexing="ij"), -1).reshape(-1, 3)
    weights = (wt[:, None, None] * wt[None, :, None] * wt[None, None, :]).reshape(-1)
    weights = weights * np.exp((grid ** 2).sum(axis=1))
    r = po.evaluate(s, grid[:, None, :].copy(), scale=False)
    quad = np.log((weights * r["s_g"]).sum() / (weights * r["s_rho"]).sum())
    exact = po.log_z_ratio_surface_diagonal(s)
    assert abs(exact) > 0.1  # the quadratic terms matter
    assert abs(quad - exact) < 1e-12, (quad, exact)


def test_closed_form_without_quadratic_terms_is_one():
    """Q = 0: V = 0, M = I and g == rho per sample (rho is the coupled model's diagonal)."""
    s = _setup(_two_surface_model(0.0, 0.0), 300.0, 12)
    assert abs(po.log_z_ratio_surface_diagonal(s)) < 1e-13


def test_closed_form_rejects_coupled_models():
    m = synthetic.generate(3, 2, quadratic=True, seed=1)
    with pytest.raises(ValueError):
        po.log_z_ratio_surface_diagonal(_setup(m, 300.0, 8))


def test_closed_form_matches_oracle_monte_carlo():
    """Harmonic-rho estimate of the oracle on the A2 N3 P12 workload model with the
    inter-surface couplings removed (finite-variance weights) within 3 sigma."""
    m = synthetic.generate(2, 3, quadratic=True, seed=0, surface_diagonal=True)
    s = _setup(m, 300.0, 12)
    rng = np.random.default_rng(5)
    ratios = []
    for _ in range(10):
        q, _ = s.draw_q(rng, 10_000)
        o = po.evaluate(s, q)
        ratios.append(o["s_g"] / o["s_rho"])
    r = np.concatenate(ratios)
    z, err = r.mean(), r.std() / np.sqrt(r.size - 1)
    exact = np.exp(po.log_z_ratio_surface_diagonal(s))
    assert abs(exact - 1.0) > 20 * err  # the estimate has something to find
    assert abs(z - exact) < 3 * err, (z, err, exact)


def test_other_surface_count_denominator_closed_form():
    """The explicit example's coupled model (A=2) with a 4-surface sampling model of the
    same frequencies: the oracle's rho (unscaled, via the qTempTensor copy) equals
    sum_a F_a^P exp(-1/2 sum_j (R_j - d_aj)^T K_j (R_j - d_aj)), with K_j the inverse of the
    reference's golden covariance for that mode (identical for every surface)."""
    A, N, E, w, L, Q = po.read_model_json(os.path.join(GOLDEN, "explicit_data", "coupled_model.json"))
    _, _, Er, wr, Lr, _ = po.read_model_json(os.path.join(GOLDEN, "explicit_data", "sampling_model.json"),
                                             diagonal=True)
    rng = np.random.default_rng(3)
    Er4 = np.concatenate([Er, Er + rng.uniform(0.0, 0.05, Er.shape)])
    Lr4 = np.concatenate([Lr, Lr + rng.normal(0.0, 0.05, Lr.shape)], axis=1)
    s = po.Setup((E, w, L, Q), (Er4, wr, Lr4), 300.0, 5)
    assert s.A_rho == 4 and s.A == 2
    cov = golden("sampling_covariance.npy")  # (A, N, P, P)
    K = np.linalg.inv(cov[0])                 # (N, P, P)
    assert np.allclose(np.linalg.inv(cov[1]), K)
    q = golden("samples.npy")[:, 0]           # (10, N, P), the reference's own samples
    r = po.evaluate(s, q.copy(), scale=False)
    _, _, pref = s.c_rho
    x = q[:, None, :, :] - s.rho["shift"][None, :, :, None]            # (B, A_rho, N, P)
    quad = np.einsum("banp,npq,banq->ba", x, K, x)
    expect = (pref[None, :] ** 5 * np.exp(-0.5 * quad)).sum(axis=1)
    np.testing.assert_allclose(r["s_rho"], expect, rtol=1e-10)


@pytest.mark.parametrize("P", [8])
def test_alternate_sampling_models_agree_on_z(P):
    """Jahn-Teller model 2 of the reference's paper examples with its three sampling models
    (A_rho = 2, 4, 8): Z_rho (closed form) * <g/rho> estimates the same Z_g for each; every
    pair within 3 sigma (oracle, 1e5 samples per model)."""
    A, N, E, w, L, Q = po.read_model_json(os.path.join(ALT, "jahnteller_2.json"))
    est = []
    for r in range(3):
        Ar, _, Er, wr, Lr, _ = po.read_model_json(os.path.join(ALT, f"jahnteller_D2_R{r}.json"),
                                                  diagonal=True)
        s = po.Setup((E, w, L, Q), (Er, wr, Lr), 300.0, P)
        assert s.A_rho == (2, 4, 8)[r]
        zr = po.z_rho((Er, wr, Lr), s.beta)[0]
        rng = np.random.default_rng(11 + r)
        x = []
        for _ in range(10):
            q, _ = s.draw_q(rng, 10_000)
            o = po.evaluate(s, q)
            x.append(o["s_g"] / o["s_rho"])
        x = np.concatenate(x)
        est.append((zr * x.mean(), zr * x.std() / np.sqrt(x.size - 1)))
    for i in range(3):
        for j in range(i + 1, 3):
            (a, ea), (b, eb) = est[i], est[j]
            assert abs(a - b) < 3 * np.hypot(ea, eb), est
