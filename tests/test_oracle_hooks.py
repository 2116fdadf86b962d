"""Pin the oracle's restatement of the training-data hooks (CPU only).

block_compute_rhoR_from_input_samples and block_compute_gR_from_raw_samples
(pibronic/pimc/pimc.py:1273-1338) evaluate rho and g at caller-supplied R after shifting R by
the coupled model's state shift d (quirk Q6). Checks:
  * d = 0 (no diagonal linear couplings): both hooks reduce to the ordinary unscaled
    evaluation, which the golden vectors pin;
  * d != 0 with off-diagonal linear couplings: V(R) is not symmetric and the reference's
    assertion (pimc.py:1169) fires;
  * d != 0 with a symmetric V: the O-factor part equals the closed form with the shift
    counted twice, and V equals the expansion with the shifts folded into E and L (the
    product path's formulation), evaluated independently here.
"""
import numpy as np
import pytest

from helpers import random_model, stage_dict_models
from oracle import pimc_oracle as po
from pibronic_amd.vibronic_model_keys import VibronicModelKeys as VMK


def _setup(tmp_path, m, T=300.0, P=7):
    FS = stage_dict_models(tmp_path, m)
    A, N, E, w, L, Q = po.read_model_json(FS.path_vib_model)
    Ar, Nr, Er, wr, Lr, _ = po.read_model_json(FS.path_rho_model, diagonal=True)
    return po.Setup((E, w, L, Q), (Er, wr, Lr), T, P)


def _no_diag_linear(m):
    L = m[VMK.G1].copy()
    for j in range(L.shape[0]):
        np.fill_diagonal(L[j], 0.0)
    m[VMK.G1] = L
    return m


def test_hooks_reduce_to_standard_evaluation_without_shift(tmp_path):
    s = _setup(tmp_path, _no_diag_linear(random_model(3, 2, seed=4, quad="full")))
    assert np.all(s.vib["shift"] == 0.0)
    R = np.random.default_rng(0).random((6, s.N, s.P))
    std = po.evaluate(s, R, scale=False)
    assert np.allclose(po.rho_of_R_hook(s, R), std["s_rho"], rtol=1e-14, atol=0)
    assert np.allclose(po.g_of_R_hook(s, R), std["s_g"], rtol=1e-13, atol=0)


def test_g_hook_asserts_on_nonsymmetric_V(tmp_path):
    s = _setup(tmp_path, random_model(3, 2, seed=5))
    assert np.any(s.vib["shift"] != 0.0)
    R = np.random.default_rng(1).random((2, s.N, s.P))
    with pytest.raises(AssertionError):
        po.g_of_R_hook(s, R)


def _symmetric_shifted_model(seed):
    """d != 0 (diagonal linear couplings) but no off-diagonal linear couplings and Q symmetric
    in the mode pair, so V(R) is symmetric for shifted coordinates; off-diagonal E and full Q
    keep V non-trivial."""
    m = random_model(3, 2, seed=seed, quad="full")
    m[VMK.G2] = 0.5 * (m[VMK.G2] + m[VMK.G2].transpose(1, 0, 2, 3))  # Q^{jk} = Q^{kj}
    L = m[VMK.G1].copy()
    for j in range(L.shape[0]):
        L[j] = np.diag(np.diag(L[j]))
    m[VMK.G1] = L
    return m


def test_hook_shift_arithmetic(tmp_path):
    s = _setup(tmp_path, _symmetric_shifted_model(6))
    d = s.vib["shift"]
    assert np.any(d != 0.0)
    R = np.random.default_rng(2).random((4, s.N, s.P))
    # rho: sampling O-factors at R - d - d_rho, closed form per surface
    coth, csch, pref = s.c_rho
    x = R[:, None] - (d + s.rho["shift"])[None, :, :, None]
    xn = np.roll(x, -1, axis=3)
    e = (coth[None, :, :, None] * (x ** 2 + xn ** 2) - 2 * csch[None, :, :, None] * x * xn).sum(axis=(2, 3))
    want = (np.prod(np.broadcast_to(pref, (s.P, s.A_rho)), axis=0)[None] * np.exp(-0.5 * e)).sum(axis=1)
    assert np.allclose(po.rho_of_R_hook(s, R), want, rtol=1e-12, atol=0)
    # g: O at R - 2d, V from E', L' (shifts folded), M = expm(-tau V), ring trace
    import scipy.linalg
    E, L, Q = s.vib["energy"], s.vib["linear"], s.vib["quadratic"]
    Lp = L - 0.5 * np.einsum("dkbc,bd->kbc", Q, d) - 0.5 * np.einsum("kebc,ce->kbc", Q, d)
    Ep = E - np.einsum("dbc,bd->bc", L, d) + 0.5 * np.einsum("debc,ce,bd->bc", Q, d, d)
    cv, sv, pv = s.c_vib
    g = np.empty(R.shape[0])
    for b in range(R.shape[0]):
        N_ = np.eye(s.A)
        for p in range(s.P):
            r = R[b, :, p]
            V = Ep + np.einsum("kbc,k->bc", Lp, r) + 0.5 * np.einsum("debc,e,d->bc", Q, r, r)
            M = scipy.linalg.expm(-s.tau * V)
            y = r[None] - 2 * d
            yn = R[b, :, (p + 1) % s.P][None] - 2 * d
            ee = (cv * (y ** 2 + yn ** 2) - 2 * sv * y * yn).sum(axis=1)
            O = np.diag(pv * np.exp(-0.5 * ee))
            N_ = N_ @ M @ O
        g[b] = np.trace(N_)
    assert np.allclose(po.g_of_R_hook(s, R), g, rtol=1e-11, atol=0)
