"""FP64 NumPy restatement of Pibronic's per-block PIMC estimator.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): this is the checker for the HIP path
and the timed CPU baseline, never part of the product.

Each function names the reference lines it restates. The arithmetic deliberately keeps the
reference's expression forms and reduction orders (the O-matrix exponent at
pibronic/pimc/pimc.py:1106, the sequential sum over modes, the in-place shift/unshift of the
coordinate tensor, the dense per-(sample, bead) `.dot` ring product at :1195-1206), because
the per-sample parity bound (1e-10 relative) is close to the rounding noise of the exponent's
cancellation when tau*omega is small.

Layouts follow the reference: q (B, A, N, P) for the per-surface coordinate copies,
O (B, P, A, A) dense diagonal matrices, V / U / M (B, P, A, A), eigenvalues (B, P, A).
"""
import json

import numpy as np
from numpy import newaxis as NEW
import scipy.linalg
import scipy.special

# pibronic/constants.py:12,24 -- k_B in eV/K, built from the same two NIST numbers
BOLTZMANN_EV_PER_K = np.float64(1.38064852e-23) / np.float64(1.6021766208e-19)
# pibronic/constants.py:28
DELTA_BETA = 2.0e-4


def beta_of_T(temperature):
    """pibronic/constants.py:41-44"""
    return 1. / (temperature * BOLTZMANN_EV_PER_K)


# --------------------------------------------------------------------------------------
# model JSON (pibronic/vibronic/vibronic_model_io.py:42-73, 575-649;
#             keys pibronic/vibronic/vibronic_model_keys.py:9-24)
# --------------------------------------------------------------------------------------
K_N, K_A = "number of modes", "number of surfaces"
K_E, K_W, K_L, K_Q = "energies", "frequencies", "linear couplings", "quadratic couplings"


def read_model_json(path, diagonal=False):
    """Return (A, N, E, w, L, Q) as float64 arrays; absent couplings are zeros.

    Coupled shapes E (A,A), L (N,A,A), Q (N,N,A,A); diagonal shapes E (A,), L (N,A), Q (N,N,A)
    (vibronic_model_io.py:42-73). Absent keys mean all-zero arrays (:575-596, :549-560).
    """
    with open(path, "r", encoding="UTF8") as fh:
        d = json.loads(fh.read())
    A, N = int(d[K_A]), int(d[K_N])
    if diagonal:
        shapes = {K_E: (A,), K_W: (N,), K_L: (N, A), K_Q: (N, N, A)}
    else:
        shapes = {K_E: (A, A), K_W: (N,), K_L: (N, A, A), K_Q: (N, N, A, A)}
    out = {}
    for k, shp in shapes.items():
        if k in d:
            arr = np.array(d[k], dtype=np.float64)
            assert arr.shape == shp, f"{k} has shape {arr.shape}, expected {shp}"
        else:
            arr = np.zeros(shp, dtype=np.float64)
        out[k] = arr
    return A, N, out[K_E], out[K_W], out[K_L], out[K_Q]


# --------------------------------------------------------------------------------------
# model folding and temperature constants
# --------------------------------------------------------------------------------------
def temperature_constants(omega, tilde_energy, tau, A):
    """coth, csch (A, N) and the O-matrix prefactor F (A,) at imaginary-time step tau.

    pibronic/pimc/pimc.py:59-83: coth = tanh(tau w)^-1, csch = sinh(tau w)^-1,
    F_a = exp(-tau E~_a) * (prod_j csch_j)^(1/2); the (2 pi)^(-1/2) factors cancel in g/rho.
    """
    w_AN = np.broadcast_to(omega, (A, omega.shape[0]))
    coth = np.tanh(1.0 * tau * w_AN) ** (-1.)
    csch = np.sinh(1.0 * tau * w_AN) ** (-1.)
    pref = np.exp(-tau * tilde_energy.copy())
    pref = pref * (np.prod(csch, axis=1) ** 0.5)
    return coth, csch, pref


def fold_coupled(E, w, L, Q):
    """Fold the diagonal linear terms of the coupled ("vib") model into displaced oscillators.

    pibronic/pimc/pimc.py:172-177 (delta_a = -1/2 sum_j L^j_aa^2 / w_j),
    :208-218 (d_aj = -L^j_aa / w_j, then zero E_aa and L^j_aa for the residual coupling V).
    Returns a dict; E_tilde = E_aa + delta_a is the energy used by the prefactor (:75-76).
    """
    A = E.shape[0]
    delta = np.zeros(A)
    for a in range(A):
        delta[a] = -0.5 * (L[:, a, a] ** 2. / w).sum(axis=0)
    e_diag = np.diag(E).copy()
    shift = np.zeros((A, w.shape[0]))
    E_res, L_res = E.copy(), L.copy()
    for a in range(A):
        shift[a, :] = -L_res[:, a, a] / w
        L_res[:, a, a] = 0.0
        E_res[a, a] = 0.0
    return dict(A=A, N=w.shape[0], omega=w.copy(), delta=delta, tilde=e_diag + delta,
                shift=shift, energy=E_res, linear=L_res, quadratic=Q.copy())


def fold_sampling(E, w, L, beta):
    """Diagonal sampling ("rho") model: displacement, energy shift and surface weights.

    pibronic/pimc/pimc.py:336-339 (delta), :341-351 (weights w_a ~ e^{-beta E~_a} /
    prod_j sinh(beta w_j / 2), normalised), :372-381 (d_aj = -L_ja / w_j). The quadratic
    terms of the sampling model are loaded but never used by the reference.
    """
    A = E.shape[0]
    delta = -0.5 * (L ** 2. / w[:, NEW]).sum(axis=0)
    weight = np.exp(-beta * (E + delta))
    weight /= np.prod(np.sinh((beta * w) / 2.))
    weight /= weight.sum()
    shift = np.zeros((A, w.shape[0]))
    for a in range(A):
        shift[a, :] = -L[:, a] / w
    return dict(A=A, N=w.shape[0], omega=w.copy(), delta=delta, tilde=E + delta,
                shift=shift, weight=weight)


def circulant_basis(P):
    """Ring adjacency matrix C (ones at +-1, periodic) and eigh(C, UPLO='L').

    pibronic/pimc/pimc.py:681-689.
    """
    assert P >= 3, "circulant matrix requires 3 or more beads"
    C = scipy.linalg.circulant([0, 1] + [0] * (P - 3) + [1])
    evals, evecs = np.linalg.eigh(C, UPLO="L")
    return C, evals, evecs


class Setup:
    """Everything the per-block kernels need for one (model, rho, T, P) combination.

    Mirrors BoxData.preprocess + initialize_models (pibronic/pimc/pimc.py:647-692, 633-645)
    and the PM variant (:695-738) without the file-system plumbing.
    """

    def __init__(self, coupled, sampling, temperature, P, delta_beta=DELTA_BETA, beta=None):
        self.P = int(P)
        self.beta = beta_of_T(temperature) if beta is None else beta
        self.tau = self.beta / self.P
        self.delta_beta = delta_beta
        self.tau_plus = (self.beta + delta_beta) / self.P
        self.tau_minus = (self.beta - delta_beta) / self.P
        E, w, L, Q = coupled
        Er, wr, Lr = sampling
        self.vib = fold_coupled(E, w, L, Q)
        self.rho = fold_sampling(Er, wr, Lr, self.beta)
        self.A, self.N, self.A_rho = self.vib["A"], self.vib["N"], self.rho["A"]
        v, r = self.vib, self.rho
        self.c_vib = temperature_constants(v["omega"], v["tilde"], self.tau, self.A)
        self.c_vib_p = temperature_constants(v["omega"], v["tilde"], self.tau_plus, self.A)
        self.c_vib_m = temperature_constants(v["omega"], v["tilde"], self.tau_minus, self.A)
        self.c_rho = temperature_constants(r["omega"], r["tilde"], self.tau, self.A_rho)
        self.C, self.circ_evals, self.circ_evecs = circulant_basis(self.P)
        # pibronic/pimc/pimc.py:401-404 -- per-(surface, mode, normal mode) standard deviation
        coth_r, csch_r, _ = self.c_rho
        inv_cov = 2. * coth_r[:, :, NEW] - csch_r[:, :, NEW] * self.circ_evals
        self.inverse_covariance = inv_cov
        self.sigma = np.sqrt(1. / inv_cov)

    # ---------------------------------------------------------------------------- sampler
    def draw_q(self, rng, B):
        """Harmonic-rho bead coordinates R (B, N, P) and their surface sources (B,).

        pibronic/pimc/pimc.py:386-389 (categorical surface draw), :329-333 (normal draws with
        sigma of the source surface), :618-624 (R = U_circ . cc + d_src). The reference draws
        the sources once for all samples with the legacy global generator; here a
        numpy Generator draws them per block, which is the same distribution.
        """
        src = rng.choice(self.A_rho, size=B, p=self.rho["weight"])
        cc = rng.normal(0.0, self.sigma[src, ...], size=(B, self.N, self.P))
        R = np.einsum("ab,ijb->ija", self.circ_evecs, cc)
        R += self.rho["shift"][src, :, NEW]
        return R, src


# --------------------------------------------------------------------------------------
# per-block kernels (pibronic/pimc/pimc.py:1062-1213)
# --------------------------------------------------------------------------------------
def o_matrix(q_all, q_tmp, shift, consts):
    """Diagonal O-matrices (B, P, A_, A_) of one model at one temperature.

    pibronic/pimc/pimc.py:1087-1129. The shift is applied to the per-surface coordinate
    tensor in place and removed again afterwards, exactly as the reference does, so the
    ulp-level drift it leaves in q (quirk Q4) is reproduced. When the model's surface count
    differs from the coordinate tensor's, the temporary copy q_tmp is used (:1094-1096).
    """
    coth, csch, pref = consts
    own = shift.shape[0] == q_all.shape[1]
    q1 = q_all if own else q_tmp
    q1 -= shift[NEW, :, :, NEW]
    q2 = np.roll(q1, shift=-1, axis=3)
    c1 = coth[NEW, :, :, NEW]
    c2 = csch[NEW, :, :, NEW]
    expo = -0.5 * np.sum(c1 * (q1 ** 2. + q2 ** 2.) - 2. * c2 * q1 * q2, axis=2).swapaxes(1, 2)
    np.exp(expo, out=expo)
    B, Ah, _, P = q1.shape
    O = np.zeros((B, P, Ah, Ah))
    for a in range(Ah):
        O[:, :, a, a] = expo[:, :, a]
    O *= np.broadcast_to(pref, (B, P, Ah))[..., NEW]
    q1 += shift[NEW, :, :, NEW]
    return O


def scaling_factors(O_rho, O_vib):
    """S12 (B, P): per-bead max over both O's (pibronic/pimc/pimc.py:1076-1084)."""
    return np.maximum(np.amax(O_rho, axis=(2, 3)), np.amax(O_vib, axis=(2, 3)))


def denominator(O_rho):
    """rho = tr prod_p O_rho(p) (pibronic/pimc/pimc.py:1132-1136)."""
    return O_rho.prod(axis=1).trace(axis1=1, axis2=2)


def coupling_eig(q_all, energy, linear, quadratic):
    """Residual diabatic coupling V (B, P, A, A) and eigh(V, UPLO='L').

    pibronic/pimc/pimc.py:1139-1174: V = 1/2 q.Q.q + L.q + E with the diagonal of E and L
    folded out; the contraction strings are the reference's.
    """
    V = np.einsum("acef, debc, abdf->afbc", q_all, 0.5 * quadratic, q_all)
    V += np.einsum("dbc, abdf->afbc", linear, q_all)
    V += energy[NEW, NEW, :, :]
    assert np.allclose(V.transpose(0, 1, 3, 2), V)
    lam, U = np.linalg.eigh(V, UPLO="L")
    return V, lam, U


def m_matrix(lam, U, tau):
    """M = U diag(exp(-tau lambda)) U^T (pibronic/pimc/pimc.py:1181-1187)."""
    return np.einsum("abcd, abd, abed->abce", U, np.exp(-tau * lam), U, optimize="optimal")


def ring_trace(M, O):
    """g = tr prod_p (M_p O_p), left to right, as the reference's (sample, bead) loop.

    pibronic/pimc/pimc.py:1190-1209.
    """
    B, P, A, _ = M.shape
    out = np.empty(B)
    acc = np.empty((A, A))
    for b in range(B):
        acc[:] = np.identity(A)
        for p in range(P):
            acc = acc.dot(M[b, p])
            acc = acc.dot(O[b, p])
        out[b] = np.trace(acc)
    return out


def evaluate(setup, q, pm=False, scale=True, keep=False):
    """Evaluate one block of host-fed coordinates q (B, N, P).

    The sequence is block_compute's body (pibronic/pimc/pimc.py:1358-1385) or, with pm=True,
    block_compute_pm's (:1413-1449) after draw_sample/transform. With scale=False the S12
    scaling is skipped, which is what the reference's explicit golden test does
    (tests/pimc/test_pimc_explicit_example.py:186-207).

    Returns a dict with s_rho, s_g (and s_gP, s_gM), plus log_scale = sum_p log S_p, the
    logarithm of the factor the scaling removed (0 when scale=False).
    """
    B, N, P = q.shape
    A, A_rho = setup.A, setup.A_rho
    q_all = np.empty((B, A, N, P))
    q_all[:] = q[:, NEW]
    q_tmp = np.empty((B, A_rho, N, P))
    q_tmp[:] = q[:, NEW]
    O_rho = o_matrix(q_all, q_tmp, setup.rho["shift"], setup.c_rho)
    O_vib = o_matrix(q_all, q_tmp, setup.vib["shift"], setup.c_vib)
    if scale:
        S = scaling_factors(O_rho, O_vib)
        O_rho /= S[..., NEW, NEW]
        O_vib /= S[..., NEW, NEW]
        log_scale = np.log(S).sum(axis=1)
    else:
        S = np.ones((B, P))
        log_scale = np.zeros(B)
    res = dict(s_rho=denominator(O_rho), log_scale=log_scale)
    V, lam, U = coupling_eig(q_all, setup.vib["energy"], setup.vib["linear"],
                             setup.vib["quadratic"])
    M = m_matrix(lam, U, setup.tau)
    res["s_g"] = ring_trace(M, O_vib)
    if pm:
        O_p = o_matrix(q_all, q_tmp, setup.vib["shift"], setup.c_vib_p)
        O_p /= S[..., NEW, NEW]
        res["s_gP"] = ring_trace(M, O_p)
        O_m = o_matrix(q_all, q_tmp, setup.vib["shift"], setup.c_vib_m)
        O_m /= S[..., NEW, NEW]
        res["s_gM"] = ring_trace(M, O_m)
    if keep:
        res.update(O_rho=O_rho, O_vib=O_vib, V=V, lam=lam, U=U, M=M, S=S)
    return res


# --------------------------------------------------------------------------------------
# closed forms for the harmonic sampling model (used by the known-answer tests)
# --------------------------------------------------------------------------------------
def z_rho(sampling, beta):
    """Z_rho(beta) = sum_a e^{-beta E~_a} prod_j 1 / (2 sinh(beta w_j / 2)).

    The partition function of the displaced-oscillator mixture the reference samples from
    (weights at pibronic/pimc/pimc.py:341-351, normalisation dropped there). Returns
    (Z, E, Cv_over_kB_T2_free) where E = -d ln Z / d beta and the last item is
    d^2 ln Z / d beta^2 (so Cv = that / (k_B T^2)).
    """
    E, w, L = sampling
    delta = -0.5 * (L ** 2. / w[:, NEW]).sum(axis=0)
    tilde = E + delta
    osc = 1.0 / (2.0 * np.sinh(beta * w / 2.0))
    terms = np.exp(-beta * tilde) * np.prod(osc)
    Z = terms.sum()
    # d/dbeta of log(term_a) = -E~_a - sum_j (w_j/2) coth(beta w_j/2)
    zpe = np.sum(0.5 * w / np.tanh(beta * w / 2.0))
    dlog = -tilde - zpe
    E_mean = -np.sum(terms * dlog) / Z
    # second derivative of log term: sum_j (w_j/2)^2 csch^2(beta w_j / 2)
    d2 = np.sum((0.5 * w) ** 2 / np.sinh(beta * w / 2.0) ** 2)
    var = np.sum(terms * dlog ** 2) / Z - (np.sum(terms * dlog) / Z) ** 2
    return Z, E_mean, var + d2


# --------------------------------------------------------------------------------------
# training-data hooks (pibronic/pimc/pimc.py:1273-1338)
# --------------------------------------------------------------------------------------
def rho_of_R_hook(setup, R):
    """block_compute_rhoR_from_input_samples' rho(R), unscaled (pibronic/pimc/pimc.py:1273-1299):
    qTensor = R - d_vib (:1293-1295), then build_o_matrix subtracts the sampling model's shift
    again (:1092, quirk Q6) and build_denominator takes the trace (:1132-1136)."""
    B = R.shape[0]
    q_all = np.empty((B, setup.A, setup.N, setup.P))
    q_all[:] = R[:, NEW]
    q_all -= setup.vib["shift"][NEW, :, :, NEW]
    assert setup.A_rho == setup.A, "the hook reads qTempTensor when A_rho != A"
    O = o_matrix(q_all, None, setup.rho["shift"], setup.c_rho)
    return denominator(O)


def g_of_R_hook(setup, R):
    """block_compute_gR_from_raw_samples' g(R), unscaled (pibronic/pimc/pimc.py:1302-1338):
    qTensor = R - d_vib (generate_random_R_values, :602-611); build_o_matrix with the coupled
    model subtracts d_vib again (quirk Q6); V is built from the shifted qTensor; no scaling."""
    B = R.shape[0]
    q_all = np.empty((B, setup.A, setup.N, setup.P))
    q_all[:] = R[:, NEW]
    q_all -= setup.vib["shift"][NEW, :, :, NEW]
    O = o_matrix(q_all, None, setup.vib["shift"], setup.c_vib)
    V, lam, U = coupling_eig(q_all, setup.vib["energy"], setup.vib["linear"], setup.vib["quadratic"])
    M = m_matrix(lam, U, setup.tau)
    return ring_trace(M, O)


# --------------------------------------------------------------------------------------
# exact Z_g / Z_rho for surface-diagonal coupled models (pins the estimators' mean)
# --------------------------------------------------------------------------------------
def _log_gauss_integral(coth, csch, tqc, shift, P):
    """log of int_{R^P} exp(-1/2 x^T K x - (tqc/2) |x + d 1|^2) dx (d = shift), where
    K = 2 coth I - csch C is the ring precision of one mode (C the adjacency matrix of the
    ring, eigenvalues 2 cos(2 pi k / P)) and tqc = tau * Q_jj,aa. With A = K + tqc I and
    K 1 = (2 coth - 2 csch) 1:
      (P/2) log 2pi - 1/2 sum_k log(2 coth - csch lam_k + tqc)
        - tqc P d^2 / 2 + (tqc d)^2 P / (2 (2 coth - 2 csch + tqc)).
    """
    lam = 2.0 * np.cos(2.0 * np.pi * np.arange(P) / P)
    ev = 2.0 * coth - csch * lam + tqc
    if np.any(ev <= 0.0):
        raise ValueError("K + tau Q is not positive definite: the integral diverges")
    ones_ev = 2.0 * coth - 2.0 * csch + tqc
    return (0.5 * P * np.log(2.0 * np.pi) - 0.5 * np.sum(np.log(ev))
            - 0.5 * tqc * P * shift * shift + 0.5 * (tqc * shift) ** 2 * P / ones_ev)


def log_z_ratio_surface_diagonal(setup):
    """log(Z_g / Z_rho) exactly, for a coupled model whose E, L^j and Q^jk are diagonal in the
    surfaces with Q^jk = 0 for j != k (no inter-surface coupling) and any sampling model.

    Z_g = int g(R) dR with g = tr prod_p M_p O_p (pibronic/pimc/pimc.py:1177-1213): here V(R)
    is diagonal, V_aa(R_p) = 1/2 sum_j Q_jj,aa R_jp^2 (the folded E_aa, L^j_aa are zero,
    :208-218; the quadratic diagonal is not folded, :1147-1160), so M_p = diag(e^{-tau V_aa})
    and g = sum_a F_a^P prod_j exp(-1/2 x_aj^T K_j x_aj - tau/2 Q_jj,aa |R_j|^2), x = R - d_a
    (O-matrix exponent summed over the ring, :1087-1129) -- a Gaussian integral per (a, j).
    Z_rho = int rho(R) dR with rho = sum_a F^rho_a^P prod_j exp(-1/2 x^T K^rho_j x)
    (:1132-1136). <g/rho> under the harmonic sampler, and <g/q> under any proposal q in
    rho's normalisation, estimate exactly this ratio at the same P (the Trotter
    approximation is shared, not an error of the estimator).
    """
    v, r = setup.vib, setup.rho
    E, L, Q = v["energy"], v["linear"], v["quadratic"]
    A, N, P = setup.A, setup.N, setup.P
    offdiag = ~np.eye(A, dtype=bool)
    if np.any(E[offdiag]) or np.any(L[:, offdiag]) or np.any(Q[..., offdiag]):
        raise ValueError("the closed form needs surface-diagonal couplings")
    jk = ~np.eye(N, dtype=bool)
    if np.any(Q[jk]):
        raise ValueError("the closed form needs Q^jk = 0 for j != k")
    coth, csch, pref = setup.c_vib
    terms_g = np.empty(A)
    for a in range(A):
        s = P * np.log(pref[a])
        for j in range(N):
            s += _log_gauss_integral(coth[a, j], csch[a, j], setup.tau * Q[j, j, a, a],
                                     v["shift"][a, j], P)
        terms_g[a] = s
    cr, sr, pr = setup.c_rho
    terms_r = np.empty(setup.A_rho)
    for a in range(setup.A_rho):
        s = P * np.log(pr[a])
        for j in range(N):
            s += _log_gauss_integral(cr[a, j], sr[a, j], 0.0, 0.0, P)
        terms_r[a] = s
    return float(scipy.special.logsumexp(terms_g) - scipy.special.logsumexp(terms_r))
