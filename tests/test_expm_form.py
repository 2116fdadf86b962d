"""M = exp(-tau V) without an eigendecomposition: the schemes of the evaluation kernels,
restated in NumPy and checked against the reference's route M = U diag(exp(-tau lambda)) U^T
from eigh (pibronic/pimc/pimc.py:1167-1187).

The kernels use a scaled and squared Taylor polynomial T_m: X = -tau V is scaled by 2^-s so
that a bound alpha on its spectral radius satisfies alpha 2^-s <= theta_m, T_m is evaluated
and squared s times. For a symmetric X every intermediate is a polynomial in X, so the
2-norm error is the largest scalar error over the eigenvalues. Schemes:
  * per-lane FP64 kernel, A <= 4: m = 12 by Paterson-Stockmeyer in X^3,
    alpha = min(|X|_F, |X^3|_F^(1/3)), theta_12 = 0.2996;
  * per-lane FP64 kernel, 5 <= A <= 8: m = 8 in three products (Sastre's form),
    alpha = |X^2|_F^(1/2), theta_8 = 0.049 (Horner's rule at degree 8 with alpha = |X|_F, the
    earlier scheme, is kept as a comparison);
  * per-lane FP32 kernel: the same three-product form in float32, theta_8 = 0.619;
  * one-wave FP64 kernel (A <= 16): m = 12 by Paterson-Stockmeyer in X^4,
    alpha = max(|X^3|_F^(1/3), |X^4|_F^(1/4)), theta_12 = 0.2996.
The bound checked is normwise: |M - M_eigh|_max <= 16 (2^s + m) max(1, |X|_2) u |M|_max,
with u the format's unit roundoff: each squaring doubles the relative error of the power it
squares, and eigh's own M carries errors of order |X|_2 u. FP32 matrices are kept below
exp(88) (float range), as the FP32 kernel's M must be.
"""
import math

import numpy as np
import pytest

from oracle import pimc_oracle as po
from pibronic_amd import synthetic
from pibronic_amd.vibronic_model_keys import VibronicModelKeys as VMK

U64 = 2.0 ** -53
U32 = 2.0 ** -24


def _squarings(alpha, theta):
    with np.errstate(divide="ignore"):
        sd = np.log2(alpha) - math.log2(theta)
    return np.where(sd > 0, np.ceil(np.nan_to_num(sd, neginf=0.0)), 0).astype(int)


def expm_horner(X, m, theta, dtype):
    """Per-lane scheme: Frobenius bound, Horner evaluation of T_m, s squarings (per matrix).
    Returns M and s."""
    X = np.asarray(X, dtype=dtype)
    n = X.shape[-1]
    s = _squarings(np.linalg.norm(X, axis=(-2, -1)), theta)
    Y = X * np.ldexp(np.ones_like(s, dtype=dtype), -s)[..., None, None]
    I = np.eye(n, dtype=dtype)
    c = [dtype(1.0 / math.factorial(k)) for k in range(m + 1)]
    T = c[m] * Y + c[m - 1] * I
    for k in range(m - 2, -1, -1):
        T = Y @ T + c[k] * I
    for k in range(int(s.max()) if s.size else 0):
        T = np.where((s > k)[..., None, None], T @ T, T)
    return T, s


def expm_ps12(X):
    """One-wave scheme: alpha from |X^3|, |X^4|, T_12 = B0 + X^4 (B1 + X^4 B2)."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[-1]
    X2 = X @ X
    X3 = X2 @ X
    X4 = X2 @ X2
    alpha = np.maximum(np.linalg.norm(X3, axis=(-2, -1)) ** (1 / 3),
                       np.linalg.norm(X4, axis=(-2, -1)) ** 0.25)
    s = _squarings(alpha, 0.2996158)
    f = np.ldexp(1.0, -s)[..., None, None]
    x1, x2, x3, x4 = X * f, X2 * f ** 2, X3 * f ** 3, X4 * f ** 4
    I = np.eye(n)
    c = [1.0 / math.factorial(k) for k in range(13)]
    B2 = c[8] * I + c[9] * x1 + c[10] * x2 + c[11] * x3 + c[12] * x4
    B1 = c[4] * I + c[5] * x1 + c[6] * x2 + c[7] * x3
    B0 = I + x1 + c[2] * x2 + c[3] * x3
    T = x4 @ (x4 @ B2 + B1) + B0
    for k in range(int(s.max()) if s.size else 0):
        T = np.where((s > k)[..., None, None], T @ T, T)
    return T, s


def expm_ps12_x3(X):
    """Per-lane FP64 scheme: alpha = min(|X|_F, |X^3|_F^(1/3)), theta_12 = 0.2996, T_12 by
    Paterson-Stockmeyer in X^3: T = ((B3 X^3 + B2) X^3 + B1) X^3 + B0 with
    B_n = c_3n I + c_3n+1 X + c_3n+2 X^2 (n < 3) and B3 = c9 I + c10 X + c11 X^2 + c12 X^3."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[-1]
    X2 = X @ X
    X3 = X2 @ X
    alpha = np.minimum(np.linalg.norm(X, axis=(-2, -1)), np.linalg.norm(X3, axis=(-2, -1)) ** (1 / 3))
    s = _squarings(alpha, 0.2996)
    f = np.ldexp(1.0, -s)[..., None, None]
    x1, x2, x3 = X * f, X2 * f ** 2, X3 * f ** 3
    I = np.eye(n)
    c = [1.0 / math.factorial(k) for k in range(13)]
    T = c[9] * I + c[10] * x1 + c[11] * x2 + c[12] * x3
    for k in (2, 1, 0):
        T = T @ x3 + (c[3 * k] * I + c[3 * k + 1] * x1 + c[3 * k + 2] * x2)
    for k in range(int(s.max()) if s.size else 0):
        T = np.where((s > k)[..., None, None], T @ T, T)
    return T, s


S8_COEF = (4.98011920555997349987007158e-3, 1.99204768222398939994802863e-2,
           7.66526532111914669031909419e-2, 8.76500980178555335977132598e-1,
           1.22552115011207473091611921e-1, 2.97430720484762666375079572e0)


def expm_s8(X, theta=0.049, dtype=np.float64):
    """Per-lane scheme for 5 <= A <= 8 (FP64, theta_8 = 0.049) and of the FP32 kernel (float32,
    theta_8 = 0.619): alpha = |X^2|_F^(1/2), T_8 in three products: Y = X2 (c1 X2 + c2 X),
    T = (Y + c3 X2 + c4 X)(Y + c5 X2) + c6 Y + X2/2 + X + I."""
    X = np.asarray(X, dtype=dtype)
    n = X.shape[-1]
    X2 = X @ X
    s = _squarings(np.linalg.norm(X2.astype(np.float64), axis=(-2, -1)) ** 0.5, theta)
    f = np.ldexp(np.ones_like(s, dtype=dtype), -s)[..., None, None]
    x1, x2 = X * f, X2 * f ** 2
    c1, c2, c3, c4, c5, c6 = (dtype(c) for c in S8_COEF)
    Y = x2 @ (c1 * x2 + c2 * x1)
    T = (Y + c3 * x2 + c4 * x1) @ (Y + c5 * x2) + (c6 * Y + dtype(0.5) * x2 + x1 + np.eye(n, dtype=dtype))
    for k in range(int(s.max()) if s.size else 0):
        T = np.where((s > k)[..., None, None], T @ T, T)
    return T, s


def expm_eigh(X):
    lam, U = np.linalg.eigh(np.asarray(X, dtype=np.float64))
    return np.einsum("...ik,...k,...jk->...ij", U, np.exp(lam), U)


SCHEMES = {  # (function, unit roundoff, degree)
    "lanes_fp64": (expm_ps12_x3, U64, 12),
    "lanes_fp64_h8": (lambda X: expm_horner(X, 8, 0.049, np.float64), U64, 8),
    "lanes_fp64_s8": (expm_s8, U64, 8),
    "lanes_fp32": (lambda X: expm_s8(X, 0.619, np.float32), U32, 8),
    "wide_fp64": (expm_ps12, U64, 12),
}


def _check(X, scheme):
    f, u, m = SCHEMES[scheme]
    M, s = f(X)
    M = M.astype(np.float64)
    R = expm_eigh(X)
    err = np.abs(M - R).max(axis=(-2, -1))
    nrm = np.abs(R).max(axis=(-2, -1))
    rho = np.abs(np.linalg.eigvalsh(np.asarray(X, dtype=np.float64))).max(axis=-1)
    bound = 16.0 * (2.0 ** s + m) * np.maximum(1.0, rho) * u * nrm
    worst = float(np.max(err / bound))
    assert worst <= 1.0, (scheme, worst, float(np.max(err / nrm)))
    return float(np.max(err / nrm)), worst


def _workload_X(A, N, P, T, off, B=24):
    kw = {} if off is None else {"off": off}
    m = synthetic.generate(A, N, quadratic=True, seed=0, **kw)
    E, w, L, Q = m[VMK.E], m[VMK.w], m[VMK.G1], m[VMK.G2]
    Er = np.diag(E).copy()
    Lr = np.stack([np.diag(L[j]) for j in range(N)])
    s = po.Setup((E, w, L, Q), (Er, w, Lr), T, P)
    q, _ = s.draw_q(np.random.default_rng(3), B)
    o = po.evaluate(s, q, keep=True)
    return -s.tau * o["V"].reshape(-1, A, A)


@pytest.mark.parametrize("shape,schemes", [
    ((4, 8, 64, 300.0, None), ["lanes_fp64", "lanes_fp32", "wide_fp64"]),
    ((8, 16, 64, 300.0, None), ["lanes_fp64_s8", "lanes_fp64_h8", "lanes_fp32", "wide_fp64"]),
    ((12, 24, 128, 250.0, 0.1), ["wide_fp64"]),
])
def test_workload_matrices(shape, schemes):
    X = _workload_X(*shape)
    for sc in schemes:
        e, worst = _check(X, sc)
        print(shape[0], sc, "max relative M error", e, "of bound", worst)


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_norm_range_and_spectra(scheme):
    """Random symmetric matrices with spectral radius 1e-3..1e2 (up to ~11 squarings),
    clustered and near-degenerate spectra, both signs."""
    rng = np.random.default_rng(11)
    mats = []
    scales = (1e-3, 0.1, 1.0, 5.0, 30.0) + (() if scheme == "lanes_fp32" else (100.0,))
    for n in (2, 4, 8):
        for scale in scales:
            Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
            for lam in (rng.uniform(-1, 1, n), np.r_[1.0, 1.0 + 1e-9, rng.uniform(-1, 1, n - 2)],
                        -np.abs(rng.uniform(0.5, 1, n))):
                mats.append(((Q * (scale * lam)) @ Q.T, n))
    for X, n in mats:
        X = 0.5 * (X + X.T)
        _check(X[None], scheme)


@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_special_cases(scheme):
    f, u, m = SCHEMES[scheme]
    # V = 0 (uncoupled model): M = I exactly
    for n in (1, 3, 8):
        M, s = f(np.zeros((1, n, n)))
        assert s[0] == 0 and np.array_equal(M.astype(np.float64)[0], np.eye(n))
    # diagonal X: exp of each entry, exact zeros off the diagonal
    d = np.array([-3.0, -0.7, 0.0, 0.2, 1.5])
    M, s = f(np.diag(d)[None])
    M = M.astype(np.float64)[0]
    assert np.allclose(np.diag(M), np.exp(d), rtol=16 * (2.0 ** s[0] + m) * u, atol=0.0)
    assert np.all(M[~np.eye(5, dtype=bool)] == 0.0)


def test_s8_coefficients_give_the_taylor_polynomial():
    """The three-product form expands to 1 + x + ... + x^8/8! (each coefficient to 1e-15
    relative), and each of its six coefficients matters: perturbing any one by 1e-6 moves some
    Taylor coefficient by more than 1e-8 relative."""
    from numpy.polynomial import polynomial as npp

    def expand(c):
        c1, c2, c3, c4, c5, c6 = c
        x, x2 = [0.0, 1.0], [0.0, 0.0, 1.0]
        y = npp.polymul(x2, npp.polyadd(npp.polymul([c1], x2), npp.polymul([c2], x)))
        left = npp.polyadd(npp.polyadd(y, npp.polymul([c3], x2)), npp.polymul([c4], x))
        right = npp.polyadd(y, npp.polymul([c5], x2))
        t = npp.polyadd(npp.polymul(left, right), npp.polymul([c6], y))
        t = npp.polyadd(t, [1.0, 1.0, 0.5])
        return np.pad(t, (0, 9 - len(t)))[:9]

    taylor = np.array([1.0 / math.factorial(k) for k in range(9)])
    assert np.max(np.abs(expand(S8_COEF) / taylor - 1.0)) < 1e-15
    for i in range(6):
        c = list(S8_COEF)
        c[i] *= 1.0 + 1e-6
        assert np.max(np.abs(expand(c) / taylor - 1.0)) > 1e-8, i
