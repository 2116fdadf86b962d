"""M = exp(-tau V) without an eigendecomposition: the schemes of the evaluation kernels,
restated in NumPy and checked against the reference's route M = U diag(exp(-tau lambda)) U^T
from eigh (pibronic/pimc/pimc.py:1167-1187).

The kernels use a scaled and squared Taylor polynomial T_m: X = -tau V is scaled by 2^-s so
that a bound alpha on its spectral radius satisfies alpha 2^-s <= theta_m, T_m is evaluated
and squared s times. For a symmetric X every intermediate is a polynomial in X, so the
2-norm error is the largest scalar error over the eigenvalues. Schemes:
  * per-lane FP64 kernel without +-dbeta (A != 6): m = 15+ (degree 16, Taylor through x^15) in
    four products, alpha = |X^2|_F^(1/2), theta_15+ = 0.6764;
  * per-lane FP64 kernel with +-dbeta, A <= 4: m = 12 by Paterson-Stockmeyer in X^3,
    alpha = min(|X|_F, |X^3|_F^(1/3)), theta_12 = 0.2996;
  * per-lane FP64 kernel with +-dbeta for 5 <= A <= 8, and A = 6 without: m = 8 in three
    products (Sastre's form), alpha = |X^2|_F^(1/2), theta_8 = 0.049 (Horner's rule at degree 8
    with alpha = |X|_F, the earlier scheme, is kept as a comparison);
  * per-lane FP32 kernel: the same three-product form in float32, theta_8 = 0.619;
  * one-wave FP64 kernel (A <= 16): the same four-product degree-15+ form, alpha = |X^2|_F^(1/2)
    (m = 12 by Paterson-Stockmeyer in X^4 with alpha = max(|X^3|_F^(1/3), |X^4|_F^(1/4)) and
    theta_12 = 0.2996, the earlier scheme, is kept as a comparison).
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


S15_COEF = (-4.01876161020103546293293287596e-4, -2.94553144027968298047549126357e-3,
            8.70906657683767596765346131523e-3, -4.01756844067356780152362494573e-1,
            -3.23076288812231208197540935626e-2, 2.33731940471157926649827413344e-2,
            2.61492797729811710806033238185e-1, -2.38107037387098722469262858218e-1,
            -4.1302763659297829110328428344e-2, -5.79236170707326052179805145605e0,
            2.22420917249637356119032703879e0, 1.04080173523135436461532671084e1,
            3.03012340073871230596564458406e0, -2.12975559049643578453357460382e0)
THETA15 = 2.0 ** -0.5640051


def expm_s15(X, theta=THETA15):
    """Per-lane FP64 scheme without +-dbeta: alpha = |X^2|_F^(1/2), degree 16 in four products
    (Taylor through x^15): y02 = X2 (c1 X2 + c2 X), y12 = (y02 + c3 X2 + c4 X)(y02 + c5 X2)
    + c6 y02 + c7 X2, T = (y12 + c8 X2 + c9 X)(y12 + c10 y02 + c11 X) + c12 y12 + c13 y02
    + c14 X2 + X + I, evaluated as the kernel does with y02 held as G = y02 + c5 X2."""
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[-1]
    X2 = X @ X
    s = _squarings(np.linalg.norm(X2, axis=(-2, -1)) ** 0.5, theta)
    f = np.ldexp(1.0, -s)[..., None, None]
    x1, x2 = X * f, X2 * f ** 2
    c1, c2, c3, c4, c5, c6, c7, c8, c9, c10, c11, c12, c13, c14 = S15_COEF
    G = x2 @ (c1 * x2 + c2 * x1) + c5 * x2
    y12 = (G + (c3 - c5) * x2 + c4 * x1) @ G + c6 * G + (c7 - c6 * c5) * x2
    F1 = y12 + c8 * x2 + c9 * x1
    F2 = y12 + c10 * G + c11 * x1 - c10 * c5 * x2
    T = F1 @ F2 + (c12 * y12 + c13 * G + (c14 - c13 * c5) * x2 + x1 + np.eye(n))
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
    "lanes_fp64_s15": (expm_s15, U64, 16),
    "lanes_fp32": (lambda X: expm_s8(X, 0.619, np.float32), U32, 8),
    "wide_fp64": (expm_s15, U64, 16),
    "wide_fp64_ps12": (expm_ps12, U64, 12),
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
    ((4, 8, 64, 300.0, None), ["lanes_fp64", "lanes_fp64_s15", "lanes_fp32", "wide_fp64"]),
    ((8, 16, 64, 300.0, None), ["lanes_fp64_s15", "lanes_fp64_s8", "lanes_fp64_h8", "lanes_fp32", "wide_fp64"]),
    ((12, 24, 128, 250.0, 0.1), ["wide_fp64", "wide_fp64_ps12"]),
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


def _s15_expand(c, one):
    """Coefficients of the four-product form as a polynomial in x (degree 16), in the
    arithmetic of `one` (float or mpmath)."""
    deg = 17

    def mul(a, b):
        r = [one * 0] * deg
        for i in range(deg):
            for j in range(deg - i):
                r[i + j] += a[i] * b[j]
        return r

    def lin(*terms):
        r = [one * 0] * deg
        for k, p in terms:
            for i in range(deg):
                r[i] += k * p[i]
        return r

    mono = lambda k: [one if i == k else one * 0 for i in range(deg)]  # noqa: E731
    c1, c2, c3, c4, c5, c6, c7, c8, c9, c10, c11, c12, c13, c14 = c
    X, X2, I = mono(1), mono(2), mono(0)
    y02 = mul(X2, lin((c1, X2), (c2, X)))
    y12 = lin((one, mul(lin((one, y02), (c3, X2), (c4, X)), lin((one, y02), (c5, X2)))), (c6, y02), (c7, X2))
    return lin((one, mul(lin((one, y12), (c8, X2), (c9, X)), lin((one, y12), (c10, y02), (c11, X)))),
               (c12, y12), (c13, y02), (c14, X2), (one, X), (one, I))


def test_s15_coefficients_give_the_taylor_polynomial():
    """The four-product form expands to 1 + x + ... + x^15/15! (each coefficient to 1e-14
    relative from the double coefficients), its x^16 coefficient is 0.5457/16!, and each of its
    fourteen coefficients matters (a 1e-6 relative change moves some Taylor coefficient by
    more than 1e-9 relative)."""
    p = _s15_expand(S15_COEF, 1.0)
    rel = [abs(p[k] * math.factorial(k) - 1.0) for k in range(16)]
    assert max(rel) < 1e-14, rel
    assert abs(p[16] * math.factorial(16) - 0.5457435) < 1e-6
    for i in range(14):
        c = list(S15_COEF)
        c[i] *= 1.0 + 1e-6
        q = _s15_expand(c, 1.0)
        assert max(abs(q[k] * math.factorial(k) - 1.0) for k in range(16)) > 1e-9, i


def test_s15_theta_is_the_backward_error_bound():
    """theta_15+ is where the relative backward error series of the form reaches the unit
    roundoff: h(x) = log(e^-x p(x)) = sum_{k>=16} h_k x^k (the truncation; the rounding of the
    coefficients to double leaves terms below x^16 of the size of rounding errors), with
    sum |h_k| theta^(k-1) <= 2^-53 at the kernel's theta and above it 1% further out."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    one = mp.mpf(1)
    p = _s15_expand([mp.mpf(c) for c in S15_COEF], one)
    n = 60
    P = p + [mp.mpf(0)] * (n - len(p))
    q = [sum(((-1) ** i / mp.factorial(i)) * P[k - i] for i in range(k + 1)) for k in range(n)]
    q[0] -= 1  # e^-x p(x) - 1: zero through x^15 but for the rounding of the double coefficients
    for k in range(16):  # that rounding: at most ~u relative at x^2, 1.2e-12 relative at x^15
        assert abs(q[k]) < 1e-11 / mp.factorial(k) + mp.mpf(10) ** -30
        q[k] = mp.mpf(0)
    q2 = [sum(q[i] * q[k - i] for i in range(k + 1)) for k in range(n)]  # from x^32
    h = [q[k] - q2[k] / 2 for k in range(n)]  # log(1 + q) through x^47 (q^3 starts at x^48)
    berr = lambda t: sum(abs(h[k]) * t ** (k - 1) for k in range(16, 48))  # noqa: E731
    u = mp.mpf(2) ** -53
    assert berr(mp.mpf(THETA15)) <= u
    assert berr(mp.mpf(THETA15) * 1.01) > u
