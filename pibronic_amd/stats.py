"""Statistics of a PIMC run: Z_MC, E, Cv with jackknife errors, and the *_thermo files.

Formulas as pibronic/stats/stats.py:38-130 and pibronic/stats/jackknife.py:14-105. The
reference reads Z_rho, E_rho, Cv_rho and alpha+- = Z_rho(beta)/Z_rho(beta +- dbeta) from a
Julia-generated analytic_results.json (pibronic/data/postprocessing.py:313-343); here they
come from the closed form of the harmonic sampling model (`analytic_sampling`). Per-block
FP64 sums reduced on the GPU (pib_block_sums) give the basic estimators without the
per-sample arrays (`estimate_from_block_sums`).

Two quirks of the reference are kept and documented: the "basic" E and Cv already estimate
the full derivative, yet add_harmonic_contribution adds E_rho and Cv_rho on top (quirk Q2;
only the "alpha" variant is consistent with that addition), and g(beta +- dbeta) uses M at
the base tau (quirk Q1).
"""
import glob
import json
import os

import numpy as np

from . import constants
from .constants import boltzman
from . import _native


# ------------------------------------------------------------------ per-sample terms
def calculate_basic_property_terms(delta_beta, rho, g, g_plus, g_minus):
    """g/rho and the symmetric first and second derivative terms (stats.py:38-55)."""
    ratio = g / rho
    d1 = (g_plus - g_minus) / rho
    d1 /= (2. * delta_beta)
    d2 = (g_plus - (2. * g) + g_minus) / rho
    d2 /= pow(delta_beta, 2)
    return [ratio, d1, d2]


def calculate_alpha_terms(delta_beta, rho, g, g_plus, g_minus, alpha_plus, alpha_minus):
    """Terms of the alpha variant (stats.py:58-81)."""
    ratio = g / rho
    ratio_plus = g_plus / rho
    ratio_minus = g_minus / rho
    d1 = np.mean(ratio_plus) * np.mean(alpha_plus)
    d1 -= np.mean(ratio_minus) * np.mean(alpha_minus)
    d1 /= (2. * delta_beta)
    d2 = np.mean(ratio_plus) * np.mean(alpha_plus)
    d2 -= 2. * np.mean(ratio)
    d2 += np.mean(ratio_minus) * np.mean(alpha_minus)
    d2 /= pow(delta_beta, 2)
    return [ratio, d1, d2, ratio_plus, alpha_plus, ratio_minus, alpha_minus]


def basic_estimate_Z_monte_carlo(g_over_rho, number_of_samples):
    """stats.py:84-90"""
    Z = np.mean(g_over_rho)
    err = np.std(g_over_rho, ddof=0) / np.sqrt(number_of_samples - 1)
    return Z, err


def basic_estimate_internal_energy(d1, Z):
    """stats.py:93-97"""
    return -1. * np.mean(d1) / np.mean(Z), 0.0


def basic_estimate_heat_capacity(d2, Z, E, T):
    """stats.py:100-106"""
    Cv = np.mean(d2) / np.mean(Z)
    Cv -= pow(E, 2.)
    Cv /= boltzman * pow(T, 2.)
    return Cv, 0.0


def estimate_basic_properties(X, T, g_r, sym1, sym2):
    """stats.py:109-123"""
    Z, Z_err = basic_estimate_Z_monte_carlo(g_r, X)
    E, E_err = basic_estimate_internal_energy(sym1, Z)
    Cv, Cv_err = basic_estimate_heat_capacity(sym2, Z, E, T)
    return {"Z": Z, "Z error": Z_err, "E": E, "E error": E_err, "Cv": Cv, "Cv error": Cv_err}


def add_harmonic_contribution(d, E_sampling, Cv_sampling):
    """stats.py:126-130"""
    d["E"] += E_sampling
    d["Cv"] += Cv_sampling


# ------------------------------------------------------------------ jackknife
def calculate_jackknife_term(length, array):
    """Leave-one-out means (jackknife.py:60-70)."""
    jk = np.full(shape=length, fill_value=np.sum(array))
    jk -= array
    jk /= (length - 1)
    return jk


def calculate_jackknife_terms(length, arrays):
    return [calculate_jackknife_term(length, a) for a in arrays]


def calculate_alpha_jackknife_terms(X, delta_beta, ratio, LN, alpha_plus, RN, alpha_minus):
    """jackknife.py:14-57"""
    jk_ratio, jk_LN, jk_RN = calculate_jackknife_terms(X, [ratio, LN, RN])
    jk_d1 = jk_LN * alpha_plus - jk_RN * alpha_minus
    jk_d1 /= (2. * delta_beta)
    jk_d2 = jk_LN * alpha_plus - 2. * jk_ratio + jk_RN * alpha_minus
    jk_d2 /= pow(delta_beta, 2)
    return [jk_ratio, jk_d1, jk_d2]


def estimate_jackknife(X, T, delta_beta, basic, jk_g_rho, jk_sym1, jk_sym2):
    """jackknife.py:77-105"""
    f_E = -1. * jk_sym1 / jk_g_rho
    E = X * basic["E"] - (X - 1.) * np.mean(f_E)
    E_err = np.sqrt(X - 1.) * np.std(f_E, ddof=1)
    f_C = jk_sym2 / jk_g_rho
    f_C -= pow(f_E, 2.)
    f_C /= boltzman * pow(T, 2.)
    Cv = X * basic["Cv"] - (X - 1.) * np.mean(f_C)
    Cv_err = np.sqrt(X - 1.) * np.std(f_C, ddof=1)
    return {"E": E, "E error": E_err, "Cv": Cv, "Cv error": Cv_err}


# ------------------------------------------------------------------ analytic sampling model
def analytic_sampling(rho, beta, delta_beta=constants.delta_beta):
    """Closed-form Z, E, Cv of the harmonic sampling distribution and alpha+-.

    rho is a ModelSampling (before or after precompute; uses energy + delta, omega). Z_rho
    = sum_a e^{-beta E~_a} prod_j 1/(2 sinh(beta w_j/2)); E = -d ln Z/d beta; Cv =
    (d^2 ln Z / d beta^2)/(k_B T^2). These replace the Julia values the reference reads
    (postprocessing.py:313-343) and agree with its stored sos_B80.json values.
    """
    w = np.asarray(rho.omega, dtype=np.float64)
    tilde = np.asarray(rho.tilde_energy, dtype=np.float64)

    def z(b):
        return np.sum(np.exp(-b * tilde) * np.prod(1.0 / (2.0 * np.sinh(b * w / 2.0))))

    Z = z(beta)
    terms = np.exp(-beta * tilde) * np.prod(1.0 / (2.0 * np.sinh(beta * w / 2.0)))
    dlog = -tilde - np.sum(0.5 * w / np.tanh(beta * w / 2.0))
    E = -np.sum(terms * dlog) / Z
    var = np.sum(terms * dlog ** 2) / Z - (np.sum(terms * dlog) / Z) ** 2
    d2 = var + np.sum((0.5 * w) ** 2 / np.sinh(beta * w / 2.0) ** 2)
    T = constants.extract_T_from_beta(beta)
    return {"Z": Z, "E": E, "Cv": d2 / (boltzman * T ** 2),
            "alpha_plus": Z / z(beta + delta_beta), "alpha_minus": Z / z(beta - delta_beta)}


# ------------------------------------------------------------------ analyses
def _arrays(result):
    return [result.scaled_rho, result.scaled_g, result.scaled_gofr_plus, result.scaled_gofr_minus]


def basic_statistical_analysis(T, result, analytic, delta_beta=constants.delta_beta):
    """stats.py:174-190"""
    terms = calculate_basic_property_terms(delta_beta, *_arrays(result))
    d = estimate_basic_properties(result.samples, T, *terms)
    add_harmonic_contribution(d, analytic["E"], analytic["Cv"])
    return d


def alpha_statistical_analysis(T, result, analytic, delta_beta=constants.delta_beta):
    """stats.py:193-211"""
    terms = calculate_alpha_terms(delta_beta, *_arrays(result), analytic["alpha_plus"],
                                  analytic["alpha_minus"])[:-4]
    d = estimate_basic_properties(result.samples, T, *terms)
    add_harmonic_contribution(d, analytic["E"], analytic["Cv"])
    return d


def basic_jackknife_analysis(T, result, analytic, delta_beta=constants.delta_beta):
    """stats.py:271-299"""
    X = result.samples
    terms = calculate_basic_property_terms(delta_beta, *_arrays(result))
    jk_terms = calculate_jackknife_terms(X, terms)
    basic = estimate_basic_properties(X, T, *terms)
    jk = estimate_jackknife(X, T, delta_beta, basic, *jk_terms)
    add_harmonic_contribution(basic, analytic["E"], analytic["Cv"])
    add_harmonic_contribution(jk, analytic["E"], analytic["Cv"])
    out = dict(basic)
    out.update({"jk_" + k: v for k, v in jk.items()})
    return out


def alpha_jackknife_analysis(T, result, analytic, delta_beta=constants.delta_beta):
    """stats.py:302-333"""
    X = result.samples
    terms = calculate_alpha_terms(delta_beta, *_arrays(result), analytic["alpha_plus"],
                                  analytic["alpha_minus"])
    jk_terms = calculate_alpha_jackknife_terms(X, delta_beta, terms[0], *terms[3:7])
    terms = terms[:-4]
    basic = estimate_basic_properties(X, T, *terms)
    jk = estimate_jackknife(X, T, delta_beta, basic, *jk_terms)
    add_harmonic_contribution(basic, analytic["E"], analytic["Cv"])
    add_harmonic_contribution(jk, analytic["E"], analytic["Cv"])
    out = dict(basic)
    out.update({"jk_" + k: v for k, v in jk.items()})
    return out


def estimate_from_block_sums(sums, T, delta_beta=constants.delta_beta):
    """Z_MC +- error, the basic E and Cv, and the Kish effective sample size from the
    FP64 per-block sums of pib_block_sums (rows of NUM_SUMS, summed over blocks).

    Z and its error equal basic_estimate_Z_monte_carlo on the per-sample arrays; E and Cv
    equal the basic estimators before the harmonic contribution is added.
    """
    s = np.asarray(sums, dtype=np.float64)
    s = s.sum(axis=0) if s.ndim == 2 else s
    n = s[_native.SUM_N]
    Z = s[_native.SUM_R] / n
    var = max(s[_native.SUM_R2] / n - Z * Z, 0.0)
    out = {"Z": Z, "Z error": np.sqrt(var) / np.sqrt(n - 1), "samples": n,
           "ESS": s[_native.SUM_ABSR] ** 2 / s[_native.SUM_R2],
           "negative fraction": s[_native.SUM_NEG] / n}
    if s[_native.SUM_D1] != 0.0 or s[_native.SUM_D2] != 0.0:
        E = -(s[_native.SUM_D1] / n) / Z
        Cv = ((s[_native.SUM_D2] / n) / Z - E * E) / (boltzman * T ** 2)
        out.update({"E": E, "Cv": Cv})
    return out


# ------------------------------------------------------------------ files
def _parse_pimc_name(path):
    """P and T of P{P}_T{T:.2f}_J{J}_data_points.npz (postprocessing.py:118-136)."""
    name = os.path.basename(path)
    P = int(name.split("_T")[0][1:])
    T = float(name.split("_T")[1].split("_J")[0])
    return P, T


def jackknife_analysis_of_pimc(FS, data_rho, method="basic", delta_beta=constants.delta_beta):
    """Write one P{P}_T{T}_X{X}_thermo JSON per (P, T) found in FS.path_rho_results.

    The reference's driver (stats.py:336-377) farms this out to a process pool and reads
    Julia analytic files; here it runs in-process and uses analytic_sampling(data_rho, beta).
    data_rho is the sampling model (ModelSampling) of the run. Returns the written paths.
    """
    from .pimc import BoxResultPM
    FS.generate_model_hashes()
    op = {"basic": basic_jackknife_analysis, "alpha": alpha_jackknife_analysis}[method]
    files = glob.glob(os.path.join(FS.path_rho_results, "P*_T*_J*_data_points.npz"))
    keys = sorted({_parse_pimc_name(f) for f in files})
    written = []
    for P, T in keys:
        # job order (J = 0, 1, ...), so the merged arrays and sums do not depend on glob order
        mine = sorted((f for f in files if _parse_pimc_name(f) == (P, T)),
                      key=lambda f: int(os.path.basename(f).split("_J")[1].split("_")[0]))
        res = BoxResultPM()
        res.load_multiple_results(mine)
        analytic = analytic_sampling(data_rho, constants.beta(T), delta_beta)
        out = op(T, res, analytic, delta_beta)
        out = {k: float(v) for k, v in out.items()}
        out["hash_vib"], out["hash_rho"] = FS.hash_vib, FS.hash_rho
        path = FS.template_jackknife.format(P=P, T=T, X=int(res.samples))
        with open(path, "w", encoding="UTF8") as fh:
            fh.write(json.dumps(out))
        written.append(path)
    return written
