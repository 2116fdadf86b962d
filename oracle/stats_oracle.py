"""Restatement of Pibronic's block statistics (TEST INFRASTRUCTURE ONLY).

pibronic/stats/stats.py:38-130 (basic and alpha terms, Z/E/Cv estimators, harmonic
contribution) and pibronic/stats/jackknife.py:14-105 (leave-one-out jackknife).
"""
import numpy as np

from .pimc_oracle import BOLTZMANN_EV_PER_K


def basic_terms(delta_beta, rho, g, g_plus, g_minus):
    """r = g/rho, d1, d2 (pibronic/stats/stats.py:38-55)."""
    ratio = g / rho
    d1 = (g_plus - g_minus) / rho
    d1 /= (2. * delta_beta)
    d2 = (g_plus - (2. * g) + g_minus) / rho
    d2 /= pow(delta_beta, 2)
    return [ratio, d1, d2]


def alpha_terms(delta_beta, rho, g, g_plus, g_minus, alpha_plus, alpha_minus):
    """pibronic/stats/stats.py:58-81."""
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


def estimate_basic(X, T, ratio, d1, d2):
    """Z_MC +- err, E, Cv (pibronic/stats/stats.py:84-123)."""
    Z = np.mean(ratio)
    Z_err = np.std(ratio, ddof=0) / np.sqrt(X - 1)
    E = -1. * np.mean(d1) / np.mean(Z)
    Cv = np.mean(d2) / np.mean(Z)
    Cv -= pow(E, 2.)
    Cv /= BOLTZMANN_EV_PER_K * pow(T, 2.)
    return {"Z": Z, "Z error": Z_err, "E": E, "E error": 0.0, "Cv": Cv, "Cv error": 0.0}


def add_harmonic(d, E_sampling, Cv_sampling):
    """pibronic/stats/stats.py:126-130."""
    d["E"] += E_sampling
    d["Cv"] += Cv_sampling


def jackknife_term(length, array):
    """pibronic/stats/jackknife.py:60-70."""
    jk = np.full(shape=length, fill_value=np.sum(array))
    jk -= array
    jk /= (length - 1)
    return jk


def estimate_jackknife(X, T, delta_beta, basic, jk_ratio, jk_d1, jk_d2):
    """pibronic/stats/jackknife.py:77-105."""
    f_E = -1. * jk_d1 / jk_ratio
    E = X * basic["E"] - (X - 1.) * np.mean(f_E)
    E_err = np.sqrt(X - 1.) * np.std(f_E, ddof=1)
    f_C = jk_d2 / jk_ratio
    f_C -= pow(f_E, 2.)
    f_C /= BOLTZMANN_EV_PER_K * pow(T, 2.)
    Cv = X * basic["Cv"] - (X - 1.) * np.mean(f_C)
    Cv_err = np.sqrt(X - 1.) * np.std(f_C, ddof=1)
    return {"E": E, "E error": E_err, "Cv": Cv, "Cv error": Cv_err}


def basic_jackknife(T, delta_beta, rho, g, gp, gm, E_sampling=0.0, Cv_sampling=0.0):
    """pibronic/stats/stats.py:271-299 on in-memory arrays."""
    X = rho.shape[0]
    terms = basic_terms(delta_beta, rho, g, gp, gm)
    jk = [jackknife_term(X, t) for t in terms]
    basic = estimate_basic(X, T, *terms)
    jkd = estimate_jackknife(X, T, delta_beta, basic, *jk)
    add_harmonic(basic, E_sampling, Cv_sampling)
    add_harmonic(jkd, E_sampling, Cv_sampling)
    out = dict(basic)
    for k, v in jkd.items():
        out["jk_" + k] = v
    return out
