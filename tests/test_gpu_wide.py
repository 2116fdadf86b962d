"""Parity of the one-wave-per-sample kernel (16x16 padded matrices, M = exp(-tau V) by a scaled
and squared Taylor polynomial and the ring products on the FP64 matrix cores) against the
oracle, on the same host-seeded coordinates.

It runs automatically for 8 < A <= 16 and can be forced for any A (BoxData.force_wide), which
lets the same inputs go through both FP64 kernels. Tolerance as for the lanes-per-sample
kernel: per-sample s_rho, s_g, s_gP, s_gM within 1e-10 relative of the oracle.
"""
import numpy as np
import pytest

from helpers import make_data, oracle_setup, random_model, rel_err, stage_dict_models
from oracle import pimc_oracle as po
from pibronic_amd import _native, pimc, synthetic
from pibronic_amd.vibronic_model_keys import VibronicModelKeys as VMK

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _parity(FS, T, P, B, pm, seed, force_wide=True, tol=RTOL, debug=False):
    data = make_data(FS, T, P, B, B, pm=pm)
    data.force_wide = force_wide
    q, _ = data.draw_host_sample(np.random.default_rng(seed), B)
    r = pimc.evaluate_samples(data, q, pm=pm, debug=debug)
    s = oracle_setup(FS, T, P)
    o = po.evaluate(s, q, pm=pm, keep=debug)
    keys = ["s_rho", "s_g"] + (["s_gP", "s_gM"] if pm else [])
    errs = {k: rel_err(r[k], o[k]) for k in keys}
    errs["log_scale"] = np.max(np.abs(r["log_scale"] - o["log_scale"]))
    print(errs)
    for k in keys:
        assert errs[k] < tol, (k, errs)
    assert errs["log_scale"] < 1e-9 * max(1.0, np.max(np.abs(o["log_scale"]))), errs
    return errs, r, o, q, data


@pytest.mark.parametrize("pm", [False, True])
def test_wide_A4_N8_P64(tmp_path, pm):
    """The main workload's model through the wide kernel."""
    _announce(f"wide A4 N8 P64 pm={pm}")
    FS = stage_dict_models(tmp_path, synthetic.generate(4, 8, quadratic=True, seed=0))
    _parity(FS, 300.0, 64, 300, pm, seed=12)


@pytest.mark.parametrize("A", [1, 2, 3, 5, 8, 9, 11, 13, 16])
def test_wide_surface_counts(tmp_path, A):
    """Every padding depth, odd bead count, full (j != k) quadratic couplings, +-dbeta."""
    _announce(f"wide surfaces A={A} N=3 P=9 full Q")
    FS = stage_dict_models(tmp_path, random_model(A, 3, seed=A, quad="full"))
    _parity(FS, 300.0, 9, 64, True, seed=14)


@pytest.mark.parametrize("A,AR", [(10, 3), (4, 12), (12, 16), (2, 1)])
def test_wide_sampling_model_with_other_surface_count(tmp_path, A, AR):
    """A_rho != A (separate coordinate copy, pibronic/pimc/pimc.py:1094-1096)."""
    _announce(f"wide A={A} A_rho={AR}")
    m = random_model(A, 3, seed=7)
    rng = np.random.default_rng(AR)
    rho = {VMK.N: 3, VMK.A: AR, VMK.w: m[VMK.w].copy(), VMK.E: rng.uniform(0, 0.3, AR),
           VMK.G1: rng.normal(0, 0.05, (3, AR))}
    FS = stage_dict_models(tmp_path, m, rho)
    _parity(FS, 300.0, 10, 64, True, seed=15)


@pytest.mark.parametrize("pm", [False, True])
def test_wide_large_surface_workload(tmp_path, pm):
    """A=12, N=24, P=128, 250 K, off-diagonal couplings at 0.1 eV (near-degenerate V):
    per-sample results and the per-bead O-factors, S_p and M = exp(-tau V).

    The kernel forms the O-exponent in the expanded form; the oracle's reference-order
    expression carries up to ~1e-11 of rounding on this workload
    (test_exponent_form.py pins both), which bounds the per-bead O and S differences."""
    _announce(f"wide A12 N24 P128 pm={pm}")
    cfg = synthetic.WORKLOADS["A12_N24_P128_T250"]
    FS = stage_dict_models(tmp_path, synthetic.generate(12, 24, quadratic=True, off=cfg["off"], seed=0))
    errs, r, o, q, data = _parity(FS, cfg["T"], cfg["P"], 256, pm, seed=16, force_wide=False, debug=not pm)
    if not pm:
        # the debug records hold O before scaling; the oracle keeps them scaled by S_p
        assert rel_err(r["S"], o["S"]) < 1e-11
        S = r["S"][..., None]
        assert rel_err(r["O_vib"] / S, np.diagonal(o["O_vib"], axis1=2, axis2=3)) < 1e-11
        assert rel_err(r["O_rho"] / S, np.diagonal(o["O_rho"], axis1=2, axis2=3)) < 1e-11
        err_m = np.max(np.abs(r["M"] - o["M"]), axis=(2, 3)) / np.max(np.abs(o["M"]), axis=(2, 3))
        print("max relative M error", err_m.max())
        assert err_m.max() < 1e-12


def test_wide_agrees_with_lane_kernel(tmp_path):
    """Both FP64 kernels on the same coordinates (A=6, full Q, +-dbeta)."""
    _announce("wide vs lanes A6")
    FS = stage_dict_models(tmp_path, random_model(6, 4, seed=3, quad="full"))
    data = make_data(FS, 300.0, 24, 256, 256, pm=True)
    q, _ = data.draw_host_sample(np.random.default_rng(5), 256)
    a = pimc.evaluate_samples(data, q, pm=True)
    data.force_wide = True
    b = pimc.evaluate_samples(data, q, pm=True)
    for k in ["s_rho", "s_g", "s_gP", "s_gM"]:
        assert rel_err(b[k], a[k]) < 1e-11, k
    assert np.max(np.abs(a["log_scale"] - b["log_scale"])) < 1e-11


def test_wide_diagonal_model_gives_g_equal_rho(tmp_path):
    """No coupling: V = 0 after folding the diagonal, M = I, g == rho per sample."""
    _announce("wide g == rho")
    m = synthetic.generate(10, 3, quadratic=False, seed=9)
    A = 10
    m[VMK.E] = np.diag(np.diag(m[VMK.E]))
    m[VMK.G1] = np.stack([np.diag(np.diag(m[VMK.G1][j])) for j in range(3)])
    FS = stage_dict_models(tmp_path, m)
    data = make_data(FS, 300.0, 8, 128, 128)
    q, _ = data.draw_host_sample(np.random.default_rng(4), 128)
    r = pimc.evaluate_samples(data, q, pm=False)
    assert rel_err(r["s_g"], r["s_rho"]) < 1e-13


def test_more_than_16_surfaces_is_rejected(tmp_path):
    _announce("A=17 rejected")
    FS = stage_dict_models(tmp_path, random_model(17, 2, seed=1))
    data = make_data(FS, 300.0, 5, 8, 8)
    q, _ = data.draw_host_sample(np.random.default_rng(1), 8)
    with pytest.raises(_native.NativeError, match="no FP64 kernel"):
        pimc.evaluate_samples(data, q, pm=False)


# ------------------------------------------------------------------ BF16x3 ring products
# Stated bounds for PIB_EVAL_BF16X3 (the ring chain on the bf16 matrix cores with a hi/lo
# split and FP32 accumulation; M = exp(-tau V) and the O-factors FP64), against the FP64 wide
# kernel on the same coordinates: r = g/rho median relative error <= 1e-3; the largest
# |r_bf - r_64| relative to mean |r| (what a block average sees) <= 1e-2 (measured 3.2e-3 at
# A=12, N=24, P=128 on 512 device samples: one sample with r ~ 30 mean|r| off by 1e-4); the block's Z estimate
# moves by less than 0.05 of its standard error; rho and the log scale are FP64 paths and
# agree to 1e-12. The largest per-sample relative error is reported, not bounded: it comes
# from samples whose trace cancels (|g| << the terms summed), where any relative bound would be
# a statement about the cancellation, not about the products.
def _check_bf(rep):
    print(rep)
    assert rep["r"]["median_rel"] <= 1e-3, rep
    assert rep["r"]["max_abs_over_mean_abs"] <= 1e-2, rep
    assert abs(rep["z_shift_sigma"]) < 0.05, rep
    assert rep["s_rho"]["max_rel"] <= 1e-12, rep
    assert rep["log_scale_max_abs"] <= 1e-9, rep


@pytest.mark.parametrize("pm", [False, True])
def test_bf16x3_ring_large_surface_workload(tmp_path, pm):
    _announce(f"bf16x3 ring A12 N24 P128 pm={pm}")
    import torch
    from pibronic_amd.precision import precision_report
    cfg = synthetic.WORKLOADS["A12_N24_P128_T250"]
    FS = stage_dict_models(tmp_path, synthetic.generate(12, 24, quadratic=True, off=cfg["off"], seed=0))
    B = 512
    data = make_data(FS, cfg["T"], cfg["P"], B, B, pm=pm)
    q = torch.empty((B, 24, cfg["P"]), dtype=torch.float64, device="cuda")
    data.context().sample_block(3, 0, B, q)
    _check_bf(precision_report(data, q, B, variant="bf16x3"))


@pytest.mark.parametrize("A", [2, 5, 16])
def test_bf16x3_ring_surface_counts(tmp_path, A):
    _announce(f"bf16x3 ring A={A}")
    import torch
    from pibronic_amd.precision import precision_report
    FS = stage_dict_models(tmp_path, random_model(A, 3, seed=A, quad="full"))
    B = 256
    data = make_data(FS, 300.0, 20, B, B, pm=True)
    q = torch.empty((B, 3, 20), dtype=torch.float64, device="cuda")
    data.context().sample_block(4, 0, B, q)
    _check_bf(precision_report(data, q, B, variant="bf16x3"))


def test_bf16x3_ring_diagonal_model(tmp_path):
    """V = 0: M = I exactly, so g == rho up to the split representation of the ring state
    (hi + lo bf16 parts carry ~17 bits: ~1e-5 relative over a few beads)."""
    _announce("bf16x3 ring g == rho")
    m = synthetic.generate(10, 3, quadratic=False, seed=9)
    m[VMK.E] = np.diag(np.diag(m[VMK.E]))
    m[VMK.G1] = np.stack([np.diag(np.diag(m[VMK.G1][j])) for j in range(3)])
    FS = stage_dict_models(tmp_path, m)
    data = make_data(FS, 300.0, 8, 64, 64)
    data.ring_bf16x3 = True
    q, _ = data.draw_host_sample(np.random.default_rng(4), 64)
    r = pimc.evaluate_samples(data, q, pm=False)
    assert rel_err(r["s_g"], r["s_rho"]) < 1e-4
