"""Reduced-precision kernel (PIB_EVAL_FP32) against the FP64 kernel and the golden vectors.

Stated FP32 bounds (per-sample, same host- or device-seeded coordinates):
  * the estimator weight r = g / rho: median relative error <= 1e-4 and largest <= 1e-2 (a
    product of P FP32 matrix factors: the error grows with P and the conditioning of the
    chain; measured 1.3e-5 / 1.6e-3 at A=4, N=8, P=64);
  * the block's Z estimate mean(r) moves by less than 0.05 of its standard error;
  * s_rho (a sum of positive terms) within 1e-4 relative, log scale within 1e-3 absolute.
The largest |r32 - r64| relative to mean |r| is reported, not bounded: with heavy-tailed
weights it is set by the one dominant sample.
"""
import os

import numpy as np
import pytest
import torch

from conftest import EXPLICIT, TEST_MODELS, golden
from helpers import make_data, random_model, rel_err, stage_dict_models, stage_models
from pibronic_amd import _native, pimc, synthetic
from pibronic_amd.precision import precision_report

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _check(rep):
    print(rep)
    assert rep["r"]["median_rel"] <= 1e-4, rep
    assert rep["r"]["max_rel"] <= 1e-2, rep
    assert abs(rep["z_shift_sigma"]) < 0.05, rep
    assert rep["s_rho"]["max_rel"] <= 1e-4, rep
    assert rep["log_scale_max_abs"] <= 1e-3, rep


@pytest.mark.parametrize("A,N,P,T,pm", [(4, 8, 64, 300.0, True), (6, 12, 256, 100.0, False),
                                        (6, 12, 52, 500.0, True), (2, 3, 12, 300.0, True)])
def test_fp32_against_fp64(tmp_path, A, N, P, T, pm):
    _announce(f"fp32 vs fp64 A{A} N{N} P{P} T{T} pm={pm}")
    FS = stage_dict_models(tmp_path, synthetic.generate(A, N, quadratic=A > 2, seed=0))
    B = 4096
    data = make_data(FS, T, P, B, B, pm=pm)
    q = torch.empty((B, N, P), dtype=torch.float64, device="cuda")
    data.context().sample_block(5, 0, B, q)
    _check(precision_report(data, q, B))


@pytest.mark.parametrize("A", [1, 3, 5, 7, 8])
def test_fp32_surface_counts(tmp_path, A):
    """Every compiled surface count with full (j != k) quadratic couplings."""
    _announce(f"fp32 surfaces A={A}")
    FS = stage_dict_models(tmp_path, random_model(A, 3, seed=A, quad="full"))
    B = 1024
    data = make_data(FS, 300.0, 9, B, B, pm=True)
    q = torch.empty((B, 3, 9), dtype=torch.float64, device="cuda")
    data.context().sample_block(6, 0, B, q)
    _check(precision_report(data, q, B))


def test_fp32_explicit_golden_vectors(tmp_path):
    """The reference's explicit example (unscaled g and rho) within FP32 accuracy."""
    _announce("fp32 explicit golden")
    FS = stage_models(tmp_path, os.path.join(EXPLICIT, "coupled_model.json"),
                      os.path.join(EXPLICIT, "sampling_model.json"))
    data = make_data(FS, 300.0, 5, 10, 10)
    data.fp32 = True
    q = golden("samples.npy")[:, 0].copy()
    r = pimc.evaluate_samples(data, q, pm=False, noscale=True)
    assert rel_err(r["s_g"], golden("numerator(g).npy")) < 1e-5
    assert rel_err(r["s_rho"], golden("denominator(rho).npy")) < 1e-5


def test_fp32_diagonal_model_gives_g_equal_rho(tmp_path):
    _announce("fp32 g == rho")
    base = os.path.join(TEST_MODELS, "data_set_0")
    FS = stage_models(tmp_path, os.path.join(base, "parameters", "coupled_model.json"),
                      os.path.join(base, "rho_0", "parameters", "sampling_model.json"))
    data = make_data(FS, 300.0, 12, 512, 512)
    data.fp32 = True
    q, _ = data.draw_host_sample(np.random.default_rng(4), 512)
    r = pimc.evaluate_samples(data, q, pm=False)
    assert rel_err(r["s_g"], r["s_rho"]) < 1e-5


def test_fp32_batch_matches_single(tmp_path):
    """An FP32 batch (one launch per temperature) equals the FP32 kernel run alone."""
    _announce("fp32 batch")
    from pibronic_amd.batch import TemperatureBatch
    FS = stage_dict_models(tmp_path, random_model(3, 4, seed=2))
    B = 256
    datas = [make_data(FS, T, P, B, B, pm=True) for T, P in [(200.0, 30), (400.0, 15)]]
    qs = [d.draw_host_sample(np.random.default_rng(i), B)[0] for i, d in enumerate(datas)]
    tb = TemperatureBatch(datas, B, fp32=True)
    assert tb.schedule()["family"] == "sequential"
    lanes = [d.context().lanes_per_sample(B) for d in datas]
    tb.set_schedule("sequential", lanes)
    res = tb.evaluate(qs)
    for i, (d, q) in enumerate(zip(datas, qs)):
        d.fp32 = True
        single = pimc.evaluate_samples(d, q, pm=True)
        for k in ["s_rho", "s_g", "s_gP", "s_gM", "log_scale"]:
            assert np.array_equal(res[i][k], single[k]), (i, k)


def test_fp32_errors(tmp_path):
    _announce("fp32 errors")
    FS = stage_dict_models(tmp_path, random_model(10, 2, seed=1))
    data = make_data(FS, 300.0, 6, 8, 8)
    data.fp32 = True
    q, _ = data.draw_host_sample(np.random.default_rng(1), 8)
    with pytest.raises(_native.NativeError, match="no FP32 kernel for A=10"):
        pimc.evaluate_samples(data, q, pm=False)
    data.force_wide = True
    with pytest.raises(_native.NativeError, match="FP32"):
        pimc.evaluate_samples(data, q, pm=False)
