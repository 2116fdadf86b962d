"""Per-sample parity against the oracle at the workload shapes themselves.

* A=8, N=16, P=64, 300 K (the learned-proposal workload's model): B = 256 host-seeded
  samples, plain and +-dbeta (block_compute / block_compute_pm,
  pibronic/pimc/pimc.py:1341-1462), within 1e-10 relative.
* A=6, N=12, sixteen temperatures 100-500 K with P_i = ceil(beta_i / tau0), tau0 = beta(100 K)/256
  (P 52-256), evaluated as one TemperatureBatch (the tuned schedule and the persistent lanes
  kernel) on B = 64 host-seeded samples per temperature: every temperature within 1e-10 of the
  oracle in FP64; the FP32 batch reports its error against FP64 per temperature.
"""
import numpy as np
import pytest

from helpers import make_data, oracle_setup, rel_err, stage_dict_models
from oracle import pimc_oracle as po
from pibronic_amd import pimc, synthetic
from pibronic_amd.batch import TemperatureBatch
from pibronic_amd.precision import _stats

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _check(r, o, pm, tol=RTOL):
    keys = ["s_rho", "s_g"] + (["s_gP", "s_gM"] if pm else [])
    errs = {k: rel_err(r[k], o[k]) for k in keys}
    errs["log_scale"] = float(np.max(np.abs(r["log_scale"] - o["log_scale"])))
    for k in keys:
        assert errs[k] < tol, (k, errs)
    assert errs["log_scale"] < 1e-9 * max(1.0, float(np.max(np.abs(o["log_scale"])))), errs
    return errs


@pytest.mark.parametrize("pm", [False, True])
def test_learned_workload_model_parity(tmp_path, pm):
    _announce(f"A8 N16 P64 T300 B=256 pm={pm}")
    cfg = synthetic.WORKLOADS["A8_N16_P64_T300"]
    FS = stage_dict_models(tmp_path, synthetic.generate(8, 16, quadratic=True, off=cfg["off"], seed=0))
    B = 256
    data = make_data(FS, cfg["T"], cfg["P"], B, B, pm=pm)
    q, _ = data.draw_host_sample(np.random.default_rng(31), B)
    r = pimc.evaluate_samples(data, q, pm=pm)
    o = po.evaluate(oracle_setup(FS, cfg["T"], cfg["P"]), q, pm=pm)
    print(_check(r, o, pm))


def _c3_batch(tmp_path, B, fp32=False):
    FS = stage_dict_models(tmp_path, synthetic.generate(6, 12, quadratic=True, off=0.05, seed=0))
    Ts, Ps, _ = synthetic.c3_temperatures()
    assert len(Ts) == 16 and Ps.min() == 52 and Ps.max() == 256
    datas = [make_data(FS, float(T), int(P), B, B) for T, P in zip(Ts, Ps)]
    qs = [d.draw_host_sample(np.random.default_rng(100 + i), B)[0] for i, d in enumerate(datas)]
    return FS, Ts, Ps, datas, qs, TemperatureBatch(datas, B, fp32=fp32)


def test_sixteen_temperature_batch_parity(tmp_path):
    _announce("A6 N12 16 temperatures P 52-256 batch B=64 FP64")
    B = 64
    FS, Ts, Ps, datas, qs, tb = _c3_batch(tmp_path, B)
    tuned = tb.tune(seed=3)
    print("tuned schedule", tuned)
    runs = {"tuned": tb.evaluate(qs)}
    tb.set_schedule("lanes", None, tile_waves=4, grid_waves=1024)
    runs["persistent lanes"] = tb.evaluate(qs)
    tb.set_schedule("wide", None, tile_waves=8, grid_waves=1024)
    runs["persistent wide"] = tb.evaluate(qs)
    for i, (T, P) in enumerate(zip(Ts, Ps)):
        o = po.evaluate(oracle_setup(FS, float(T), int(P)), qs[i])
        for name, res in runs.items():
            errs = _check(res[i], o, False)
            print(f"T={T:.1f} P={P} {name}: {errs}", flush=True)


def test_sixteen_temperature_batch_fp32_error(tmp_path):
    """FP32 batch against the FP64 batch on the same samples, per temperature (the bounds of
    the FP32 kernel: r median relative error <= 1e-4, Z shift < 0.05 sigma)."""
    _announce("A6 N12 16 temperatures batch FP32 vs FP64")
    B = 2048
    FS, Ts, Ps, datas, qs, tb64 = _c3_batch(tmp_path, B)
    r64 = tb64.evaluate(qs)
    tb32 = TemperatureBatch(datas, B, fp32=True)
    r32 = tb32.evaluate(qs)
    for i, (T, P) in enumerate(zip(Ts, Ps)):
        a = r64[i]["s_g"] / r64[i]["s_rho"]
        b = r32[i]["s_g"] / r32[i]["s_rho"]
        st = _stats(b, a)
        st["max_abs_over_mean_abs"] = float(np.max(np.abs(b - a)) / np.mean(np.abs(a)))
        shift = (b.mean() - a.mean()) / (a.std() / np.sqrt(B - 1))
        print(f"T={T:.1f} P={P}: r error {st}, Z shift {shift:+.2e} sigma", flush=True)
        assert st["median_rel"] <= 1e-4 and abs(shift) < 0.05, (T, P, st, shift)
