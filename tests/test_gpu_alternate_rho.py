"""Sampling models with another surface count than the coupled model, from the reference's own
paper examples: Jahn-Teller model 2 (A = 2) with its three sampling models
(examples/paper_1.5025058/alternate_rhos/jahnteller_D2_R{0,1,2}.json, A_rho = 2, 4, 8), the
reference's qTempTensor path (pibronic/pimc/pimc.py:625-626, 643-644, 1094-1096).

* per-sample s_rho, s_g, s_gP, s_gM within 1e-10 of the oracle (pinned for A_rho != A in
  tests/test_closed_form.py), through pib_eval_block and through the batched launcher;
* Z_rho (closed form) * <g/rho> from device samples: the same Z_g for all three sampling
  models, every pair within 2 sigma.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from helpers import make_data, oracle_setup, rel_err, stage_models
from oracle import pimc_oracle as po
from pibronic_amd import pimc, stats
from pibronic_amd.batch import TemperatureBatch

pytestmark = pytest.mark.gpu

ALT = os.path.join(GOLDEN, "alternate_rhos")


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _fs(tmp_path, r):
    d = tmp_path / f"R{r}"
    d.mkdir()
    return stage_models(d, os.path.join(ALT, "jahnteller_2.json"),
                        os.path.join(ALT, f"jahnteller_D2_R{r}.json"), id_rho=r)


@pytest.mark.parametrize("r", [0, 1, 2])
def test_alternate_rho_per_sample_parity(tmp_path, r):
    _announce(f"jahnteller_2 rho R{r} parity")
    FS = _fs(tmp_path, r)
    B, T, P = 512, 300.0, 16
    data = make_data(FS, T, P, B, B, pm=True)
    assert data.rho.states == (2, 4, 8)[r]
    q, _ = data.draw_host_sample(np.random.default_rng(40 + r), B)
    res = pimc.evaluate_samples(data, q, pm=True)
    o = po.evaluate(oracle_setup(FS, T, P), q, pm=True)
    for k in ("s_rho", "s_g", "s_gP", "s_gM"):
        assert rel_err(res[k], o[k]) < 1e-10, k


def test_alternate_rho_batched_launch(tmp_path):
    """A = 2, A_rho = 8 at three temperatures in one persistent launch (lanes and wide
    families) against the oracle."""
    _announce("jahnteller_2 rho R2 batched")
    FS = _fs(tmp_path, 2)
    B = 128
    TP = [(200.0, 24), (300.0, 16), (450.0, 11)]
    datas = [make_data(FS, T, P, B, B, pm=True) for T, P in TP]
    qs = [d.draw_host_sample(np.random.default_rng(60 + i), B)[0] for i, d in enumerate(datas)]
    tb = TemperatureBatch(datas, B)
    for fam in ("lanes", "wide"):
        tb.set_schedule(fam, None, tile_waves=2, grid_waves=256)
        res = tb.evaluate(qs)
        for i, (T, P) in enumerate(TP):
            o = po.evaluate(oracle_setup(FS, T, P), qs[i], pm=True)
            for k in ("s_rho", "s_g", "s_gP", "s_gM"):
                assert rel_err(res[i][k], o[k]) < 1e-10, (fam, T, k)


def test_alternate_rhos_agree_on_z(tmp_path):
    _announce("jahnteller_2 Z over three sampling models")
    T, P, B, blocks = 300.0, 16, 1 << 17, 8
    est = []
    for r in range(3):
        FS = _fs(tmp_path, r)
        data = make_data(FS, T, P, B * blocks, B, seed=900 + r)
        res = pimc.block_compute(data, pimc.BoxResult(data=data))
        x = res.scaled_g / res.scaled_rho
        zr = stats.analytic_sampling(data.rho, data.beta)["Z"]
        est.append((zr * x.mean(), zr * x.std() / np.sqrt(x.size - 1)))
    print("Z_g from R0, R1, R2:", est, flush=True)
    for i in range(3):
        for j in range(i + 1, 3):
            (a, ea), (b, eb) = est[i], est[j]
            assert abs(a - b) < 2 * np.hypot(ea, eb), est
