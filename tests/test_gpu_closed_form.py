"""The learned proposal's Z estimate against an exact value.

The estimator of pibronic/pimc/pimc.py:1341-1385 with the statistics of
pibronic/stats/stats.py:84-90 gives Z_g/Z_rho as the mean weight. With every inter-surface
coupling removed, g(R) is a sum of Gaussians and Z_g/Z_rho has a closed form
(oracle.pimc_oracle.log_z_ratio_surface_diagonal, pinned in tests/test_closed_form.py), while
the mode-diagonal quadratic draws of the workload models keep rho's tails as light relative to
g as in the coupled models. A trained flow (8 coupling layers, hidden 256, the workloads'
architecture) must then reproduce the exact value within 2 sigma of its own error estimate,
and so must the harmonic rho wherever its effective sample count reaches 1e3.
"""
import math

import numpy as np
import pytest

from helpers import make_data, oracle_setup, stage_dict_models
from oracle import pimc_oracle as po
from pibronic_amd import learn, flow, synthetic

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _max_share(logw):
    lw = logw.double().cpu().numpy()
    lw = lw[np.isfinite(lw)]
    w = np.exp(lw - lw.max())
    return float(w.max() / w.sum())


@pytest.mark.parametrize("name,steps", [("A4_N8_P64_T300_diag", 600), ("A8_N16_P64_T300_diag", 600)])
def test_learned_z_matches_closed_form(tmp_path, name, steps):
    _announce(f"closed-form Z {name}, {steps} training steps")
    cfg = synthetic.WORKLOADS[name]
    m = synthetic.generate(cfg["A"], cfg["N"], cfg["quadratic"], cfg["off"], 0, surface_diagonal=True)
    FS = stage_dict_models(tmp_path, m)
    data = make_data(FS, cfg["T"], cfg["P"], 4096, 4096, seed=11)
    exact = math.exp(po.log_z_ratio_surface_diagonal(oracle_setup(FS, cfg["T"], cfg["P"])))
    fl = flow.Flow(data, layers=8, hidden=256, seed=3, gemm="bf16x3")
    learn.train(fl, steps, 2048, 1e-3, 5)
    fl.draw_precision = "bf16"
    n = 1 << 20
    _, lw_f, sg_f = learn._draw_flow(fl, n, 21, 7 * 10 ** 9, 65536)
    _, lw_r, sg_r = learn._draw_rho(data, n, 22, 8 * 10 ** 9, 65536)
    f, r = learn.weight_stats(lw_f, sg_f), learn.weight_stats(lw_r, sg_r)
    dev_f = (f["z"] - exact) / f["z_err"]
    dev_r = (r["z"] - exact) / r["z_err"]
    print(f"{name}: exact {exact:.8g}; flow {f['z']:.8g} +- {f['z_err']:.3g} ({dev_f:+.2f} sigma, "
          f"ESS frac {f['ess_frac']:.4f}, max weight share {_max_share(lw_f):.2e}); harmonic "
          f"{r['z']:.8g} +- {r['z_err']:.3g} ({dev_r:+.2f} sigma, ESS {r['ess_frac'] * n:.4g}, "
          f"max weight share {_max_share(lw_r):.2e})", flush=True)
    assert f["dropped"] == 0 and f["ess_frac"] > 0.2
    assert abs(dev_f) < 2.0, (f, exact)
    if r["ess_frac"] * n >= 1e3:
        assert abs(dev_r) < 2.0, (r, exact)
