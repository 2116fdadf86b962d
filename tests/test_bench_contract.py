"""bench.py's pure parts on the CPU: the algorithmic flop counts its rooflines divide by, the
roofline record of the evaluation kernel, and the refusal of more than one GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _by_hand(A, N, P, nq, pm):
    """The per-sample count written out term by term: V assembly P A^2 (2N + 2nq + 1), the
    eigendecomposition's 9 A^3 and M's 2 A^3 + A^2 per bead, A (7N + 1) per O-factor set and
    bead (2 sets, 4 with +-dbeta), 2 A^3 + A^2 per ring product and bead (1, or 3 with
    +-dbeta), 5 P A for rho and the scaling."""
    sets, prods = (4, 3) if pm else (2, 1)
    per_bead = (A * A * (2 * N + 2 * nq + 1) + 9 * A ** 3 + 2 * A ** 3 + A * A + sets * A * (7 * N + 1)
                + prods * (2 * A ** 3 + A * A) + 5 * A)
    return P * per_bead


@pytest.mark.parametrize("A,N,P,nq,pm", [(4, 8, 64, 0, False), (8, 16, 64, 0, False), (4, 8, 64, 0, True),
                                         (12, 24, 128, 3, False), (2, 3, 12, 0, True)])
def test_algorithmic_flops(A, N, P, nq, pm):
    fl = bench.algorithmic_flops(A, N, P, nq, pm)
    assert fl["eval"] == _by_hand(A, N, P, nq, pm)
    assert fl["sample"] == 2 * N * P * P
    assert fl["total"] == fl["eval"] + fl["sample"]


def test_quoted_workloads():
    # the per-sample figures the measured table and the profiles quote; n_quad_pairs counts the
    # nonzero (j, k) blocks of the quadratic couplings, N of them for the seeded workloads'
    # mode-diagonal couplings
    assert bench.algorithmic_flops(4, 8, 64, 8)["eval"] == 119_552
    assert bench.algorithmic_flops(8, 16, 64, 16)["eval"] == 818_688


def test_eval_roofline_record():
    B, ms = 100_000, 0.25
    r = bench.eval_roofline(4, 8, 64, 8, False, B, ms, "k_eval_fp64<4, 4, false, false>")
    assert (r["bound"], r["unit"], r["peak"]) == ("valu", "TFLOP/s", bench.FP64_PEAK_TFLOPS)
    assert r["achieved"] == pytest.approx(119_552 * B / (ms * 1e-3) / 1e12)
    assert r["frac"] == pytest.approx(r["achieved"] / bench.FP64_PEAK_TFLOPS)
    assert r["flops_per_launch"] == 119_552 * B
    # coordinates in, rho, g and log S out (two more outputs with +-dbeta)
    assert r["algorithmic_bytes"] == B * 8 * 64 * 8 + B * 3 * 8
    rp = bench.eval_roofline(4, 8, 64, 8, True, B, ms, "k")
    assert rp["algorithmic_bytes"] == B * 8 * 64 * 8 + B * 5 * 8
    assert r["traffic"] is None and r["hbm_frac"] is None
    t = bench.eval_roofline(4, 8, 64, 8, False, B, ms, "k", traffic=4e8)
    assert t["hbm_frac"] == pytest.approx(4e8 / (ms * 1e-3) / bench.HBM_PEAK_BPS)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 2
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
