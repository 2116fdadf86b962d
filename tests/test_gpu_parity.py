"""Parity of the HIP path against the oracle and the reference's golden vectors.

Every test runs the product entry points (pibronic_amd.pimc -> libpibronic_hip.so) on the
GPU and checks against the CPU oracle on the same host-seeded coordinates.

Tolerances (FP64 path): per-sample s_rho, s_g, s_gP, s_gM within 1e-10 relative of the
oracle. The remaining differences are the O-matrix exponent (evaluated in a
cancellation-free form, closer to the exact value than the reference's expression; the
difference is the reference's own rounding, pinned in test_exponent_form.py), exp()
implementations (1 ulp), M = exp(-tau V) (a scaled and squared Taylor polynomial against
LAPACK's eigendecomposition) and the association order of the bead products.
"""
import os

import numpy as np
import pytest

from conftest import EXPLICIT, TEST_MODELS, golden
from helpers import make_data, oracle_setup, random_model, rel_err, stage_dict_models, stage_models
from oracle import pimc_oracle as po
from oracle import stats_oracle as so
from pibronic_amd import _native, pimc, synthetic

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def explicit_fs(tmp_path_factory):
    root = tmp_path_factory.mktemp("explicit")
    return stage_models(root, os.path.join(EXPLICIT, "coupled_model.json"),
                        os.path.join(EXPLICIT, "sampling_model.json"))


def test_explicit_golden_vectors(explicit_fs):
    """O-matrices, M, g and rho of the reference's explicit example (A=2, N=2, P=5, 300 K)
    (tests/pimc/test_pimc_explicit_example.py:168-211), unscaled as that test does."""
    _announce("explicit golden A2 N2 P5")
    data = make_data(explicit_fs, 300.0, 5, 10, 2)
    q = golden("samples.npy")[:, 0].copy()
    r = pimc.evaluate_samples(data, q, pm=False, noscale=True, debug=True)
    Or, Ov, M = golden("rho_oMat.npy"), golden("vib_oMat.npy"), golden("vib_mMat.npy")
    dr = np.diagonal(Or, axis1=2, axis2=3)
    dv = np.diagonal(Ov, axis1=2, axis2=3)
    # the fixtures were written by an older NumPy; the oracle agrees with them to 8e-13
    assert rel_err(r["O_rho"], dr) < 1e-11
    assert rel_err(r["O_vib"], dv) < 1e-11
    assert np.max(np.abs(r["M"] - M)) < 1e-14
    assert rel_err(r["s_g"], golden("numerator(g).npy")) < 1e-12
    assert rel_err(r["s_rho"], golden("denominator(rho).npy")) < 1e-12
    # and against the oracle at the same inputs: the oracle's exponent (the reference's
    # expression) carries up to 9e-13 of rounding on these coordinates, the kernel's 2e-14
    # (test_exponent_form.py), so O agrees to ~5e-13
    s = oracle_setup(explicit_fs, 300.0, 5)
    o = po.evaluate(s, q, scale=False, keep=True)
    assert rel_err(r["O_rho"], np.diagonal(o["O_rho"], axis1=2, axis2=3)) < 1e-12
    assert rel_err(r["O_vib"], np.diagonal(o["O_vib"], axis1=2, axis2=3)) < 1e-12
    assert rel_err(r["s_g"], o["s_g"]) < 3e-12  # products of five such O-factors
    assert rel_err(r["s_rho"], o["s_rho"]) < 3e-12


def _parity(FS, T, P, B, pm, seed, lanes=None, tol=RTOL):
    data = make_data(FS, T, P, B, B, pm=pm)
    if lanes is not None:
        data.context().set_lanes_per_sample(lanes)
    q, _ = data.draw_host_sample(np.random.default_rng(seed), B)
    r = pimc.evaluate_samples(data, q, pm=pm)
    s = oracle_setup(FS, T, P)
    o = po.evaluate(s, q, pm=pm)
    keys = ["s_rho", "s_g"] + (["s_gP", "s_gM"] if pm else [])
    errs = {k: rel_err(r[k], o[k]) for k in keys}
    errs["log_scale"] = np.max(np.abs(r["log_scale"] - o["log_scale"]))
    for k in keys:
        assert errs[k] < tol, (k, errs)
    assert errs["log_scale"] < 1e-9 * max(1.0, np.max(np.abs(o["log_scale"]))), errs
    return errs, r, o


@pytest.mark.parametrize("pm", [False, True])
def test_parity_A2_N3_P12(tmp_path, pm):
    """Small workload: A=2, N=3, P=12, 300 K, linear couplings, host-seeded q."""
    _announce(f"parity A2 N3 P12 pm={pm}")
    cfg = synthetic.WORKLOADS["A2_N3_P12_T300"]
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 3, quadratic=False, seed=0))
    _parity(FS, cfg["T"], cfg["P"], 4000, pm, seed=11)


@pytest.mark.parametrize("pm", [False, True])
def test_parity_A4_N8_P64(tmp_path, pm):
    """Main workload shape: A=4, N=8, P=64, 300 K, linear + diagonal quadratic."""
    _announce(f"parity A4 N8 P64 pm={pm}")
    FS = stage_dict_models(tmp_path, synthetic.generate(4, 8, quadratic=True, seed=0))
    errs, r, o = _parity(FS, 300.0, 64, 600, pm, seed=12)
    print("A4N8P64 errors", errs)


@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32, 64])
def test_lane_split_invariance(tmp_path, lanes):
    """Every lanes-per-sample split gives the same per-sample results (ordered tree)."""
    _announce(f"lanes={lanes} A3 N4 P37")
    FS = stage_dict_models(tmp_path, random_model(3, 4, seed=5))
    _parity(FS, 300.0, 37, 300, True, seed=13, lanes=lanes)


@pytest.mark.parametrize("pm", [False, True])
def test_empty_and_single_sample_blocks(tmp_path, pm):
    """Edge sizes: a block of no samples is a no-op (the evaluation and the sampler return 0 and
    write nothing, with null coordinate pointers too); block sums over no samples are refused
    (B < 1); a block of one sample matches the oracle under several lanes-per-sample splits."""
    import torch
    _announce(f"empty and single-sample blocks pm={pm}")
    FS = stage_dict_models(tmp_path, synthetic.generate(4, 8, quadratic=True, seed=0))
    data = make_data(FS, 300.0, 64, 4, 4, pm=pm)
    ctx = data.context()
    out = torch.full((_native.NUM_OUT, 4), 7.0, dtype=torch.float64, device="cuda")
    ctx.eval_block(0, 0, _native.EVAL_PM if pm else 0, out, 0)
    ctx.sample_block(1, 0, 0, 0)
    torch.cuda.synchronize()
    assert bool((out == 7.0).all())
    sums = torch.zeros(_native.NUM_SUMS, dtype=torch.float64, device="cuda")
    with pytest.raises(_native.NativeError, match="B < 1"):
        ctx.block_sums(out, 4, 0, pm, 1e-4, sums)
    for lanes in (1, 8, 64):
        _parity(FS, 300.0, 64, 1, pm, seed=21, lanes=lanes)


@pytest.mark.parametrize("quad", ["diag", "full"])
def test_pm_passes_match_one_pass_bit_for_bit(tmp_path, quad):
    """+-dbeta at A = 8 runs as three single-chain passes (g, g+, g-); the debug path keeps the
    kernel that carries the three chains together. Both form the same M and scaling, so every
    output agrees bit for bit, and the passes match the oracle."""
    _announce(f"A=8 +-dbeta passes vs one pass, {quad} Q")
    FS = stage_dict_models(tmp_path, random_model(8, 4, seed=31, quad=quad))
    data = make_data(FS, 300.0, 16, 300, 300, pm=True)
    q, _ = data.draw_host_sample(np.random.default_rng(32), 300)
    r3 = pimc.evaluate_samples(data, q, pm=True)
    r1 = pimc.evaluate_samples(data, q, pm=True, debug=True)
    for k in ("s_rho", "s_g", "s_gP", "s_gM", "log_scale"):
        assert np.array_equal(r3[k], r1[k]), k
    o = po.evaluate(oracle_setup(FS, 300.0, 16), q, pm=True)
    for k in ("s_rho", "s_g", "s_gP", "s_gM"):
        assert rel_err(r3[k], o[k]) < RTOL, k


@pytest.mark.parametrize("A", [1, 2, 3, 5, 6, 7, 8])
def test_surface_counts(tmp_path, A):
    """Every compiled surface count, odd bead count, full (j != k) quadratic couplings."""
    _announce(f"surfaces A={A} N=3 P=9 full Q")
    FS = stage_dict_models(tmp_path, random_model(A, 3, seed=A, quad="full"))
    _parity(FS, 300.0, 9, 200, True, seed=14)


@pytest.mark.parametrize("A", [1, 3, 5, 6, 7, 8])
def test_surface_counts_plain(tmp_path, A):
    """The plain kernels (no +-dbeta) at every compiled surface count: the four-product degree-15+
    exponential (the degree-8 form at A = 6), odd bead count, full (j != k) quadratic couplings."""
    _announce(f"surfaces A={A} N=3 P=9 full Q, plain")
    FS = stage_dict_models(tmp_path, random_model(A, 3, seed=A, quad="full"))
    _parity(FS, 300.0, 9, 200, False, seed=14)


@pytest.mark.parametrize("A,pm", [(3, False), (5, False), (8, False), (5, True)])
def test_parity_many_squarings(tmp_path, A, pm):
    """A cold, coarse ring (T = 10 K, P = 4, tau = 290 / eV) so that the exponentials square four
    to five times in the plain kernels' degree-15+ form (seven to eight in the +-dbeta kernels'
    degree-8 form), with g up to 1e124."""
    _announce(f"many squarings A={A} pm={pm}")
    FS = stage_dict_models(tmp_path, random_model(A, 3, seed=20 + A, quad="full"))
    _parity(FS, 10.0, 4, 200, pm, seed=16)


@pytest.mark.parametrize("A,AR", [(2, 1), (4, 2), (2, 4), (4, 8)])
def test_sampling_model_with_other_surface_count(tmp_path, A, AR):
    """A_rho != A: the sampling model's O-matrices use a separate coordinate copy
    (pibronic/pimc/pimc.py:1094-1096)."""
    _announce(f"A={A} A_rho={AR}")
    m = random_model(A, 3, seed=7)
    rng = np.random.default_rng(AR)
    from pibronic_amd.vibronic_model_keys import VibronicModelKeys as VMK
    rho = {VMK.N: 3, VMK.A: AR, VMK.w: m[VMK.w].copy(), VMK.E: rng.uniform(0, 0.3, AR),
           VMK.G1: rng.normal(0, 0.05, (3, AR))}
    FS = stage_dict_models(tmp_path, m, rho)
    _parity(FS, 300.0, 10, 300, True, seed=15)


def test_noscale_matches_scaled_ratio(tmp_path):
    """S12 scaling changes s_g and s_rho by the same factor prod_p S_p."""
    _announce("noscale vs scaled A4 N8 P16")
    FS = stage_dict_models(tmp_path, synthetic.generate(4, 8, seed=3))
    data = make_data(FS, 300.0, 16, 500, 500, pm=True)
    q, _ = data.draw_host_sample(np.random.default_rng(3), 500)
    a = pimc.evaluate_samples(data, q, pm=True)
    b = pimc.evaluate_samples(data, q, pm=True, noscale=True)
    assert rel_err(a["s_g"] / a["s_rho"], b["s_g"] / b["s_rho"]) < 1e-12
    assert rel_err(a["s_g"] * np.exp(a["log_scale"]), b["s_g"]) < 1e-11
    assert np.all(b["log_scale"] == 0.0)


def test_diagonal_model_gives_g_equal_rho(tmp_path, test_models_dir=TEST_MODELS):
    """Uncoupled model sampled by its own diagonal: V = 0, M = I, so g == rho per sample
    (data_set_0 / rho_0 of the reference's test models)."""
    _announce("g == rho special case")
    base = os.path.join(test_models_dir, "data_set_0")
    FS = stage_models(tmp_path, os.path.join(base, "parameters", "coupled_model.json"),
                      os.path.join(base, "rho_0", "parameters", "sampling_model.json"))
    data = make_data(FS, 300.0, 12, 2000, 2000, pm=False)
    q, _ = data.draw_host_sample(np.random.default_rng(4), 2000)
    r = pimc.evaluate_samples(data, q, pm=False)
    assert rel_err(r["s_g"], r["s_rho"]) < 1e-13


def test_block_sums_match_host(tmp_path):
    """pib_block_sums against numpy sums of the same per-sample outputs."""
    _announce("block sums")
    import torch
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 3, quadratic=False, seed=0))
    B = 20000
    data = make_data(FS, 300.0, 12, B, B, pm=True)
    res = pimc.block_compute_pm(data, pimc.BoxResultPM(data=data))
    s = res.block_sums[0]
    rho, g, gp, gm = res.scaled_rho, res.scaled_g, res.scaled_gofr_plus, res.scaled_gofr_minus
    r, d1, d2 = so.basic_terms(data.delta_beta, rho, g, gp, gm)
    ref = {_native.SUM_N: B, _native.SUM_R: r.sum(), _native.SUM_R2: (r * r).sum(),
           _native.SUM_ABSR: np.abs(r).sum(), _native.SUM_NEG: (r < 0).sum(),
           _native.SUM_D1: d1.sum(), _native.SUM_D1SQ: (d1 * d1).sum(),
           _native.SUM_D2: d2.sum(), _native.SUM_D2SQ: (d2 * d2).sum(),
           _native.SUM_RP: (gp / rho).sum(), _native.SUM_RM: (gm / rho).sum()}
    for k, v in ref.items():
        assert abs(s[k] - v) <= 1e-10 * max(1.0, abs(v)), (k, s[k], v)


@pytest.mark.parametrize("P", [5, 64])
def test_device_sampler_distribution(tmp_path, P):
    """Device Philox/Box-Muller sampler against the closed-form mixture moments:
    surface frequencies ~ w_a, mean sum_a w_a d_a, per-mode covariance
    sum_a w_a (C + d_a d_a^T) - mu mu^T with C = (2 coth I - csch C_ring)^-1
    (tests/pimc/explicit_data/sampling_covariance.npy holds the same C for the
    explicit model at P = 5). At P = 64 the draws are also whitened back,
    y = (R - d_src) U / sigma, and the y (i.i.d. N(0, 1) for an exact sampler) are checked for
    mean, variance, fourth moment, tail frequencies beyond 3 and 4 sigma, and the largest |y|."""
    _announce(f"device sampler moments P={P}")
    import torch
    FS = stage_models(tmp_path, os.path.join(EXPLICIT, "coupled_model.json"),
                      os.path.join(EXPLICIT, "sampling_model.json"))
    B = 400_000
    data = make_data(FS, 300.0, P, B, B, seed=99)
    data.draw_sample(slice(0, B))
    torch.cuda.synchronize()
    dev = data.device_buffers(B)
    q = dev["q"][:B].cpu().numpy()
    src = dev["src"][:B].cpu().numpy()
    w = data.rho.state_weight
    freq = np.bincount(src, minlength=len(w)) / B
    assert np.all(np.abs(freq - w) < 5 * np.sqrt(w * (1 - w) / B) + 1e-12), (freq, w)
    coth, csch = data.rho.const.cothAN[0], data.rho.const.cschAN[0]
    d = data.rho.state_shift
    for j in range(data.modes):
        C = np.linalg.inv(2 * coth[j] * np.eye(P) - csch[j] * data.circulant_matrix)
        if P == 5:
            assert np.allclose(C, golden("sampling_covariance.npy")[0, j], rtol=1e-10, atol=1e-12)
        mu = np.sum(w * d[:, j]) * np.ones(P)
        cov = sum(w[a] * (C + np.outer(d[a, j] * np.ones(P), d[a, j] * np.ones(P)))
                  for a in range(len(w))) - np.outer(mu, mu)
        x = q[:, j, :]
        m_hat = x.mean(axis=0)
        c_hat = np.cov(x, rowvar=False)
        se_m = np.sqrt(np.diag(cov) / B)
        assert np.all(np.abs(m_hat - mu) < 6 * se_m), (j, m_hat, mu)
        se_c = np.sqrt((np.outer(np.diag(cov), np.diag(cov)) + cov ** 2) / B)
        assert np.all(np.abs(c_hat - cov) < 6 * se_c + 1e-12), (j, c_hat, cov)
    # the surface-conditional mean: samples from source a are centred on d_a
    for a in range(len(w)):
        sel = src == a
        if sel.sum() > 1000:
            assert np.allclose(q[sel].mean(axis=(0, 2)), d[a], atol=6 * np.sqrt(np.max(np.diag(C)) / sel.sum()))
    if P == 64:
        # whitened draws: i.i.d. standard normals for an exact sampler
        y = (q - d[src][:, :, None]) @ data.circulant_eigvects / data.rho.standard_deviation[0][None]
        y = y.ravel()
        n = y.size
        m1, m2, m4 = y.mean(), (y * y).mean(), (y ** 4).mean()
        print(f"whitened: n={n} mean {m1:.2e} var {m2:.6f} m4 {m4:.5f} max|y| {np.abs(y).max():.3f}")
        assert abs(m1) < 6 / np.sqrt(n)
        assert abs(m2 - 1) < 6 * np.sqrt(2 / n)
        assert abs(m4 - 3) < 6 * np.sqrt(96 / n)
        for t, p_tail in ((3.0, 2.699796063260e-3), (4.0, 6.334248366624e-5)):
            k = int(np.sum(np.abs(y) > t))
            assert abs(k - n * p_tail) < 6 * np.sqrt(n * p_tail), (t, k, n * p_tail)
        # the largest of 5e7 standard normals is ~6 (a 32-bit radius uniform alone stops at 6.7)
        assert 5.0 < np.abs(y).max() < 7.5


def test_counter_based_sampler_is_split_invariant(tmp_path):
    """Sample i of a run is the same whether drawn in one block or in several."""
    _announce("sampler split invariance")
    import torch
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 3, quadratic=False, seed=0))
    data = make_data(FS, 300.0, 12, 3000, 3000, seed=5)
    ctx = data.context()
    dev = torch.device("cuda", 0)
    whole = torch.empty((3000, 3, 12), dtype=torch.float64, device=dev)
    parts = torch.empty((3000, 3, 12), dtype=torch.float64, device=dev)
    ctx.sample_block(5, 0, 3000, whole)
    ctx.sample_block(5, 0, 1000, parts[:1000])
    ctx.sample_block(5, 1000, 2000, parts[1000:])
    torch.cuda.synchronize()
    assert torch.equal(whole, parts)


def test_statistics_match_oracle_sampling(tmp_path):
    """Independent runs agree statistically: device-sampled GPU run vs host-sampled oracle
    run of the same model; Z_MC and the jackknife E within 2 combined standard errors
    (measured with these seeds: 0.08 and 0.8 sigma)."""
    _announce("statistical agreement A2 N3 P12")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 3, quadratic=False, seed=0))
    X = 200_000
    data = make_data(FS, 300.0, 12, X, 50_000, pm=True, seed=21)
    res = pimc.block_compute_pm(data, pimc.BoxResultPM(data=data))
    s = oracle_setup(FS, 300.0, 12)
    Xo = 20_000
    q, _ = s.draw_q(np.random.default_rng(77), Xo)
    o = po.evaluate(s, q, pm=True)

    def est(rho, g, gp, gm):
        r, d1, d2 = so.basic_terms(po.DELTA_BETA, rho, g, gp, gm)
        n = len(r)
        Z, Zerr = r.mean(), r.std() / np.sqrt(n - 1)
        jk = [so.jackknife_term(n, t) for t in (r, d1, d2)]
        basic = so.estimate_basic(n, 300.0, r, d1, d2)
        j = so.estimate_jackknife(n, 300.0, po.DELTA_BETA, basic, *jk)
        return Z, Zerr, j["E"], j["E error"]

    Zg, eg, Eg, sEg = est(res.scaled_rho, res.scaled_g, res.scaled_gofr_plus, res.scaled_gofr_minus)
    Zo, eo, Eo, sEo = est(o["s_rho"], o["s_g"], o["s_gP"], o["s_gM"])
    print(f"Z gpu {Zg:.6f}+-{eg:.2e} oracle {Zo:.6f}+-{eo:.2e}; E gpu {Eg:.6f}+-{sEg:.2e} oracle {Eo:.6f}+-{sEo:.2e}")
    assert abs(Zg - Zo) < 2 * np.hypot(eg, eo)
    assert abs(Eg - Eo) < 2 * np.hypot(sEg, sEo)


def test_estimates_agree_within_jackknife_error(tmp_path):
    """Z(beta)/Z_rho and the jackknife <E> of block_compute_pm on host-seeded samples agree
    with the oracle's on the same samples far inside 2 sigma of the jackknife error (the
    criterion the estimates are judged by); the difference is roundoff only."""
    _announce("estimates within 2 sigma A4 N8 P64")
    FS = stage_dict_models(tmp_path, synthetic.generate(4, 8, quadratic=True, seed=0))
    X, B = 8000, 2000
    data = make_data(FS, 300.0, 64, X, B, pm=True, seed=31)
    data.sampler = "host"
    res = pimc.block_compute_pm(data, pimc.BoxResultPM(data=data))
    s = oracle_setup(FS, 300.0, 64)
    # the oracle on exactly the samples the host sampler produced, block by block
    o = {"s_rho": [], "s_g": [], "s_gP": [], "s_gM": []}
    for k in range(X // B):
        qk, _ = data.draw_host_sample(np.random.default_rng([31, k * B]), B)  # draw_sample's host stream
        ok = po.evaluate(s, qk, pm=True)
        for key in o:
            o[key].append(ok[key])
    o = {k: np.concatenate(v) for k, v in o.items()}

    def est(rho, g, gp, gm):
        r, d1, d2 = so.basic_terms(po.DELTA_BETA, rho, g, gp, gm)
        n = len(r)
        jk = [so.jackknife_term(n, t) for t in (r, d1, d2)]
        basic = so.estimate_basic(n, 300.0, r, d1, d2)
        j = so.estimate_jackknife(n, 300.0, po.DELTA_BETA, basic, *jk)
        return r.mean(), r.std() / np.sqrt(n - 1), j["E"], j["E error"]

    Zg, eg, Eg, sEg = est(res.scaled_rho, res.scaled_g, res.scaled_gofr_plus, res.scaled_gofr_minus)
    Zo, eo, Eo, sEo = est(o["s_rho"], o["s_g"], o["s_gP"], o["s_gM"])
    print(f"Z gpu {Zg:.10e}+-{eg:.2e} oracle {Zo:.10e}; E gpu {Eg:.10e}+-{sEg:.2e} oracle {Eo:.10e}")
    assert abs(Zg - Zo) < 2 * eo and abs(Eg - Eo) < 2 * sEo
    assert abs(Zg - Zo) <= 1e-9 * abs(Zo) and abs(Eg - Eo) <= 1e-8 * max(abs(Eo), sEo)


def test_errors_are_reported(tmp_path):
    """Invalid calls return an error through the ABI instead of faulting."""
    _announce("error paths")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 3, quadratic=False, seed=0))
    data = make_data(FS, 300.0, 12, 10, 10)
    ctx = data.context()
    with pytest.raises(_native.NativeError, match="out_stride"):
        ctx.eval_block(0, 10, 0, 0, 5)
    with pytest.raises(_native.NativeError, match="3 or more beads"):
        ctx.set_beads(2, 0.1)
    with pytest.raises(_native.NativeError, match="unknown flag"):
        ctx.eval_block(0, 10, 32, 0, 10)
