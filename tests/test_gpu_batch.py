"""Batched multi-temperature launch: the persistent kernels give the same per-sample results
as one launch per problem, the autotuned schedule keeps them, and the hipGraph block step
reproduces the eager sample -> evaluate -> reduce sequence on every replay."""
import numpy as np
import pytest
import torch

from helpers import make_data, oracle_setup, random_model, rel_err, stage_dict_models
from oracle import pimc_oracle as po
from pibronic_amd import _native, pimc, synthetic
from pibronic_amd.batch import TemperatureBatch

pytestmark = pytest.mark.gpu

TEMPS_P = [(150.0, 47), (250.0, 29), (400.0, 18), (600.0, 12)]


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _datas(FS, B, pm, temps=TEMPS_P):
    return [make_data(FS, T, P, B, B, pm=pm) for T, P in temps]


@pytest.mark.parametrize("pm", [False, True])
def test_batch_matches_single_launches(tmp_path, pm):
    """Lanes family, four temperatures with uneven P: the batched launch equals one
    pib_eval_block per temperature bit for bit (same lanes per sample), and the oracle to
    1e-10."""
    _announce(f"batch lanes A3 four temperatures pm={pm}")
    FS = stage_dict_models(tmp_path, random_model(3, 4, seed=2, quad="full"))
    B = 200
    datas = _datas(FS, B, pm)
    qs = [d.draw_host_sample(np.random.default_rng(10 + i), B)[0] for i, d in enumerate(datas)]
    tb = TemperatureBatch(datas, B)
    lanes = [d.context().lanes_per_sample(B) for d in datas]
    tb.set_schedule("lanes", lanes, tile_waves=3, grid_waves=512)
    res = tb.evaluate(qs)
    keys = ["s_rho", "s_g", "log_scale"] + (["s_gP", "s_gM"] if pm else [])
    for i, (d, q) in enumerate(zip(datas, qs)):
        single = pimc.evaluate_samples(d, q, pm=pm)
        for k in keys:
            assert np.array_equal(res[i][k], single[k]), (i, k)
        T, P = TEMPS_P[i]
        o = po.evaluate(oracle_setup(FS, T, P), q, pm=pm)
        for k in keys[:2] + keys[3:]:
            assert rel_err(res[i][k], o[k]) < 1e-10, (i, k)
    # the same problems as separate launches through the batch object
    seq = tb.evaluate(qs, sequential=True)
    for i in range(len(datas)):
        assert np.array_equal(seq[i]["s_g"], res[i]["s_g"])
    # and as the batch's own one-launch-per-problem family with the same lanes
    tb.set_schedule("sequential", lanes)
    fam = tb.evaluate(qs)
    for i in range(len(datas)):
        for k in keys:
            assert np.array_equal(fam[i][k], res[i][k]), (i, k)


def test_batch_wide_family_mixed_models(tmp_path):
    """Wide family: problems with different A, N and P in one launch, bit for bit equal to
    the wide kernel run per problem."""
    _announce("batch wide mixed models")
    B = 48
    datas, qs = [], []
    for k, (A, N, P) in enumerate([(3, 3, 9), (5, 2, 14), (10, 2, 6), (12, 3, 11)]):
        (tmp_path / f"m{k}").mkdir()
        FS = stage_dict_models(tmp_path / f"m{k}", random_model(A, N, seed=30 + k, quad="full"))
        d = make_data(FS, 300.0, P, B, B, pm=True)
        d.force_wide = True
        datas.append(d)
        qs.append(d.draw_host_sample(np.random.default_rng(k), B)[0])
    tb = TemperatureBatch(datas, B)
    assert tb.schedule()["family"] == "wide"
    tb.set_schedule("wide", tile_waves=5, grid_waves=300)
    res = tb.evaluate(qs)
    for i, (d, q) in enumerate(zip(datas, qs)):
        single = pimc.evaluate_samples(d, q, pm=True)
        for k in ["s_rho", "s_g", "s_gP", "s_gM", "log_scale"]:
            assert np.array_equal(res[i][k], single[k]), (i, k)


def test_batch_tune_and_graph_replays(tmp_path):
    """Autotune, then run blocks as hipGraph replays: every replay equals the eager
    sample_block -> eval_block -> block_sums sequence of the same block, also with eager
    work and synchronisations between the replays."""
    _announce("batch tune + graph replays")
    FS = stage_dict_models(tmp_path, synthetic.generate(4, 8, quadratic=True, seed=0))
    B = 4096
    datas = _datas(FS, B, True, temps=[(300.0, 64), (200.0, 96), (500.0, 40)])
    tb = TemperatureBatch(datas, B, max_blocks=4)
    t = tb.tune(seed=3)
    print("tune", t)
    assert t["batched_ms"] > 0 and t["sequential_ms"] > 0
    sched = tb.schedule()
    seed = 77
    tb.run(seed, first_block=0, nblocks=2)
    first = tb.block_sums()[:2].copy()
    # eager work between replays, then replay the same blocks and one more
    scratch = torch.randn(1 << 20, device="cuda").sum().item()
    assert np.isfinite(scratch)
    tb.run(seed, first_block=0, nblocks=3)
    again = tb.block_sums()
    assert np.array_equal(first, again[:2])
    # eager reference with the tuned lanes per sample
    for i, d in enumerate(datas):
        ctx = d.context()
        if sched["family"] in ("lanes", "sequential"):
            ctx.set_lanes_per_sample(sched["lanes"][i])
        else:
            d.force_wide = True
        q = torch.empty((B, d.modes, d.beads), dtype=torch.float64, device="cuda")
        out = torch.empty((_native.NUM_OUT, B), dtype=torch.float64, device="cuda")
        sums = torch.zeros(_native.NUM_SUMS, dtype=torch.float64, device="cuda")
        for k in range(3):
            ctx.sample_block(seed, k * B, B, q)
            d.evaluate_device(q, B, True, out, B)
            ctx.block_sums(out, B, B, True, d.delta_beta, sums)
            torch.cuda.synchronize()
            ref = sums.cpu().numpy()
            assert np.allclose(again[k, i], ref, rtol=1e-12, atol=0), (i, k)
        ctx.set_lanes_per_sample(0)
        d.force_wide = False
    assert np.all(again[:3, :, _native.SUM_N] == B)


def test_batch_errors(tmp_path):
    _announce("batch errors")
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    FS3 = stage_dict_models(tmp_path / "a", random_model(3, 2, seed=1))
    FS4 = stage_dict_models(tmp_path / "b", random_model(4, 2, seed=1))
    d3 = make_data(FS3, 300.0, 8, 16, 16)
    d4 = make_data(FS4, 300.0, 8, 16, 16)
    b = _native.Batch([d3.context(), d4.context()], 16)
    with pytest.raises(_native.NativeError, match="not bound"):
        b.eval()
    with pytest.raises(_native.NativeError, match="lanes family"):
        b.set_schedule(_native.BATCH_LANES, [4, 4], 4, 256)
    b.close()


def test_batch_refuses_contexts_set_again(tmp_path):
    """A setter on a context after pib_batch_create may reallocate the device buffers whose
    pointers the batch copied: every launching entry point must then fail with
    PIB_ERR_STATE, before any kernel is queued."""
    _announce("batch refuses re-set contexts")
    FS = stage_dict_models(tmp_path, random_model(3, 4, seed=2, quad="full"))
    B = 64
    datas = _datas(FS, B, False, TEMPS_P[:2])
    tb = TemperatureBatch(datas, B)
    qs = [d.draw_host_sample(np.random.default_rng(i), B)[0] for i, d in enumerate(datas)]
    tb.evaluate(qs)  # fine before
    d = datas[1]
    d.context().set_beads(d.beads, d.tau)
    for call in (lambda: tb.batch.eval(), lambda: tb.batch.eval_sequential(), lambda: tb.batch.tune(1),
                 lambda: tb.batch.run(1, 0, 1)):
        with pytest.raises(_native.NativeError, match="set again") as ei:
            call()
        assert ei.value.code == _native.PIB_ERR_STATE
