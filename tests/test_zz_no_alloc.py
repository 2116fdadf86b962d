"""pib_block_sums_noalloc (partial rows in the library's static per-GPU array) against
pib_block_sums / pib_block_sums_q (partial rows the context allocates), through the C-ABI:
identical sums, bit for bit, on two calls in a row, the second on a smaller block (so stale
rows of the first call would show), with and without the +-dbeta terms and a proposal
density."""
import numpy as np
import pytest
import torch

from helpers import make_data, stage_dict_models
from pibronic_amd import _native, synthetic

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("pm,use_q", [(False, False), (True, False), (False, True)])
def test_noalloc_sums_equal_allocating_sums(tmp_path, pm, use_q):
    print(f"[gpu-case] block_sums_noalloc pm={pm} logq={use_q}", flush=True)
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 3, quadratic=False, seed=0))
    B1, B2 = 300_000, 70_000          # 1024 partial rows (capped), then 274
    data = make_data(FS, 300.0, 12, B1, B1, pm=pm, seed=3)
    ctx = data.context()
    dev = torch.device("cuda", 0)
    q = torch.empty((B1, 3, 12), dtype=torch.float64, device=dev)
    ctx.sample_block(3, 0, B1, q)
    out = torch.empty((_native.NUM_OUT, B1), dtype=torch.float64, device=dev)
    data.evaluate_device(q, B1, pm, out, B1)
    logq = None
    if use_q:  # some density in rho's normalisation: rho itself, perturbed
        logq = (torch.log(out[_native.OUT_RHO]) + out[_native.OUT_LOGS]
                + 0.01 * torch.randn(B1, dtype=torch.float64, device=dev, generator=torch.Generator(dev).manual_seed(1)))
    db = data.delta_beta if pm else 1.0
    for B in (B1, B2):
        ref = torch.full((_native.NUM_SUMS,), float("nan"), dtype=torch.float64, device=dev)
        got = torch.full((_native.NUM_SUMS,), float("nan"), dtype=torch.float64, device=dev)
        if use_q:
            ctx.block_sums_q(out, B1, B, pm, db, logq, ref)
        else:
            ctx.block_sums(out, B1, B, pm, db, ref)
        ctx.block_sums_noalloc(out, B1, B, pm, db, got, logq=logq)
        torch.cuda.synchronize(dev)
        a, b = ref.cpu().numpy(), got.cpu().numpy()
        assert a[_native.SUM_N] == B and np.isfinite(a).all()
        assert np.array_equal(a, b), (B, a, b)


def test_noalloc_sums_rejects_bad_arguments(tmp_path):
    print("[gpu-case] block_sums_noalloc argument checks", flush=True)
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 3, quadratic=False, seed=0))
    data = make_data(FS, 300.0, 12, 64, 64, seed=3)
    ctx = data.context()
    dev = torch.device("cuda", 0)
    out = torch.zeros((_native.NUM_OUT, 64), dtype=torch.float64, device=dev)
    sums = torch.zeros((_native.NUM_SUMS,), dtype=torch.float64, device=dev)
    with pytest.raises(_native.NativeError, match="out_stride < B"):
        ctx.block_sums_noalloc(out, 64, 65, False, 1.0, sums)
    with pytest.raises(_native.NativeError, match="delta_beta"):
        ctx.block_sums_noalloc(out, 64, 64, True, 0.0, sums)
