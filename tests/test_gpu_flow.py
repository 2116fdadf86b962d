"""Learned-proposal kernels on the GPU against the float64 torch reference (FP32 bounds).

GEMM (all transposes, bias, ReLU, mask, accumulate), whitening, log rho, the flow in both
directions, and the training gradient; then a short training run that must lower the
weight variance relative to the harmonic rho on the same model.
"""
import numpy as np
import pytest
import torch

from flow_helpers import random_params, ref_for
from helpers import make_data, stage_dict_models
from pibronic_amd import _native, flow, synthetic

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_f32(tmp_path, ta, tb):
    _announce(f"gemm ta={ta} tb={tb}")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 2, quadratic=False, seed=0))
    ctx = make_data(FS, 300.0, 64, 10, 10).context()
    M, N, K = 128, 192, 48
    g = torch.Generator().manual_seed(ta * 2 + tb)
    A = torch.randn((K, M) if ta else (M, K), generator=g, dtype=torch.float32)
    B = torch.randn((N, K) if tb else (K, N), generator=g, dtype=torch.float32)
    bias = torch.randn(N, generator=g)
    mask = (torch.rand((M, N), generator=g) > 0.5).float()
    C0 = torch.randn((M, N), generator=g)
    ref = (A.double().T if ta else A.double()) @ (B.double().T if tb else B.double())
    Ad, Bd, bd, md = A.to(DEV), B.to(DEV), bias.to(DEV), mask.to(DEV)
    lda, ldb = A.shape[1], B.shape[1]
    C = torch.empty((M, N), device=DEV)
    ctx.gemm_f32(Ad, Bd, C, M, N, K, lda, ldb, N, trans_a=ta, trans_b=tb)
    torch.cuda.synchronize()
    assert torch.allclose(C.cpu().double(), ref, rtol=1e-5, atol=1e-4)
    ctx.gemm_f32(Ad, Bd, C, M, N, K, lda, ldb, N, trans_a=ta, trans_b=tb, bias=bd, relu=True)
    assert torch.allclose(C.cpu().double(), torch.relu(ref + bias.double()), rtol=1e-5, atol=1e-4)
    ctx.gemm_f32(Ad, Bd, C, M, N, K, lda, ldb, N, trans_a=ta, trans_b=tb, mask=md, ldmask=N)
    assert torch.allclose(C.cpu().double(), ref * mask.double(), rtol=1e-5, atol=1e-4)
    Cacc = C0.to(DEV).clone()
    ctx.gemm_f32(Ad, Bd, Cacc, M, N, K, lda, ldb, N, trans_a=ta, trans_b=tb, accumulate=True)
    assert torch.allclose(Cacc.cpu().double(), ref + C0.double(), rtol=1e-5, atol=1e-4)
    with pytest.raises(_native.NativeError, match="multiples"):
        ctx.gemm_f32(Ad, Bd, C, 100, N, K, lda, ldb, N, trans_a=ta, trans_b=tb)


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_gemm_f32_split_k(tmp_path, ta, tb):
    """Long-K products (the weight gradients, K = batch) run split over K with an ordered
    second pass; same epilogue options."""
    _announce(f"gemm split-K ta={ta} tb={tb}")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 2, quadratic=False, seed=0))
    ctx = make_data(FS, 300.0, 64, 10, 10).context()
    M, N, K = 128, 128, 2048
    g = torch.Generator().manual_seed(7 + ta * 2 + tb)
    A = torch.randn((K, M) if ta else (M, K), generator=g, dtype=torch.float32)
    B = torch.randn((N, K) if tb else (K, N), generator=g, dtype=torch.float32)
    bias = torch.randn(N, generator=g)
    mask = (torch.rand((M, N), generator=g) > 0.5).float()
    C0 = torch.randn((M, N), generator=g)
    ref = (A.double().T if ta else A.double()) @ (B.double().T if tb else B.double())
    Ad, Bd = A.to(DEV), B.to(DEV)
    lda, ldb = A.shape[1], B.shape[1]
    C = C0.to(DEV).clone()
    ctx.gemm_f32(Ad, Bd, C, M, N, K, lda, ldb, N, trans_a=ta, trans_b=tb, bias=bias.to(DEV),
                 relu=True, mask=mask.to(DEV), ldmask=N, accumulate=True)
    torch.cuda.synchronize()
    want = torch.relu(ref + bias.double()) * mask.double() + C0.double()
    assert torch.allclose(C.cpu().double(), want, rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("ta,tb,K", [(0, 0, 64), (0, 1, 96), (1, 0, 2048), (1, 1, 160)])
def test_gemm_bf16x3(tmp_path, ta, tb, K):
    """BF16x3 split on the bf16 matrix cores: elementwise |C - AB| <= 1e-4 (|A||B|), the
    standard GEMM error shape with the split's ~2^-16 per-term error; all epilogue options;
    K = 2048 runs split over K."""
    _announce(f"gemm bf16x3 ta={ta} tb={tb} K={K}")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 2, quadratic=False, seed=0))
    ctx = make_data(FS, 300.0, 64, 10, 10).context()
    M, N = 128, 192
    g = torch.Generator().manual_seed(11 + ta * 2 + tb)
    A = torch.randn((K, M) if ta else (M, K), generator=g, dtype=torch.float32)
    B = torch.randn((N, K) if tb else (K, N), generator=g, dtype=torch.float32)
    bias = torch.randn(N, generator=g)
    mask = (torch.rand((M, N), generator=g) > 0.5).float()
    C0 = torch.randn((M, N), generator=g)
    Ad = (A.double().T if ta else A.double())
    Bd = (B.double().T if tb else B.double())
    ref, bound = Ad @ Bd, 1e-4 * (Ad.abs() @ Bd.abs())
    lda, ldb = A.shape[1], B.shape[1]
    C = torch.empty((M, N), device=DEV)
    ctx.gemm_bf16x3(A.to(DEV), B.to(DEV), C, M, N, K, lda, ldb, N, trans_a=ta, trans_b=tb)
    torch.cuda.synchronize()
    err = (C.cpu().double() - ref).abs()
    print("max err / bound", float((err / bound).max()))
    assert torch.all(err <= bound)
    C = C0.to(DEV).clone()
    ctx.gemm_bf16x3(A.to(DEV), B.to(DEV), C, M, N, K, lda, ldb, N, trans_a=ta, trans_b=tb,
                    bias=bias.to(DEV), relu=True, mask=mask.to(DEV), ldmask=N, accumulate=True)
    want = torch.relu(ref + bias.double()) * mask.double() + C0.double()
    assert torch.all((C.cpu().double() - want).abs() <= bound + 1e-5 * want.abs() + 1e-6)
    with pytest.raises(_native.NativeError, match="K of 32"):
        ctx.gemm_bf16x3(A.to(DEV), B.to(DEV), C, M, N, 48, lda, ldb, N, trans_a=ta, trans_b=tb)


def test_gemm_bf16x3_unaligned_operand(tmp_path):
    """A view that is not 16-byte aligned takes the scalar-staging kernel: same results."""
    _announce("gemm bf16x3 unaligned")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 2, quadratic=False, seed=0))
    ctx = make_data(FS, 300.0, 64, 10, 10).context()
    M, N, K = 64, 128, 64
    g = torch.Generator().manual_seed(5)
    Afull = torch.randn((M, K + 1), generator=g).to(DEV)
    A = Afull[:, 1:]                       # 4-byte offset, lda = K + 1
    B = torch.randn((K, N), generator=g).to(DEV)
    C1 = torch.empty((M, N), device=DEV)
    C2 = torch.empty((M, N), device=DEV)
    ctx.gemm_bf16x3(A, B, C1, M, N, K, K + 1, N, N)
    ctx.gemm_bf16x3(A.contiguous(), B, C2, M, N, K, K, N, N)
    torch.cuda.synchronize()
    assert torch.equal(C1, C2)


def test_split_perm32_layout(tmp_path):
    """PIB_SPLIT_PERM32: input 32b + w of every output row sits at 32b + 8((w%16)/4) +
    4(w/16) + w%4, values identical to the plain transposed split; bad modes are refused."""
    _announce("split perm32")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 2, quadratic=False, seed=0))
    ctx = make_data(FS, 300.0, 64, 10, 10).context()
    K, N = 96, 48
    W = torch.randn((K, N), generator=torch.Generator().manual_seed(3)).to(DEV)
    th, tl = torch.empty((N, K), dtype=torch.int16, device=DEV), torch.empty((N, K), dtype=torch.int16, device=DEV)
    ph, pl = torch.empty_like(th), torch.empty_like(tl)
    ctx.split_bf16(K, N, W, N, th, tl, K, transpose=True)
    ctx.split_bf16(K, N, W, N, ph, pl, K, transpose=True, perm32=True)
    torch.cuda.synchronize()
    k = np.arange(K)
    w = k % 32
    pos = (k - w) + 8 * ((w % 16) // 4) + 4 * (w // 16) + w % 4
    assert sorted(pos.tolist()) == list(range(K))
    assert torch.equal(ph[:, torch.as_tensor(pos, device=DEV)], th)
    assert torch.equal(pl[:, torch.as_tensor(pos, device=DEV)], tl)
    with pytest.raises(_native.NativeError, match="PERM32"):
        ctx.split_bf16(K, N, W, N, ph, pl, K, transpose=False, perm32=True)
    with pytest.raises(_native.NativeError, match="PERM32"):
        ctx.split_bf16(40, N, W, N, ph, pl, K, transpose=True, perm32=True)


def test_gemm_presplit_matches_converting(tmp_path):
    """Pre-split operands give bit-for-bit the converting kernel's result (same splits, same
    product order), and the split output equals splitting the FP32 output."""
    _announce("gemm presplit")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 2, quadratic=False, seed=0))
    ctx = make_data(FS, 300.0, 64, 10, 10).context()
    M, N, K = 128, 192, 96
    g = torch.Generator().manual_seed(21)
    A = torch.randn((M, K), generator=g).to(DEV)
    W = torch.randn((K, N), generator=g).to(DEV)
    bias = torch.randn(N, generator=g).to(DEV)
    ref = torch.empty((M, N), device=DEV)
    ctx.gemm_bf16x3(A, W, ref, M, N, K, K, N, N, bias=bias, relu=True)
    ah, al = torch.empty((M, K), dtype=torch.int16, device=DEV), torch.empty((M, K), dtype=torch.int16, device=DEV)
    wh, wl = torch.empty((N, K), dtype=torch.int16, device=DEV), torch.empty((N, K), dtype=torch.int16, device=DEV)
    ctx.split_bf16(M, K, A, K, ah, al, K)
    ctx.split_bf16(K, N, W, N, wh, wl, K, transpose=True)
    C = torch.empty((M, N), device=DEV)
    oh, ol = torch.empty((M, N), dtype=torch.int16, device=DEV), torch.empty((M, N), dtype=torch.int16, device=DEV)
    ctx.gemm_bf16x3_presplit(M, N, K, ah, al, K, wh, wl, K, C=C, ldc=N, O_hi=oh, O_lo=ol, ldo=N,
                             bias=bias, relu=True)
    rh, rl = torch.empty_like(oh), torch.empty_like(ol)
    ctx.split_bf16(M, N, ref, N, rh, rl, N)
    torch.cuda.synchronize()
    assert torch.equal(C, ref)
    assert torch.equal(oh, rh) and torch.equal(ol, rl)
    with pytest.raises(_native.NativeError, match="multiples"):
        ctx.gemm_bf16x3_presplit(M, N, 48, ah, al, K, wh, wl, K, C=C, ldc=N)


def test_colsum_f32(tmp_path):
    """Two-pass column sums, ragged shapes, strided rows."""
    _announce("colsum")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 2, quadratic=False, seed=0))
    ctx = make_data(FS, 300.0, 64, 10, 10).context()
    M, N, ld = 2085, 200, 211
    X = torch.randn((M, ld), generator=torch.Generator().manual_seed(3), dtype=torch.float32)
    out = torch.empty(N, device=DEV)
    ctx.colsum_f32(M, N, X.to(DEV), ld, out)
    torch.cuda.synchronize()
    assert torch.allclose(out.cpu().double(), X[:, :N].double().sum(dim=0), rtol=1e-6, atol=1e-4)


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    FS = stage_dict_models(tmp_path_factory.mktemp("flow"), synthetic.generate(3, 2, quadratic=True, seed=1))
    data = make_data(FS, 300.0, 64, 1024, 1024, seed=3)
    return data


def _flow_and_ref(data, layers=4, hidden=64, seed=7, scale=0.05, gemm="f32"):
    fl = flow.Flow(data, layers=layers, hidden=hidden, smax=2.0, gemm=gemm)
    ref = ref_for(data, layers, hidden, 2.0)
    flat = random_params(ref, seed, scale)
    fl.set_params(flat)
    return fl, ref, flat


def _host_q(data, B, seed):
    q, _ = data.draw_host_sample(np.random.default_rng(seed), B)
    return q


def test_whitening_and_logrho(model):
    _announce("flow whitening + log rho")
    data = model
    fl, ref, _ = _flow_and_ref(data)
    B = 128
    q = _host_q(data, B, 1)
    R = torch.as_tensor(q, device=DEV)
    zp = torch.empty((B, fl.D), device=DEV)
    work = torch.empty(2 * B * fl.D, device=DEV)
    fl.ctx.flow_whiten(_native.FLOW_WHITEN, B, R, work, zp)
    zref = ref.whiten(torch.as_tensor(q))
    assert torch.allclose(zp.cpu().double(), zref, rtol=1e-4, atol=1e-4)
    back = torch.empty_like(R)
    fl.ctx.flow_whiten(_native.FLOW_UNWHITEN, B, zp, work, back)
    assert torch.allclose(back.cpu(), torch.as_tensor(q), atol=1e-4)
    lr = torch.empty(B, dtype=torch.float64, device=DEV)
    fl.ctx.flow_logrho(B, R, logrho=lr)
    torch.cuda.synchronize()
    assert torch.allclose(lr.cpu(), ref.log_rho(torch.as_tensor(q)), rtol=0, atol=1e-10)
    # the same rho as the FP64 evaluation kernel (scaled rho times the removed scale)
    r = __import__("pibronic_amd.pimc", fromlist=["x"]).evaluate_samples(data, q, pm=False)
    assert np.allclose(lr.cpu().numpy(), np.log(r["s_rho"]) + r["log_scale"], atol=1e-9)


@pytest.mark.parametrize("P", [128, 192])
def test_whitening_other_bead_counts(tmp_path, P):
    """P = 128 takes the fused whitening kernel, P = 192 the pre/product/post sequence; both
    against the float64 reference, and the inverse returns R."""
    _announce(f"flow whitening P={P}")
    FS = stage_dict_models(tmp_path, synthetic.generate(3, 2, quadratic=True, seed=1))
    data = make_data(FS, 300.0, P, 64, 64, seed=3)
    fl, ref, _ = _flow_and_ref(data)
    B = 96
    q = _host_q(data, B, 1)
    R = torch.as_tensor(q, device=DEV)
    zp = torch.empty((B, fl.D), device=DEV)
    work = torch.empty(2 * B * fl.D, device=DEV)
    fl.ctx.flow_whiten(_native.FLOW_WHITEN, B, R, work, zp)
    back = torch.empty_like(R)
    fl.ctx.flow_whiten(_native.FLOW_UNWHITEN, B, zp, work, back)
    torch.cuda.synchronize()
    assert torch.allclose(zp.cpu().double(), ref.whiten(torch.as_tensor(q)), rtol=1e-4, atol=1e-4)
    assert torch.allclose(back.cpu(), torch.as_tensor(q), atol=1e-4)


@pytest.mark.parametrize("gemm", ["f32", "bf16x3"])
def test_flow_forward_and_density(model, gemm):
    _announce(f"flow forward + log q gemm={gemm}")
    data = model
    fl, ref, flat = _flow_and_ref(data, gemm=gemm)
    B = 256
    q = _host_q(data, B, 2)
    R0 = torch.as_tensor(q, device=DEV)
    R, ld = fl.forward(R0)
    Rr, ldr = ref.forward(torch.as_tensor(q), flat)
    torch.cuda.synchronize()
    tf = 1.0 if gemm == "f32" else 10.0  # bf16x3: ~2^-16 per product term
    assert torch.allclose(R.cpu(), Rr, rtol=1e-4 * tf, atol=2e-4 * tf)
    assert torch.allclose(ld.cpu(), ldr, rtol=1e-4 * tf, atol=2e-3 * tf)
    lq, R0b = fl.log_q(R)
    lqr, _ = ref.log_q(R.cpu(), flat)
    assert torch.allclose(R0b.cpu(), torch.as_tensor(q), atol=5e-4 * tf)
    assert torch.allclose(lq.cpu(), lqr, rtol=1e-5 * tf, atol=5e-3 * tf)


@pytest.mark.parametrize("gemm", ["f32", "bf16x3"])
def test_training_gradient(model, gemm):
    _announce(f"flow gradient gemm={gemm}")
    data = model
    fl, ref, flat = _flow_and_ref(data, gemm=gemm)
    B = 256
    q = _host_q(data, B, 3)
    c = np.random.default_rng(4).uniform(0.5, 1.5, B)
    c /= c.sum()
    loss = fl.loss_and_grad(torch.as_tensor(q, device=DEV), torch.as_tensor(c, device=DEV))
    torch.cuda.synchronize()
    lref, gref = ref.loss_and_grad(q, c, flat)
    g = fl.grad.cpu().numpy().astype(np.float64)
    assert float(loss) == pytest.approx(lref, rel=1e-5, abs=1e-3)
    rel = np.linalg.norm(g - gref) / np.linalg.norm(gref)
    print(f"gradient relative error {rel:.2e}")
    assert rel < (2e-3 if gemm == "f32" else 1e-2)
    # every parameter block gets a gradient
    for l in range(fl.layers):
        for name, _ in fl.shapes:
            assert torch.count_nonzero(fl.p(l, name, fl.grad)) > 0, (l, name)


def test_adam_step_matches_formula(model):
    _announce("adam")
    data = model
    fl, ref, flat = _flow_and_ref(data, layers=2, hidden=64)
    rng = np.random.default_rng(0)
    g = rng.normal(size=fl.n_params).astype(np.float32)
    fl.grad.copy_(torch.as_tensor(g))
    p0 = fl.params.cpu().numpy().copy()
    fl.adam_step(1e-2)
    fl.grad.copy_(torch.as_tensor(2 * g))
    fl.adam_step(1e-2)
    torch.cuda.synchronize()
    m = 0.1 * g
    v = 0.001 * g * g
    p1 = p0 - 1e-2 * (m / 0.1) / (np.sqrt(v / 0.001) + 1e-8)
    m = 0.9 * m + 0.1 * 2 * g
    v = 0.999 * v + 0.001 * 4 * g * g
    p2 = p1 - 1e-2 * (m / (1 - 0.81)) / (np.sqrt(v / (1 - 0.999 ** 2)) + 1e-8)
    assert np.allclose(fl.params.cpu().numpy(), p2, rtol=1e-5, atol=1e-6)


def test_flow_fast_path_matches(model):
    """The draw's pre-split path equals the layer-by-layer bf16x3 path bit for bit, and
    recomputes its weight split after a parameter update."""
    _announce("flow fast path")
    data = model
    fl, ref, flat = _flow_and_ref(data, gemm="bf16x3")
    q = torch.as_tensor(_host_q(data, 128, 8), device=DEV)
    R_fast, ld_fast = fl.forward(q)
    R_slow, ld_slow = _forward_slow(fl, q)  # layer by layer with the converting kernel
    torch.cuda.synchronize()
    assert torch.equal(R_fast, R_slow) and torch.equal(ld_fast, ld_slow)
    fl.set_params(random_params(ref, 99, 0.05))
    R_new, _ = fl.forward(q)
    torch.cuda.synchronize()
    assert not torch.equal(R_new, R_fast)


def _forward_slow(fl, R0):
    B, D, h = R0.shape[0], fl.D, fl.hidden
    zp = torch.empty((B, D), device=DEV)
    work = torch.empty(2 * B * D, device=DEV)
    h1, h2, o = torch.empty((B, h), device=DEV), torch.empty((B, h), device=DEV), torch.empty((B, D), device=DEV)
    logdet = torch.zeros(B, dtype=torch.float64, device=DEV)
    R = torch.empty_like(R0)
    fl.ctx.flow_whiten(_native.FLOW_WHITEN, B, R0, work, zp)
    for l in range(fl.layers):
        a_off, b_off = fl.halves(l)
        fl._mlp(l, zp, B, h1, h2, o, None)
        fl.ctx.flow_coupling(False, B, D, a_off, b_off, fl.smax, zp, o, logdet)
    fl.ctx.flow_whiten(_native.FLOW_UNWHITEN, B, zp, work, R)
    return R, logdet


@pytest.mark.parametrize("N,P", [(4, 128), (8, 64)])
def test_fused_layer_matches_layer_by_layer(tmp_path, N, P):
    """One-launch coupling layers (D = 512 as N x P = 4 x 128 and 8 x 64) against the layer-by-layer pre-split
    path. The fused layer sums the same BF16x3 products in another order (transposed tiles,
    permuted inputs of the second and third products), so per layer z agrees to FP32 rounding
    (measured 1.5e-6 of max|z|) and log|det| to the sum of those roundings over D/2 features;
    the fused forward-inverse round trip of a layer returns its input to FP32 rounding. Through
    the whole 4-layer flow (random weights, smax = 2, which amplify rounding ~1e4-fold in the
    inverse direction) the forward R agrees to 8e-6 of max|R|."""
    _announce(f"fused coupling layer D={N * P}")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, N, quadratic=False, seed=2))
    data = make_data(FS, 300.0, P, 64, 64)
    fl = flow.Flow(data, layers=4, hidden=256, gemm="bf16x3")
    ref = ref_for(data, 4, 256, 2.0)
    fl.set_params(random_params(ref, 5, 0.05))
    B, D = 256, fl.D
    fl._split_weights(None)
    z0 = torch.randn((B, D), generator=torch.Generator().manual_seed(1)).to(DEV)
    o = torch.empty((B, D), device=DEV)
    for l in range(2):
        a_off, b_off = fl.halves(l)
        for inverse in (False, True):
            z1, l1 = z0.clone(), torch.zeros(B, dtype=torch.float64, device=DEV)
            z2, l2 = z0.clone(), torch.zeros(B, dtype=torch.float64, device=DEV)
            fl._layer_fused(inverse, l, z1, B, l1, None)
            fl._mlp_fast(l, z2, B, o, None)
            fl.ctx.flow_coupling(inverse, B, D, a_off, b_off, fl.smax, z2, o, l2, None)
            torch.cuda.synchronize()
            scale = float(z2.abs().max())
            dz, dl = float((z1 - z2).abs().max()), float((l1 - l2).abs().max())
            print(f"layer {l} inverse={inverse}: max |dz| {dz:.3e} (|z| <= {scale:.1f}), |d logdet| {dl:.3e}")
            assert torch.equal(z1[:, a_off:a_off + fl.H], z0[:, a_off:a_off + fl.H])
            assert dz <= 1e-5 * scale and dl <= 5e-4
            if not inverse:
                fl._layer_fused(True, l, z1, B, l1, None)
                torch.cuda.synchronize()
                assert float((z1 - z0).abs().max()) <= 1e-6 * scale
                assert float(l1.abs().max()) <= 1e-4
    q = torch.as_tensor(_host_q(data, B, 3), device=DEV)
    fl.use_fused = True
    R1, ld1 = fl.forward(q)
    fl.use_fused = False
    R2, ld2 = fl.forward(q)
    torch.cuda.synchronize()
    scale = float(R2.abs().max())
    dR, dld = float((R1 - R2).abs().max()), float((ld1 - ld2).abs().max())
    print(f"flow: max |dR| {dR:.3e} (|R| <= {scale:.2f}), |d logdet| {dld:.3e}")
    assert dR <= 5e-5 * scale and dld <= 1e-2


@pytest.mark.parametrize("N,frag", [(8, True), (8, False), (16, True), (16, False)])
def test_single_bf16_layer_is_an_exact_coupling(tmp_path, N, frag):
    """One-launch layers with single bf16 products (lo parts null; at D = 512 with frag the
    fragment-ordered weights of pib_flow_layer_fused_frag): the network differs from
    the BF16x3 one by bf16 rounding, but the layer is still an exact coupling. The conditioning
    half passes bit for bit; the inverse layer (the same scales, computed from the same half)
    returns the input to FP32 rounding; and the forward and inverse log-determinants cancel
    exactly (same scales, same summation order). The transformed half stays within bf16
    rounding of the network of the BF16x3 layer's (at D = 1024, which has no one-launch BF16x3
    layer, of the layer-by-layer BF16x3 path's)."""
    _announce(f"fused coupling layer, single bf16, D={N * 64}")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, N, quadratic=False, seed=2))
    data = make_data(FS, 300.0, 64, 64, 64)
    fl = flow.Flow(data, layers=2, hidden=256, gemm="bf16x3")
    fl.set_params(random_params(ref_for(data, 2, 256, 2.0), 5, 0.05))
    fl.use_frag = frag
    B, D = 256, fl.D
    fl._split_weights(None)
    z0 = torch.randn((B, D), generator=torch.Generator().manual_seed(1)).to(DEV)
    for l in range(2):
        a_off, b_off = fl.halves(l)
        z1, l1 = z0.clone(), torch.zeros(B, dtype=torch.float64, device=DEV)
        z3, l3 = z0.clone(), torch.zeros(B, dtype=torch.float64, device=DEV)
        fl._layer_fused(False, l, z1, B, l1, None, single_bf16=True)
        if D == 512:
            fl._layer_fused(False, l, z3, B, l3, None)
        else:
            o = torch.empty((B, D), device=DEV)
            fl._mlp_fast(l, z3, B, o, None)
            fl.ctx.flow_coupling(False, B, D, a_off, b_off, fl.smax, z3, o, l3, None)
        torch.cuda.synchronize()
        scale = float(z3.abs().max())
        dz, dl = float((z1 - z3).abs().max()), float((l1 - l3).abs().max())
        print(f"layer {l}: single bf16 vs bf16x3 max |dz| {dz:.3e} (|z| <= {scale:.1f}), |d logdet| {dl:.3e}")
        assert torch.equal(z1[:, a_off:a_off + fl.H], z0[:, a_off:a_off + fl.H])
        assert 0.0 < dz <= 2e-2 * scale
        fl._layer_fused(True, l, z1, B, l1, None, single_bf16=True)
        torch.cuda.synchronize()
        assert float((z1 - z0).abs().max()) <= 1e-6 * scale
        assert torch.equal(l1, torch.zeros_like(l1))
    # the lo parts are all given or all null
    fl._split_weights_fused(None)
    w = fl._wfused[0]
    with pytest.raises(_native.NativeError):
        fl.ctx.flow_layer_fused(False, B, D, 256, 0, D // 2, fl.smax, z0, l1,
                                ((w[0][0], None), w[1], w[2]), fl.p(0, "b1"), fl.p(0, "b2"), fl.p(0, "b3"))


@pytest.mark.parametrize("N,layer,frag", [(8, 0, True), (8, 1, True), (8, 0, False), (16, 0, True), (16, 1, True),
                                           (16, 0, False)])
def test_single_bf16_layer_matches_restatement(tmp_path, N, layer, frag):
    """The one-launch single-bf16 layer at D = 512 and 1024 (both kernels: fragment-ordered weights
    with frag, pib_flow_layer_fused with null lo parts without) against the network it computes, restated in
    float64 torch: the conditioning half and the weights rounded to bf16 (round to nearest
    even), h1 = bf16(relu(x W1 + b1)), h2 = bf16(relu(h1 W2 + b2)), (raw, t) = h2 W3 + b3,
    s = smax tanh(raw / smax), y = z e^s + t and log-det sum s. The kernel accumulates in FP32,
    so a hidden activation within an FP32 rounding of a bf16 rounding boundary may round the
    other way; such flips are rare and each moves y by about a bf16 ulp of one hidden unit.
    Bounds: 99% of the y elements within 1e-5 of the largest |y| (median within 1e-6), all
    within 2e-2; 95% of the log-dets within 1e-4 relative (or absolute below 1)."""
    _announce(f"single-bf16 layer {layer} (frag {frag}) vs float64 restatement, D={N * 64}")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, N, quadratic=False, seed=2))
    data = make_data(FS, 300.0, 64, 64, 64)
    fl = flow.Flow(data, layers=2, hidden=256, gemm="bf16x3")
    fl.set_params(random_params(ref_for(data, 2, 256, 2.0), 7, 0.05))
    fl.use_frag = frag
    B, D, H = 2048, fl.D, fl.H
    assert D == 64 * N
    fl._split_weights(None)
    z0 = torch.randn((B, D), generator=torch.Generator().manual_seed(3)).to(DEV)
    z1, l1 = z0.clone(), torch.zeros(B, dtype=torch.float64, device=DEV)
    fl._layer_fused(False, layer, z1, B, l1, None, single_bf16=True)
    torch.cuda.synchronize()

    def bf(t):
        return t.float().to(torch.bfloat16).double()

    a_off, b_off = fl.halves(layer)
    x = bf(z0[:, a_off:a_off + H].cpu())
    W1, W2, W3 = (bf(fl.p(layer, n).cpu()) for n in ("W1", "W2", "W3"))
    b1, b2, b3 = (fl.p(layer, n).cpu().double() for n in ("b1", "b2", "b3"))
    h1 = bf(torch.relu(x @ W1 + b1))
    h2 = bf(torch.relu(h1 @ W2 + b2))
    o = h2 @ W3 + b3
    sc = fl.smax * torch.tanh(o[:, :H] / fl.smax)
    y = z0[:, b_off:b_off + H].cpu().double() * torch.exp(sc) + o[:, H:]
    ld = sc.sum(1)
    got = z1[:, b_off:b_off + H].cpu().double()
    scale = float(y.abs().max())
    dz = ((got - y).abs() / scale).flatten()
    dl = (l1.cpu() - ld).abs() / ld.abs().clamp_min(1.0)
    q99, med = float(dz.quantile(0.99)), float(dz.median())
    print(f"y: median {med:.2e} q99 {q99:.2e} max {float(dz.max()):.2e} of {scale:.2f}; "
          f"log-det q95 {float(dl.quantile(0.95)):.2e} max {float(dl.max()):.2e}")
    assert torch.equal(z1[:, a_off:a_off + H], z0[:, a_off:a_off + H])
    assert med <= 1e-6 and q99 <= 1e-5 and float(dz.max()) <= 2e-2
    assert float(dl.quantile(0.95)) <= 1e-4


def test_split_frag32_layout(tmp_path):
    """pib_split_bf16 with PIB_SPLIT_FRAG32 against the layout restated in NumPy: element (n, k)
    of the transposed matrix (RNE bf16) at ((n/32 * K/16 + k/16) * 64 + 32 ((k % 16) / 8) +
    n % 32) * 8 + k % 8, for a K x N input (K inputs, N outputs); bit for bit. Also: the lo
    output may be omitted, and shapes outside the layout are refused."""
    _announce("split_bf16 frag32 layout")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 2, quadratic=False, seed=0))
    ctx = make_data(FS, 300.0, 64, 10, 10).context()
    K, N = 48, 96
    X = torch.randn((K, N), generator=torch.Generator().manual_seed(4), dtype=torch.float32)
    out = torch.empty(N * K, dtype=torch.int16, device=DEV)
    ctx.split_bf16(K, N, X.to(DEV), N, out, None, K, transpose=True, frag32=True)
    torch.cuda.synchronize()
    bits = X.to(torch.bfloat16).view(torch.int16).numpy()  # (K, N), round to nearest even
    n, k = np.meshgrid(np.arange(N), np.arange(K), indexing="xy")  # k rows, n columns
    off = ((n // 32 * (K // 16) + k // 16) * 64 + 32 * ((k % 16) // 8) + n % 32) * 8 + k % 8
    want = np.empty(N * K, dtype=np.int16)
    want[off.ravel()] = bits.ravel()
    assert np.array_equal(out.cpu().numpy(), want)
    with pytest.raises(_native.NativeError):
        ctx.split_bf16(K, 80, X.to(DEV), N, out, None, K, transpose=True, frag32=True)  # N % 32


def test_single_bf16_draws_after_training(tmp_path):
    """A trained flow at A=4, N=8, P=64 (D = 512: the one-launch layers) drawn with single-bf16
    layers: the same unbiased Z/Z_rho as its BF16x3 draws within their errors, and an ESS
    fraction within 10% of theirs (log q is exact for the map applied either way)."""
    _announce("single-bf16 draws of a trained flow")
    FS = stage_dict_models(tmp_path, synthetic.generate(4, 8, quadratic=True, seed=0))
    data = make_data(FS, 300.0, 64, 1024, 1024, seed=3)
    from pibronic_amd import learn
    fl = flow.Flow(data, layers=8, hidden=256, seed=1, gemm="bf16x3")
    learn.train(fl, 400, 2048, 1e-3, 2)
    B, n = 65536, 6
    lws = {}
    for prec in ("bf16x3", "bf16"):
        fl.draw_precision = prec
        parts = []
        for i in range(n):
            R, lq = fl.draw(B, 77, i * B)
            lw, sg = flow.log_weights(data, R, lq)  # log|g/q| (-inf where g underflows), sign
            parts.append((lw.cpu().numpy(), sg.cpu().numpy()))
        lws[prec] = (np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]))
    ref = max(v[0][np.isfinite(v[0])].max() for v in lws.values())
    res = {}
    for prec, (lw, sg) in lws.items():
        w = sg * np.exp(lw - ref)
        res[prec] = (w.mean(), w.std() / np.sqrt(w.size), w.sum() ** 2 / (w * w).sum() / w.size)
    print(res)
    (z3, e3, f3), (z1, e1, f1) = res["bf16x3"], res["bf16"]
    assert abs(z1 - z3) < 4 * np.hypot(e1, e3)
    assert f1 >= 0.9 * f3


def test_base_draw_density_matches_oracle(model):
    """pib_flow_draw_base: log rho at the drawn point equals the reference's log rho(R0) for
    R0 = m + (Z sigma) U^T formed in float64 from the float Z (oracle: the O-matrix product of
    pimc.py:1087-1136 in log form), to FP64 rounding."""
    _announce("flow base draw: log rho")
    data = model
    fl, ref, _ = _flow_and_ref(data)
    B = 512
    zp = torch.empty((B, fl.D), device=DEV)
    lr = torch.empty(B, dtype=torch.float64, device=DEV)
    fl.ctx.flow_draw_base(9, 100, B, zp, lr)
    torch.cuda.synchronize()
    Z = ref.unperm(zp.cpu().double())
    R0 = ref.m[None, :, None] + (Z * ref.sigma) @ ref.U.T
    want = ref.log_rho(R0)
    err = (lr.cpu() - want).abs().max()
    print("max |d log rho|", float(err), "|log rho| ~", float(want.abs().mean()))
    assert float(err) <= 1e-9 * (1.0 + float(want.abs().max()))


def test_base_draw_moments(model):
    """Every coordinate of the base draw has the mixture's mean sum_a w_a mu_a and variance
    1 + sum_a w_a mu_a^2 - (sum_a w_a mu_a)^2 (within 6 standard errors, B = 65536); draws
    with another offset are different numbers."""
    _announce("flow base draw: moments")
    data = model
    fl, ref, _ = _flow_and_ref(data)
    B = 65536
    zp = torch.empty((B, fl.D), device=DEV)
    lr = torch.empty(B, dtype=torch.float64, device=DEV)
    fl.ctx.flow_draw_base(3, 0, B, zp, lr)
    z2 = torch.empty((64, fl.D), device=DEV)
    fl.ctx.flow_draw_base(3, 64, 64, z2, lr[:64])
    torch.cuda.synchronize()
    z = zp.cpu().double()
    assert not torch.equal(z[:64], z2.cpu().double())
    assert torch.equal(z[64:128], z2.cpu().double())
    rho = data.rho
    m = np.asarray(rho.state_weight, dtype=np.float64) @ np.asarray(rho.state_shift)  # the flow's centre
    w = np.asarray(rho.state_weight, dtype=np.float64)
    w = w / w.sum()                                                                    # surface probabilities
    sigma = np.asarray(rho.standard_deviation[0])
    U = np.asarray(data.circulant_eigvects)
    mu = (np.asarray(rho.state_shift) - m[None, :])[:, :, None] * U.sum(axis=0)[None, None, :] / sigma[None]
    mu = ref.perm(torch.as_tensor(mu)).numpy()                     # (A_rho, D)
    mean = w @ mu
    var = 1.0 + w @ (mu * mu) - mean * mean
    se_m = np.sqrt(var / B)
    se_v = var * np.sqrt(2.0 / B) * 1.5
    zm, zv = z.mean(dim=0).numpy(), z.var(dim=0).numpy()
    print("max |mean err|/se", float(np.max(np.abs(zm - mean) / se_m)), "max |var err|/se", float(np.max(np.abs(zv - var) / se_v)))
    assert np.all(np.abs(zm - mean) < 6 * se_m)
    assert np.all(np.abs(zv - var) < 6 * se_v)


def test_identity_flow_reproduces_harmonic_weights(model):
    _announce("zero-initialised flow = rho")
    data = model
    fl = flow.Flow(data, layers=2, hidden=64)
    B = 512
    R, lq = fl.draw(B, seed=5, offset=0)
    lw, sg = flow.log_weights(data, R, lq)
    lh, sh = flow.harmonic_log_weights(data, R)
    torch.cuda.synchronize()
    assert torch.allclose(lw, lh, atol=1e-4)


def test_training_lowers_weight_variance(tmp_path):
    """A short on-policy training run (samples and weights drawn by the method itself): the
    trained proposal's Kish ESS fraction on fresh samples beats the harmonic rho's, and both
    give the same unbiased Z/Z_rho."""
    _announce("flow training gain A2 N2 P64")
    FS = stage_dict_models(tmp_path, synthetic.generate(2, 2, quadratic=True, seed=5))
    data = make_data(FS, 300.0, 64, 1024, 1024, seed=11)
    from pibronic_amd import learn
    res = learn.train_and_compare(data, layers=4, hidden=64, steps=300, batch=1024,
                                  eval_samples=65536, lr=1e-3, seed=11)
    print({k: v for k, v in res.items() if k not in ("flow", "train")})
    for st in res["train"]:
        print(st)
    assert res["ess_frac_flow"] > 1.5 * res["ess_frac_rho"]
    assert abs(res["z_flow"] - res["z_rho"]) < 4 * np.hypot(res["z_flow_err"], res["z_rho_err"])


def test_block_sums_with_logq_matches_host(model):
    """pib_block_sums_q: the device reduction with q as the denominator equals the same sums
    formed on the host from the per-sample g, log scale and log q."""
    _announce("block sums with log q")
    data = model
    fl = flow.Flow(data, layers=2, hidden=64)
    fl.set_params(random_params(ref_for(data, 2, 64, 2.0), 9, 0.05))
    B = 1024
    R, lq = fl.draw(B, seed=21, offset=0)
    out = torch.empty((_native.NUM_OUT, B), dtype=torch.float64, device=DEV)
    data.evaluate_device(R, B, False, out, B)
    sums = torch.zeros(_native.NUM_SUMS, dtype=torch.float64, device=DEV)
    data.context().block_sums_q(out, B, B, False, 1.0, lq, sums)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    r = o[_native.OUT_G] / np.exp(lq.cpu().numpy() - o[_native.OUT_LOGS])
    s = sums.cpu().numpy()
    assert s[_native.SUM_N] == B
    assert s[_native.SUM_R] == pytest.approx(r.sum(), rel=1e-12)
    assert s[_native.SUM_R2] == pytest.approx((r * r).sum(), rel=1e-12)
    assert s[_native.SUM_ABSR] == pytest.approx(np.abs(r).sum(), rel=1e-12)
    assert s[_native.SUM_NEG] == (r < 0).sum()
    # the same as the flow's own log-weight path
    lw, sg = flow.log_weights(data, R, lq)
    assert np.allclose(np.log(np.abs(r)), lw.cpu().numpy(), rtol=0, atol=1e-9)
    with pytest.raises(_native.NativeError):
        data.context().block_sums_q(out, B, B, False, 1.0, None, sums)


@pytest.mark.parametrize("which", ["A3_N2_P64", "A8_N16_P64"])
def test_offset_draws_match_double_draws(tmp_path, model, which):
    """Flow.draw(offsets=True) writes float offsets dR with m + dR (float64) equal to the double
    path's R bit for bit, the same log q, and pib_eval_block_flow_offsets on dR gives
    pib_eval_block's outputs on R bit for bit (the lanes kernel at A = 3 and at A = 8);
    +-dbeta is refused before any launch."""
    _announce(f"flow offset draws {which}")
    if which == "A3_N2_P64":
        data = model
    else:
        FS = synthetic.write_workload(str(tmp_path), "A8_N16_P64_T300")
        data = make_data(FS, 300.0, 64, 256, 256, seed=4)
    fl = flow.Flow(data, layers=2, hidden=64, smax=2.0, seed=3)
    assert fl.offsets_ok()
    B = 256
    R, lq1 = fl.draw(B, 5, 1000)
    dq, lq2 = fl.draw(B, 5, 1000, offsets=True)
    assert dq.dtype == torch.float32 and dq.shape == R.shape
    m = torch.as_tensor(np.asarray(data.rho.state_weight) @ np.asarray(data.rho.state_shift), dtype=torch.float64)
    assert torch.equal(lq1.cpu(), lq2.cpu())
    assert torch.equal(R.cpu(), m[None, :, None] + dq.cpu().double())
    out1 = torch.empty((_native.NUM_OUT, B), dtype=torch.float64, device=DEV)
    out2 = torch.full((_native.NUM_OUT, B), float("nan"), dtype=torch.float64, device=DEV)
    data.evaluate_device(R, B, False, out1, B)
    data.evaluate_offsets_device(dq, B, out2, B)
    torch.cuda.synchronize()
    rows = [_native.OUT_RHO, _native.OUT_G, _native.OUT_LOGS]
    assert torch.equal(out1[rows].cpu(), out2[rows].cpu())
    with pytest.raises(_native.NativeError, match="flags"):
        data.context().eval_block_flow_offsets(dq, B, _native.EVAL_PM, out2, B)


@pytest.mark.parametrize("name,B", [("A8_N16_P64_T300", 262144), ("A4_N8_P64_T300", 131072)])
def test_fragment_layer_draws_repeat_bit_for_bit(tmp_path, name, B):
    """Single-bf16 draws through k_flow_layer_ws at full size: the same draw twice gives the
    same R and log q bit for bit, and every R finite. The layer moves z through per-wave LDS
    slots with no barrier between its phases; a read of a slot ahead of the write that fills it
    shows as a few changed samples per million that differ from run to run."""
    _announce(f"fragment-layer draws repeat {name} B={B}")
    FS = synthetic.write_workload(str(tmp_path), name)
    cfg = synthetic.WORKLOADS[name]
    data = make_data(FS, cfg["T"], cfg["P"], 256, 256, seed=4)
    fl = flow.Flow(data, layers=8, hidden=256, seed=3, gemm="bf16x3")
    g = torch.Generator().manual_seed(11)
    fl.params += 0.02 * torch.randn(fl.params.shape, generator=g).to(fl.params.device, fl.params.dtype)
    fl._version += 1
    fl.draw_precision = "bf16"
    assert fl.use_frag
    R1, lq1 = fl.draw(B, 5, 0)
    R2, lq2 = fl.draw(B, 5, 0)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(R1).all())
    same = (R1 == R2).all(dim=2).all(dim=1) & (lq1 == lq2)
    assert bool(same.all()), f"{int((~same).sum())} of {B} samples differ between two identical draws"
