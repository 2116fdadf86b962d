"""Benchmark of the PIMC block estimator on one MI355X.

Default workload: A=8 surfaces, N=16 modes, P=64 beads, T=300 K, seeded synthetic model
(linear + mode-diagonal quadratic couplings). The metric is effective samples per second at a
fixed Z-ratio standard error, learned proposal against harmonic rho on the same GPU:

  * learned leg: the affine-coupling flow (8 layers, hidden 256, over the N*P = 1024 bead
    normal-mode coordinates) is trained on-policy on the method's own samples and weights
    (reported as train_seconds, outside the timed region), then one timed step is one draw of
    262,144 samples (base draw with its log rho, 8 coupling layers with single-bf16 products,
    unwhitening) + the FP64 evaluation of g + the on-device block sums with q as the
    denominator; the default K steps evaluate >= 1e7 samples;
  * harmonic leg: the same number of samples from rho in blocks of 1e5 (device sampler, FP64
    evaluation, block sums);
  * value = Kish ESS of w = g/q (|w|) over the timed learned samples / timed wall time.

Also in the line: the same timed draws with FP32-accurate flow products (BF16x3, and FP32
matrix-core products), tail diagnostics of both proposals' weights (largest weight share,
Pareto k-hat), and the plain FP64 kernel's samples*beads/s and roofline on the A=4, N=8, P=64
workload (blocks of 1e5) with the CPU oracle timed beside each. Inputs and outputs stay in HBM;
nothing is copied to the host inside a timed region.

The CPU baseline is the FP64 NumPy restatement of the reference algorithm (the oracle), one
process per core, on a bounded sample, before the GPU is touched.
"""
import argparse
import json
import math
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix peaks coincide)
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 matrix cores
HBM_PEAK_BPS = 8.0e12  # MI355X HBM3E
METRIC = "effective samples/sec at fixed Z-ratio std error; learned ρ vs harmonic ρ"
CPU_WORKER_CAP = 16  # worker processes per GPU box (its CPU share for one GPU)


def algorithmic_flops(A, N, P, n_quad_pairs, pm=False):
    """Algorithmic flops per sample, split by kernel (counting conventions: transform
    2 N P^2; V P A^2 (2N + 2Nq + 1); eig 9 A^3 per bead; M 2A^3 + A^2 per bead; O-factors
    A (7N + 1) per bead per set; ring 2A^3 + A^2 per bead per product; rho and scaling
    5 P A)."""
    sets = 4 if pm else 2
    prods = 3 if pm else 1
    sample = 2 * N * P * P
    ev = (P * A * A * (2 * N + 2 * n_quad_pairs + 1) + P * 9 * A ** 3 + P * (2 * A ** 3 + A * A)
          + P * sets * A * (7 * N + 1) + P * prods * (2 * A ** 3 + A * A) + 5 * P * A)
    return {"sample": sample, "eval": ev, "total": sample + ev}


def profiled_traffic(path, kernel_prefix):
    """HBM bytes per launch of a kernel from a committed rocprofv3 summary: FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each divided by what the same counter reads for a plain
    256 MiB device copy in the same run. Returns (bytes, detail) or (None, reason)."""
    if not os.path.exists(path):
        return None, "no profile"
    rows = {}
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0].startswith("`"):
            try:
                rows[(cells[0].strip("`"), cells[1])] = (float(cells[2]), float(cells[3]))
            except ValueError:
                pass
    copy = [v for (k, g), v in rows.items() if k.startswith("__amd_rocclr_copyBuffer")]
    kern = [v for (k, g), v in rows.items() if k.startswith(kernel_prefix)]
    if not copy or not kern:
        return None, "kernel or calibration copy missing from profile"
    ref_kib = 256 * 1024.0
    f_fetch, f_write = copy[0][0] / ref_kib, copy[0][1] / ref_kib
    fetch, write = kern[0][0] / f_fetch, kern[0][1] / f_write
    return (fetch + write) * 1024.0, {"fetch_kib_raw": kern[0][0], "write_kib_raw": kern[0][1],
                                      "fetch_factor": f_fetch, "write_factor": f_write,
                                      "profile": os.path.relpath(path, ROOT)}


# --------------------------------------------------------------------------------------
# CPU baseline (oracle = FP64 NumPy restatement of the reference algorithm)
# --------------------------------------------------------------------------------------
def _cpu_worker(args):
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    vib_path, rho_path, T, P, B, seconds, seed = args
    import numpy as np
    from oracle import pimc_oracle as po
    A, N, E, w, L, Q = po.read_model_json(vib_path)
    Ar, Nr, Er, wr, Lr, _ = po.read_model_json(rho_path, diagonal=True)
    s = po.Setup((E, w, L, Q), (Er, wr, Lr), T, P)
    rng = np.random.default_rng(seed)
    n, t0 = 0, time.perf_counter()
    sw, sw2 = 0.0, 0.0
    while True:
        q, _ = s.draw_q(rng, B)
        r = po.evaluate(s, q)
        ratio = np.abs(r["s_g"] / r["s_rho"])
        sw += ratio.sum()
        sw2 += (ratio * ratio).sum()
        n += B
        el = time.perf_counter() - t0
        if el >= seconds:
            return n, el, sw, sw2


def host_cores():
    """(cores used, os.cpu_count(), CPUs this process may run on)."""
    total = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = total
    return min(avail, CPU_WORKER_CAP), total, avail


def cpu_baseline(vib_path, rho_path, T, P, seconds, cores, B=500):
    import multiprocessing as mp
    import numpy as np
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[k] = "1"
    ctx = mp.get_context("spawn")
    jobs = [(vib_path, rho_path, T, P, B, seconds, 1000 + i) for i in range(cores)]
    with ctx.Pool(cores) as pool:
        res = pool.map(_cpu_worker, jobs)
    sbs = sum(n * P / el for n, el, _, _ in res)
    one = res[0][0] * P / res[0][1]
    n_tot = sum(r[0] for r in res)
    sw = sum(r[2] for r in res)
    sw2 = sum(r[3] for r in res)
    return dict(samples_beads_per_s=sbs, one_core=one, samples=n_tot, ess_frac=(sw * sw / sw2) / n_tot,
                seconds=max(r[1] for r in res), numpy=np.__version__, T=T, P=P, cores=cores)


def cpu_line(cpu, gpu_sbs, what):
    used, total, avail = host_cores()
    return {"value": cpu["samples_beads_per_s"], "unit": "samples*beads/s", "cores": cpu["cores"],
            "kind": "port",
            "sample": f"oracle (FP64 NumPy restatement of the reference algorithm, per-(sample, bead) "
                      f"dot loop, batched eigh) on {what}, {cpu['samples']} samples in blocks of 500 over "
                      f"{cpu['seconds']:.1f} s, one process per core, numpy {cpu['numpy']}",
            "host_cpu_count": total, "host_cpus_available": avail,
            "cores_note": f"one worker per available CPU, capped at {CPU_WORKER_CAP} (the GPU box's CPU "
                          "share for one GPU)",
            "one_core_value": cpu["one_core"], "gpu_over_cpu": gpu_sbs / cpu["samples_beads_per_s"]}


# --------------------------------------------------------------------------------------
def weight_summary(s, elapsed):
    """Effective-sample statistics from summed block sums over a timed region."""
    from pibronic_amd import _native
    n = s[_native.SUM_N]
    ess = s[_native.SUM_ABSR] ** 2 / s[_native.SUM_R2]
    mean_w = s[_native.SUM_ABSR] / n
    cv2 = max(s[_native.SUM_R2] / n - mean_w ** 2, 0.0) / mean_w ** 2
    n_needed = cv2 / 1e-6
    z = s[_native.SUM_R] / n
    z_err = math.sqrt(max(s[_native.SUM_R2] / n - z * z, 0.0) / (n - 1))
    return {"samples": n, "seconds": elapsed, "samples_per_s": n / elapsed,
            "ess_fraction": ess / n, "ess_per_s": ess / elapsed,
            "negative_fraction": s[_native.SUM_NEG] / n, "z_ratio": z, "z_ratio_err": z_err,
            "samples_for_1e-3_z_error": n_needed, "time_to_1e-3_z_error_s": n_needed * elapsed / n}


def pareto_khat(logw):
    """Shape of a generalized Pareto fit to the largest weights (Zhang & Stephens 2009 profile
    estimate, with the weakly informative adjustment of Vehtari et al., "Pareto smoothed
    importance sampling"): < 0.5 the weights have finite variance, > 0.7 the mean weight is not
    a reliable estimate. logw: numpy array of log |w|."""
    import numpy as np
    lw = np.sort(logw[np.isfinite(logw)])
    n = lw.size
    M = int(min(0.2 * n, 3.0 * math.sqrt(n)))
    if M < 10:
        return float("nan")
    top = np.exp(lw[-M:] - lw[-1])
    x = top - math.exp(lw[-M - 1] - lw[-1])      # exceedances over the (M+1)-th largest, sorted
    x = x[x > 0]
    m = x.size
    if m < 10:
        return float("nan")
    grid = 20 + int(math.sqrt(m))
    theta = 1.0 / x[-1] + (1.0 - np.sqrt(grid / (np.arange(1, grid + 1) - 0.5))) / (3.0 * x[int(m / 4 + 0.5) - 1])
    k = -np.mean(np.log1p(-theta[:, None] * x[None, :]), axis=1)
    prof = m * (np.log(theta / k) + k - 1.0)
    wts = 1.0 / np.exp(np.minimum(prof[None, :] - prof[:, None], 700.0)).sum(axis=1)
    th = float(np.sum(theta * wts))
    xi = float(np.mean(np.log1p(-th * x)))
    return (m * xi + 10 * 0.5) / (m + 10)


def tail_diagnostics(draw, n, chunk):
    """Largest weight share and Pareto k-hat over n fresh samples of a proposal (untimed).
    draw(B, offset) -> (log|w| tensor, sign tensor)."""
    import numpy as np
    lws = []
    for s in range(0, n, chunk):
        lw, _ = draw(min(chunk, n - s), s)
        lws.append(lw.double().cpu().numpy())
    lw = np.concatenate(lws)
    lw = lw[np.isfinite(lw)]
    w = np.exp(lw - lw.max())
    return {"samples": int(lw.size), "max_weight_share": float(w.max() / w.sum()),
            "pareto_khat": pareto_khat(lw)}


# --------------------------------------------------------------------------------------
def harmonic_leg(data, B, nblocks, W, seed, flags, pm):
    """nblocks timed blocks of B samples from rho (device sampler -> evaluation -> block
    sums), after W warm-up blocks. HIP events on the launch stream time every evaluation."""
    import torch
    from pibronic_amd import _native
    ctx = data.context()
    dev = data.device_buffers(B)
    stream = torch.cuda.current_stream()
    sums = torch.zeros((nblocks + W, _native.NUM_SUMS), dtype=torch.float64, device="cuda")
    ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(nblocks)]
    ev1 = [torch.cuda.Event(enable_timing=True) for _ in range(nblocks)]

    def step(i, timed):
        ctx.sample_block(seed, i * B, B, dev["q"], None, stream)
        if timed:
            ev0[i - W].record(stream)
        ctx.eval_block(dev["q"], B, flags, dev["out"], dev["out"].shape[1], None, stream)
        if timed:
            ev1[i - W].record(stream)
        ctx.block_sums(dev["out"], dev["out"].shape[1], B, pm, data.delta_beta, sums[i], stream)

    for i in range(W):
        step(i, False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(W, W + nblocks):
        step(i, True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    res = weight_summary(sums[W:].sum(dim=0).cpu().numpy(), elapsed)
    res["ms_per_step"] = elapsed / nblocks * 1e3
    res["block"] = B
    res["steps"] = nblocks
    res["eval_kernel_ms"] = sum(a.elapsed_time(b) for a, b in zip(ev0, ev1)) / nblocks
    return res


def eval_roofline(A, N, P, nq, pm, B, eval_ms, kernel, traffic=None, tdetail=None, coord_bytes=8):
    fl = algorithmic_flops(A, N, P, nq, pm)
    achieved = fl["eval"] * B / (eval_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
            "hbm_GBps": (traffic / (eval_ms * 1e-3) / 1e9) if traffic else None,
            "hbm_frac": (traffic / (eval_ms * 1e-3) / HBM_PEAK_BPS) if traffic else None,
            "kernel": kernel, "launch_samples": B, "kernel_ms": eval_ms,
            "flops_per_sample": fl["eval"], "flops_per_launch": fl["eval"] * B,
            "algorithmic_bytes": B * N * P * coord_bytes + B * (5 if pm else 3) * 8,
            "traffic_detail": tdetail or "not profiled for this launch size",
            "peak_note": "78.6 TFLOP/s: MI355X FP64 vector (= FP64 matrix) peak"}


def timed_draws(fl, data, B, K, W, seed, variant, eval_events=True):
    """K timed steps: flow draw + FP64 g + block sums with q as the denominator. The draws come
    as float offsets from the basis centre where the FP64 lanes kernel takes them (the same
    coordinates in FP64, half the bytes), else as R."""
    import torch
    from pibronic_amd import _native
    out = torch.empty((_native.NUM_OUT, B), dtype=torch.float64, device="cuda")
    sums = torch.zeros((K + W, _native.NUM_SUMS), dtype=torch.float64, device="cuda")
    ctx = data.context()
    stream = torch.cuda.current_stream()
    mk = lambda: [torch.cuda.Event(enable_timing=True) for _ in range(K)]  # noqa: E731
    d0, d1, l0, l1, g0, g1 = mk(), mk(), mk(), mk(), mk(), mk()
    offsets = fl.offsets_ok() and not data.fp32

    def step(i, timed):
        j = i - W
        if timed:
            d0[j].record(stream)
        R, logq = fl.draw(B, seed, (1 << 40) + i * B, stream, layer_events=(l0[j], l1[j]) if timed else None,
                          offsets=offsets)
        if timed:
            d1[j].record(stream)
            g0[j].record(stream)
        if offsets:
            data.evaluate_offsets_device(R, B, out, B, stream=stream)
        else:
            data.evaluate_device(R, B, False, out, B, stream=stream)
        if timed:
            g1[j].record(stream)
        ctx.block_sums_q(out, B, B, False, 1.0, logq, sums[i], stream)

    for i in range(W):
        step(i, False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(W, W + K):
        step(i, True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    avg = lambda a, b: sum(x.elapsed_time(y) for x, y in zip(a, b)) / K  # noqa: E731
    res = weight_summary(sums[W:].sum(dim=0).cpu().numpy(), el)
    res.update(variant=variant, coordinates="float offsets from the basis centre" if offsets else "R (double)",
               steps=K, samples_per_step=B, ms_per_step=el / K * 1e3,
               flow_draw_ms=avg(d0, d1), layers_ms=avg(l0, l1), eval_kernel_ms=avg(g0, g1))
    return res


def fp32_supported(data):
    """The FP32 kernel covers A <= 8 (lanes family)."""
    return data.states <= 8 and data.rho.states <= 8


def learned_leg(data, args, log):
    """Train the flow on-policy, then time K steps of the single-bf16 draws; the same timed
    steps with FP32-accurate flow products are reported beside them."""
    import torch
    from pibronic_amd import flow, learn
    t0 = time.perf_counter()
    fl = flow.Flow(data, layers=args.flow_layers, hidden=args.flow_hidden, seed=args.seed, gemm="bf16x3")
    hist = learn.train(fl, args.train_steps, args.train_batch, args.lr, args.seed + 1,
                       log=lambda st: log(f"[bench] train step {st['step']} ess {st['ess_frac']:.3e}"),
                       log_every=max(1, args.train_steps // 5))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    B, K, W = args.flow_block, args.steps, args.warmup
    fl.draw_precision = "bf16" if fl._fused_ok(B, True) else "bf16x3"
    main = timed_draws(fl, data, B, K, W, args.seed + 2, f"{fl.draw_precision} draws")
    log(f"[bench] learned ({fl.draw_precision}): {K} steps, ess fraction {main['ess_fraction']:.4f}")
    H, hdn, D = fl.H, fl.hidden, fl.D
    mlp_flops = 2.0 * B * fl.layers * (H * hdn + hdn * hdn + hdn * D)
    tf = mlp_flops / (main["layers_ms"] * 1e-3) / 1e12
    single = fl.draw_precision == "bf16"
    main["roofline_layers"] = {
        "kernel": "k_flow_layer_ws (one launch per coupling layer)" if single else "BF16x3 layer-by-layer",
        "bound": "mfma", "achieved": tf, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": tf / BF16_PEAK_TFLOPS, "matrix_core_frac": (1 if single else 3) * tf / BF16_PEAK_TFLOPS,
        "layers_ms": main["layers_ms"], "flops_per_draw": mlp_flops,
        "note": "HIP events around the coupling layers of each timed draw"}
    if single:
        # a one-launch layer reads both halves of z (B x D float) and writes the transformed half,
        # adds to the log-det (read and write of B doubles) and reads its bf16 weights once
        layer_bytes = B * (3 * H * 4 + 16) + 2 * (H * hdn + hdn * hdn + hdn * D)
        hbm = fl.layers * layer_bytes / (main["layers_ms"] * 1e-3)
        main["roofline_layers"].update({"hbm_frac": hbm / HBM_PEAK_BPS, "bytes_per_draw": fl.layers * layer_bytes})
        if layer_bytes / HBM_PEAK_BPS > mlp_flops / fl.layers / (BF16_PEAK_TFLOPS * 1e12):
            # HBM is the larger floor: name it as the bound (the matrix-core fraction stays beside it)
            main["roofline_layers"].update({"bound": "hbm", "achieved": hbm / 1e9, "peak": HBM_PEAK_BPS / 1e9,
                                            "unit": "GB/s", "frac": hbm / HBM_PEAK_BPS})
    variants = []
    if args.variants:
        Kv = max(1, min(K, args.variant_steps))
        fl.draw_precision = "bf16x3"
        variants.append(timed_draws(fl, data, B, Kv, 1, args.seed + 3, "bf16x3 draws (FP32-accurate products)"))
        fl.gemm = "f32"
        variants.append(timed_draws(fl, data, B, Kv, 1, args.seed + 4, "fp32 draws (FP32 matrix-core products)"))
        fl.gemm = "bf16x3"
        # the precision pattern on the learned step: g by the FP32 kernel (PIB_EVAL_FP32), with
        # its error against the FP64 kernel on one draw's samples
        if fp32_supported(data):
            import torch
            from pibronic_amd.precision import precision_report
            fl.draw_precision = "bf16" if single else "bf16x3"
            data.fp32 = True
            try:
                v = timed_draws(fl, data, B, Kv, 1, args.seed + 5, f"{fl.draw_precision} draws, g by the FP32 kernel")
            finally:
                data.fp32 = False
            R, _ = fl.draw(B, args.seed + 6, 1 << 42)
            rep = precision_report(data, R, B, pm=False, variant="fp32")
            v["fp32_error_vs_fp64"] = {"r_median_rel": rep["r"]["median_rel"], "r_max_rel": rep["r"]["max_rel"],
                                       "r_max_abs_over_mean_abs": rep["r"]["max_abs_over_mean_abs"],
                                       "z_shift_sigma": rep["z_shift_sigma"]}
            variants.append(v)
            del R
            torch.cuda.synchronize()
        for v in variants:
            log(f"[bench] learned {v['variant']}: ess fraction {v['ess_fraction']:.4f}, "
                f"{v['ess_per_s']:.3e} ESS/s")
    fl.draw_precision = "bf16" if single else "bf16x3"
    main.update(train_seconds=train_s, train_steps=args.train_steps, train_batch=args.train_batch,
                layers=args.flow_layers, hidden=args.flow_hidden,
                train_final_batch_ess=hist[-1]["ess_frac"] if hist else None,
                variants=variants)
    main["time_to_1e-3_z_error_incl_training_s"] = train_s + main["time_to_1e-3_z_error_s"]
    return fl, main


def flow_tails(fl, data, n, seed):
    from pibronic_amd import flow as flow_mod

    def draw(B, off):
        R, lq = fl.draw(B, seed, (1 << 44) + off)
        return flow_mod.log_weights(data, R, lq)
    return tail_diagnostics(draw, n, 65536)


def rho_tails(data, n, seed):
    import torch
    from pibronic_amd import flow as flow_mod

    def draw(B, off):
        R = torch.empty((B, data.modes, data.beads), dtype=torch.float64, device="cuda")
        data.context().sample_block(seed, (1 << 44) + off, B, R)
        return flow_mod.harmonic_log_weights(data, R)
    return tail_diagnostics(draw, n, 65536)


def setup_data(FS, cfg, B, blocks, seed, pm=False):
    from pibronic_amd import pimc
    data = (pimc.BoxDataPM if pm else pimc.BoxData).from_FileStructure(FS)
    data.samples, data.beads, data.temperature = B * blocks, cfg["P"], cfg["T"]
    data.block_size, data.blocks = B, blocks
    data.seed = seed
    data.preprocess()
    return data


def n_quad_pairs(data):
    import numpy as np
    return int(np.count_nonzero(np.abs(data.vib.quadratic).sum(axis=(2, 3))))


def pm_leg(FS, cfg, B, steps, W, seed, log, label):
    """The same harmonic blocks with g at beta and beta +- dbeta (PIB_EVAL_PM, the evaluation of
    block_compute_pm, the reference's job path), timed like the plain leg."""
    from pibronic_amd import _native
    A, N, P = cfg["A"], cfg["N"], cfg["P"]
    data = setup_data(FS, cfg, B, steps + W, seed, pm=True)
    h = harmonic_leg(data, B, steps, W, seed, _native.EVAL_PM, True)
    sbs = h["samples"] * P / h["seconds"]
    kname = ("k_eval_fp64_chain<8, 8, 1..3> (three single-chain passes)" if A == 8
             else f"k_eval_fp64<{A}, {A}, true, false>" if A <= 8 else "k_eval_wide (+-dbeta)")
    log(f"[bench] FP64 +-dbeta {label}: {sbs:.3e} samples*beads/s, eval {h['eval_kernel_ms']:.3f} ms")
    return {"samples_beads_per_s": sbs, "ms_per_step": h["ms_per_step"], "steps": h["steps"], "block": B,
            "eval_kernel_ms": h["eval_kernel_ms"],
            "roofline": eval_roofline(A, N, P, n_quad_pairs(data), True, B, h["eval_kernel_ms"], kname),
            "note": "harmonic blocks with g at beta and beta +- dbeta (block_compute_pm's evaluation); "
                    "algorithmic flops count M once and three ring products"}


def fp64_line(FS, name, cfg, args, cpu, log):
    """The plain FP64 kernel on one workload: harmonic blocks of 1e5, samples*beads/s and the
    evaluation kernel's roofline (traffic from the committed counter profile when there is one)."""
    import torch
    from pibronic_amd import pimc
    A, N, P = cfg["A"], cfg["N"], cfg["P"]
    B = args.block
    data = setup_data(FS, cfg, B, args.fp64_steps + args.warmup, args.seed)
    lanes = data.context().lanes_per_sample(B)
    # this line runs first on an idle GPU: untimed blocks for args.fp64_warm_s seconds bring the
    # clocks up before the W warm-up blocks and the timed ones
    ctx, dev = data.context(), data.device_buffers(B)
    t_end = time.perf_counter() + args.fp64_warm_s
    while time.perf_counter() < t_end:
        for _ in range(8):
            ctx.sample_block(args.seed, 0, B, dev["q"], None)
            ctx.eval_block(dev["q"], B, 0, dev["out"], dev["out"].shape[1], None)
        torch.cuda.synchronize()
    h = harmonic_leg(data, B, args.fp64_steps, args.warmup, args.seed, 0, False)
    sbs = h["samples"] * P / h["seconds"]
    kname = f"k_eval_fp64<{A}, {A}, false, false>" if A <= 8 else "k_eval_wide"
    traffic, tdetail = profiled_traffic(os.path.join(ROOT, "profiles", f"k_eval_fp64_A{A}_N{N}_P{P}_B{B}.md"),
                                        f"void pib::k_eval_fp64<{A}, {A}, false, false>")
    log(f"[bench] FP64 {name}: {sbs:.3e} samples*beads/s, eval {h['eval_kernel_ms']:.3f} ms")
    line = {"workload": f"{name}: A={A} N={N} P={P} T={cfg['T']:g}K block={B}, FP64, harmonic rho",
            "samples_beads_per_s": sbs, "samples_per_s": h["samples_per_s"], "ms_per_step": h["ms_per_step"],
            "steps": h["steps"], "lanes_per_sample": lanes, "eval_kernel_ms": h["eval_kernel_ms"],
            "ess_fraction": h["ess_fraction"],
            "roofline": eval_roofline(A, N, P, n_quad_pairs(data), False, B, h["eval_kernel_ms"], kname,
                                      traffic, tdetail),
            "cpu_baseline": cpu_line(cpu, sbs, f"the {name} model") if cpu else None,
            "pm": pm_leg(FS, cfg, B, args.fp64_steps, args.warmup, args.seed + 3, log, name)}
    # the reference-API path: block_compute (device sampling, evaluation, sums, and the copy of
    # every block's per-sample results to the host arrays of BoxResult)
    api = setup_data(FS, cfg, B, args.fp64_steps, args.seed + 1)
    res = pimc.BoxResult(data=api)
    pimc.block_compute(api, res)  # warm-up: context, buffers
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pimc.block_compute(api, res)
    el = time.perf_counter() - t0
    line["block_compute_api"] = {"samples_beads_per_s": B * args.fp64_steps * P / el,
                                 "ms_per_block": el / args.fp64_steps * 1e3,
                                 "note": "pimc.block_compute end to end, incl. the per-block device-to-host "
                                         "copies into BoxResult"}
    return line


def batched_leg(FS, cfg, args, cpu, log):
    """Sixteen temperatures of one model (T_i = 100 * 5^(i/15) K, P_i = ceil(beta_i / tau0))
    in one block step per call, the block step replayed as a hipGraph."""
    import numpy as np
    import torch
    from pibronic_amd import pimc, synthetic
    from pibronic_amd.batch import TemperatureBatch
    Ts, Ps, tau0 = synthetic.c3_temperatures()
    B, K, W = args.block, args.fp64_steps, args.warmup
    datas = []
    for T, P in zip(Ts, Ps):
        d = pimc.BoxData.from_FileStructure(FS)
        d.samples, d.beads, d.temperature = B * (K + W), int(P), float(T)
        d.block_size, d.blocks = B, K + W
        d.preprocess()
        datas.append(d)
    tb = TemperatureBatch(datas, B, max_blocks=K + W, fp32=args.fp32)
    tune = tb.tune(seed=args.seed)
    log(f"[bench] batched tune: {tune}")
    tb.run(args.seed, 0, W)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tb.run(args.seed, W, K)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    sums = tb.block_sums()[W:W + K]
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        tb.batch.eval(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    eval_ms = e0.elapsed_time(e1) / 3
    nq = n_quad_pairs(datas[0])
    flops = sum(algorithmic_flops(cfg["A"], cfg["N"], int(P), nq, False)["eval"] for P in Ps) * B
    errors = None
    if args.fp32:
        from pibronic_amd.precision import precision_report
        errors = []
        for i, d in enumerate(datas):
            rep = precision_report(d, tb.q[i], B, pm=False)
            errors.append({"T": float(Ts[i]), "P": int(Ps[i]), "r_median_rel": rep["r"]["median_rel"],
                           "r_max_abs_over_mean_abs": rep["r"]["max_abs_over_mean_abs"],
                           "s_rho_max_rel": rep["s_rho"]["max_rel"], "z_shift_sigma": rep["z_shift_sigma"]})
    achieved = flops / (eval_ms * 1e-3) / 1e12
    peak = FP32_PEAK_TFLOPS if args.fp32 else FP64_PEAK_TFLOPS
    per_t, ess_tot = [], 0.0
    for i in range(len(Ts)):
        w = weight_summary(sums[:, i].sum(axis=0), el)
        ess_tot += w["ess_per_s"] * el
        per_t.append({"T": float(Ts[i]), "P": int(Ps[i]), "ess_fraction": w["ess_fraction"],
                      "z_ratio": w["z_ratio"], "z_ratio_err": w["z_ratio_err"]})
    beads = int(np.sum(Ps))
    sbs = B * beads * K / el
    return {
        "metric": METRIC, "value": ess_tot / el, "unit": "effective samples/s (harmonic rho, all temperatures)",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": el / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.fp32 else "f64", "data": "synthetic (seeded model, device Philox samples)",
        "config": {"workload": f"A6_N12_T100-500: A={cfg['A']} N={cfg['N']}, 16 temperatures "
                               f"100-500 K, P_i = ceil(beta_i/tau0), tau0 = beta(100 K)/256 "
                               f"(P {int(Ps.min())}-{int(Ps.max())}), block={B} per temperature",
                   "samples_per_step": B * len(Ts), "schedule": tune["schedule"],
                   "value_from": "harmonic rho; block step (16 samplers, evaluation, 16 reductions) "
                                 "replayed as one hipGraph, evaluation schedule chosen by the autotuner"},
        "samples_beads_per_s": sbs, "samples_per_s": B * len(Ts) * K / el,
        "batched_eval_ms": eval_ms, "tuned_batched_ms": tune["batched_ms"],
        "sequential_eval_ms": tune["sequential_ms"],
        "batched_speedup_vs_sequential": tune["sequential_ms"] / tune["batched_ms"],
        "temperatures": per_t, "fp32_error_vs_fp64": errors,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None,
                     "kernel": {"lanes": "k_eval_batched (persistent)", "wide": "k_eval_wide_batched (persistent)",
                                "sequential": f"k_eval_{'fp32' if args.fp32 else 'fp64'}, one launch per "
                                              "temperature in the graph step"}[tune["schedule"]["family"]],
                     "peak_note": ("157.3 TFLOP/s: MI355X FP32 vector peak (the FP32 kernel)" if args.fp32
                                   else "78.6 TFLOP/s: MI355X FP64 vector (= FP64 matrix) peak"),
                     "flops_per_launch": flops, "traffic_detail": "not profiled for this workload"},
        "cpu_baseline": cpu_line(cpu, sbs, f"the middle temperature (T={cpu['T']:.1f} K, P={cpu['P']})")
        if cpu else None,
    }


def calibration_copy(mib):
    """Three device copies of a `mib` MiB buffer (the known transfer that calibrates the
    profiler's FETCH_SIZE / WRITE_SIZE counters); nothing when mib is 0."""
    if not mib:
        return
    import torch
    nel = mib * (1 << 20) // 8
    src_t = torch.ones(nel, dtype=torch.float64, device="cuda")
    dst_t = torch.empty_like(src_t)
    for _ in range(3):
        dst_t.copy_(src_t)
    torch.cuda.synchronize()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=39,
                    help="timed learned steps of --flow-block samples (39 x 262144 >= 1e7 samples)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--block", type=int, default=100_000, help="harmonic block size")
    ap.add_argument("--workload", default="A8_N16_P64_T300")
    ap.add_argument("--fp64-workload", default="A4_N8_P64_T300",
                    help="workload of the plain FP64 kernel's samples*beads/s line ('none' to skip)")
    ap.add_argument("--fp64-steps", type=int, default=20, help="timed blocks of the FP64 line")
    ap.add_argument("--fp64-warm-s", type=float, default=0.5,
                    help="seconds of untimed blocks before the FP64 line's warm-up and timed blocks")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--no-variants", dest="variants", action="store_false",
                    help="skip the FP32-accurate flow variants")
    ap.add_argument("--variant-steps", type=int, default=13, help="timed steps of each flow-precision variant")
    ap.add_argument("--tail-samples", type=int, default=1 << 22,
                    help="fresh samples per proposal for the tail diagnostics (0 to skip)")
    ap.add_argument("--fp32", action="store_true", help="batched workload: the FP32 kernel")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--train-batch", type=int, default=2048)
    ap.add_argument("--flow-block", type=int, default=262144,
                    help="samples per learned step (per draw and evaluation launch); 262,144 measured 3%% "
                         "faster per sample than 65,536 or 131,072")
    ap.add_argument("--flow-layers", type=int, default=8)
    ap.add_argument("--flow-hidden", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the timed regions, copy a buffer of this many MiB 3 times "
                         "(calibrates profiler byte counters against a known transfer)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not "
                                   "supported (samples shard trivially, one process per GPU)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    from pibronic_amd import synthetic
    name = args.workload
    cfg = synthetic.WORKLOADS[name]
    tmp = tempfile.mkdtemp(prefix="pib_bench_")
    FS = synthetic.write_workload(tmp, name)
    cores = host_cores()[0]

    if cfg["P"] is None:  # the sixteen-temperature batch
        cpu = None
        if not args.no_cpu:
            Ts, Ps, _ = synthetic.c3_temperatures()
            cpu = cpu_baseline(FS.path_vib_model, FS.path_rho_model, float(Ts[8]), int(Ps[8]),
                               args.cpu_seconds, cores)
        line = batched_leg(FS, cfg, args, cpu, log)
        calibration_copy(args.calib_copy_mib)
        print(json.dumps(line))
        return

    fp64_name = None if args.fp64_workload == "none" else args.fp64_workload
    FS64 = synthetic.write_workload(tempfile.mkdtemp(prefix="pib_bench64_"), fp64_name) if fp64_name else None
    cpu = cpu64 = None
    if not args.no_cpu:
        cpu = cpu_baseline(FS.path_vib_model, FS.path_rho_model, cfg["T"], cfg["P"], args.cpu_seconds, cores)
        log(f"[bench] cpu baseline {name}: {cpu['samples_beads_per_s']:.3e} samples*beads/s on {cores} cores")
        if FS64:
            c64 = synthetic.WORKLOADS[fp64_name]
            cpu64 = cpu_baseline(FS64.path_vib_model, FS64.path_rho_model, c64["T"], c64["P"],
                                 args.cpu_seconds, cores)
            log(f"[bench] cpu baseline {fp64_name}: {cpu64['samples_beads_per_s']:.3e} samples*beads/s")

    import torch  # noqa: F401  (the GPU is first touched below)
    # the plain FP64 kernel's line first, on a GPU that has not yet run the flow's matrix-core work
    fp64 = fp64_line(FS64, fp64_name, synthetic.WORKLOADS[fp64_name], args, cpu64, log) if FS64 else None
    A, N, P, T = cfg["A"], cfg["N"], cfg["P"], cfg["T"]
    B_l, K, W = args.flow_block, args.steps, args.warmup
    # harmonic rho on the same model and at least as many samples as the timed learned steps
    nblocks = max(1, -(-K * B_l // args.block))
    data = setup_data(FS, cfg, args.block, nblocks + W, args.seed)
    harm = harmonic_leg(data, args.block, nblocks, W, args.seed, 0, False)
    log(f"[bench] harmonic {name}: {nblocks} blocks, ess fraction {harm['ess_fraction']:.3e}")
    nq = n_quad_pairs(data)
    kname = f"k_eval_fp64<{A}, {A}, false, false>" if A <= 8 else "k_eval_wide"
    harm_pm = pm_leg(FS, cfg, args.block, max(1, min(nblocks, args.fp64_steps)), W, args.seed + 4, log, name)

    learned = fl = None
    if not args.no_learned and P % 64 == 0:
        fl, learned = learned_leg(data, args, log)
    elif not args.no_learned:
        log(f"[bench] learned leg skipped: P = {P} is not a multiple of 64 (the flow's whitening tiles)")

    tails = None
    if args.tail_samples:
        tails = {"harmonic": rho_tails(data, args.tail_samples, args.seed + 7)}
        if fl is not None:
            tails["learned"] = flow_tails(fl, data, args.tail_samples, args.seed + 8)

    calibration_copy(args.calib_copy_mib)

    harm_sbs = harm["samples"] * P / harm["seconds"]
    if learned:
        off = learned.get("coordinates", "").startswith("float offsets")
        kfull = f"void pib::k_eval_fp64_offsets<{A}>" if off else f"void pib::k_eval_fp64<{A}, {A}, false, false>"
        roof = eval_roofline(A, N, P, nq, False, B_l, learned["eval_kernel_ms"], kfull[len("void pib::"):],
                             *profiled_traffic(os.path.join(ROOT, "profiles", f"k_eval_fp64_A{A}_N{N}_P{P}_B{B_l}.md"),
                                               kfull), coord_bytes=4 if off else 8)
        roof["step_note"] = "the FP64 evaluation of g inside the timed learned step (its largest kernel)"
        value = learned["ess_per_s"]
    else:
        roof = eval_roofline(A, N, P, nq, False, args.block, harm["eval_kernel_ms"], kname)
        value = harm["ess_per_s"]
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "effective samples/s" + ("" if learned else " (harmonic rho)"),
        "n_gpus": 1, "steps": learned["steps"] if learned else harm["steps"], "warmup": W,
        "ms_per_step": learned["ms_per_step"] if learned else harm["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("bf16 flow draws (single-bf16 matrix-core products, FP32 accumulation and flow "
                  "state), f64 g and weights") if learned else "f64",
        "data": "synthetic (seeded model, device Philox samples)",
        "config": {"workload": f"{name}: A={A} N={N} P={P} T={T:g}K",
                   "samples_per_step": learned["samples_per_step"] if learned else args.block,
                   "flow": (f"{args.flow_layers} affine-coupling layers, hidden {args.flow_hidden}, "
                            f"D = N*P = {N * P}, trained on-policy with BF16x3 products "
                            f"({args.train_steps} steps of {args.train_batch})") if learned else None,
                   "harmonic_block": args.block,
                   "value_from": ("learned proposal: Kish ESS over the timed samples / timed wall time "
                                  "(ms_per_step: one draw of samples_per_step + FP64 g + sums)")
                                 if learned else "harmonic rho"},
        "evaluated_samples": learned["samples"] if learned else harm["samples"],
        "learned": learned,
        "harmonic": harm,
        "learned_over_harmonic_ess_per_s": (learned["ess_per_s"] / harm["ess_per_s"]) if learned else None,
        "learned_over_harmonic_incl_training": (
            harm["time_to_1e-3_z_error_s"] / learned["time_to_1e-3_z_error_incl_training_s"]) if learned else None,
        "tail_diagnostics": tails,
        "samples_beads_per_s": harm_sbs,
        "samples_beads_per_s_note": f"harmonic FP64 step on {name} (blocks of {args.block})",
        "harmonic_pm": harm_pm,
        "roofline": roof,
        "cpu_baseline": cpu_line(cpu, harm_sbs, f"the {name} model") if cpu else None,
        "fp64_kernel": fp64,
    }
    if cpu and learned:
        line["cpu_baseline"]["effective_samples_per_s"] = cpu["samples_beads_per_s"] / P * cpu["ess_frac"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
