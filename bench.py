"""Benchmark of the PIMC block estimator on one MI355X.

Workload (default): A=4 surfaces, N=8 modes, P=64 beads, T=300 K, seeded synthetic model
(linear + mode-diagonal quadratic couplings), blocks of 1e5 samples, FP64. One step is one
block through the hot path: device bead sampling, the fused evaluation of g and rho
(block_compute's per-block body) and the on-device block sums. Inputs and outputs stay
in HBM; the per-sample results are not copied to the host inside the timed region.

Prints one JSON line. `value` is effective samples per second (Kish ESS of w = |g/rho|
over the timed samples, divided by the timed wall time); samples*beads/s is reported
beside it. The CPU baseline is the FP64 NumPy restatement of the reference algorithm
(the oracle), one process per host core, timed on a bounded sample before the GPU is
touched.
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
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16 matrix cores
HBM_PEAK_BPS = 8.0e12  # MI355X HBM3E
HBM_PEAK_GBS = 8000.0


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
    256 MiB device copy in the same run (FETCH_SIZE reads 0.5x, WRITE_SIZE 1.0x on MI355X).
    Returns (bytes, detail) or (None, reason)."""
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


def cpu_baseline(vib_path, rho_path, T, P, seconds, cores, B=500):
    import multiprocessing as mp
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
    import numpy as np
    return dict(samples_beads_per_s=sbs, one_core=one, samples=n_tot, ess_frac=(sw * sw / sw2) / n_tot,
                seconds=max(r[1] for r in res), numpy=np.__version__)


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


def learned_leg(data, args, log):
    """Train the flow on-policy, then time K steps of: flow draw (device sampler + log rho +
    flow forward), FP64 evaluation of g, block sums with q as the denominator."""
    import torch
    from pibronic_amd import _native, flow, learn
    t0 = time.perf_counter()
    fl = flow.Flow(data, layers=args.flow_layers, hidden=args.flow_hidden, seed=args.seed,
                   gemm=args.flow_gemm)
    fl.use_fused = bool(args.flow_fused)
    hist = learn.train(fl, args.train_steps, args.train_batch, args.lr, args.seed + 1,
                       log=lambda st: log(f"[bench] train step {st['step']} ess {st['ess_frac']:.3e}"),
                       log_every=max(1, args.train_steps // 5))
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    # the timed draws: single-bf16 layers only where the one-launch layer runs
    draw_prec = args.flow_draw if (args.flow_gemm == "bf16x3" and fl._fused_ok(args.flow_block, True)) else "bf16x3"
    fl.draw_precision = draw_prec
    B = args.flow_block
    K, W = args.steps, args.warmup
    out = torch.empty((_native.NUM_OUT, B), dtype=torch.float64, device="cuda")
    sums = torch.zeros((K + W, _native.NUM_SUMS), dtype=torch.float64, device="cuda")
    ctx = data.context()
    stream = torch.cuda.current_stream()

    e0 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    e1 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    l0 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    l1 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]

    def step(i, timed):
        if timed:
            e0[i - W].record(stream)
        R, logq = fl.draw(B, args.seed + 2, (1 << 40) + i * B, stream,
                          layer_events=(l0[i - W], l1[i - W]) if timed else None)
        if timed:
            e1[i - W].record(stream)
        data.evaluate_device(R, B, False, out, B, stream=stream)
        ctx.block_sums_q(out, B, B, False, 1.0, logq, sums[i], stream)

    for i in range(W):
        step(i, False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for i in range(W, W + K):
        step(i, True)
    torch.cuda.synchronize()
    el = time.perf_counter() - t1
    draw_ms = sum(a.elapsed_time(b) for a, b in zip(e0, e1)) / K
    layers_ms = sum(a.elapsed_time(b) for a, b in zip(l0, l1)) / K
    H, hdn, D = fl.H, fl.hidden, fl.D
    mlp_flops = 2.0 * B * fl.layers * (H * hdn + hdn * hdn + hdn * D)
    fused = fl.use_fused and fl._fused_ok(B, draw_prec == "bf16")
    # the coupling layers (the learned step's largest cost) against the bf16 matrix cores:
    # algorithmic flops are the dense products' 2 m n k; BF16x3 issues three bf16 products
    # for each (single-bf16 draws one), so the matrix cores execute that multiple
    layer_tflops = mlp_flops / (layers_ms * 1e-3) / 1e12
    products = 1 if draw_prec == "bf16" else 3
    layer_roofline = {"kernel": ("k_flow_layer (one launch per coupling layer)" if fused
                                 else "k_gemm_bf16x3_pre + k_coupling per layer"),
                      "bound": "mfma", "achieved": layer_tflops, "peak": BF16_PEAK_TFLOPS,
                      "unit": "TFLOP/s", "frac": layer_tflops / BF16_PEAK_TFLOPS,
                      "matrix_core_frac": products * layer_tflops / BF16_PEAK_TFLOPS,
                      "layers_ms": layers_ms, "flops_per_draw": mlp_flops,
                      "note": "HIP events around the coupling layers of each timed draw"}
    res = weight_summary(sums[W:].sum(dim=0).cpu().numpy(), el)
    res.update(train_seconds=train_s, train_steps=args.train_steps, train_batch=args.train_batch,
               samples_per_step=B, layers=args.flow_layers, hidden=args.flow_hidden,
               gemm=args.flow_gemm, draw_precision=draw_prec, flow_draw_ms=draw_ms,
               flow_mlp_tflops=mlp_flops / (draw_ms * 1e-3) / 1e12,
               flow_mlp_flops_per_draw=mlp_flops, roofline=layer_roofline,
               train_final_batch_ess=hist[-1]["ess_frac"] if hist else None)
    return res


def batched_leg(FS, cfg, args, cpu, log):
    """Sixteen temperatures of one model (T_i = 100 * 5^(i/15) K, P_i = ceil(beta_i / tau0))
    in one persistent launch per block, the block step replayed as a hipGraph."""
    import numpy as np
    import torch
    from pibronic_amd import _native, pimc, synthetic
    from pibronic_amd.batch import TemperatureBatch
    Ts, Ps, tau0 = synthetic.c3_temperatures()
    B, K, W = args.block, args.steps, args.warmup
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
    # evaluation launch alone (HIP events on the launch stream), for the roofline
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        tb.batch.eval(stream)
    e1.record(stream)
    torch.cuda.synchronize()
    eval_ms = e0.elapsed_time(e1) / 3
    nq = int(np.count_nonzero(np.abs(datas[0].vib.quadratic).sum(axis=(2, 3))))
    flops = sum(algorithmic_flops(cfg["A"], cfg["N"], int(P), nq, False)["eval"] for P in Ps) * B
    errors = None
    if args.fp32:  # FP32 vs FP64 on the last block's samples, per temperature
        from pibronic_amd.precision import precision_report
        errors = []
        for i, d in enumerate(datas):
            rep = precision_report(d, tb.q[i], B, pm=False)
            errors.append({"T": float(Ts[i]), "P": int(Ps[i]), "r_median_rel": rep["r"]["median_rel"],
                           "r_max_rel": rep["r"]["max_rel"], "s_rho_max_rel": rep["s_rho"]["max_rel"],
                           "z_shift_sigma": rep["z_shift_sigma"]})
    achieved = flops / (eval_ms * 1e-3) / 1e12
    per_t = []
    ess_tot = 0.0
    for i in range(len(Ts)):
        s = sums[:, i].sum(axis=0)
        w = weight_summary(s, el)
        ess_tot += w["ess_per_s"] * el
        per_t.append({"T": float(Ts[i]), "P": int(Ps[i]), "ess_fraction": w["ess_fraction"],
                      "z_ratio": w["z_ratio"], "z_ratio_err": w["z_ratio_err"]})
    beads = int(np.sum(Ps))
    line = {
        "metric": "effective samples/sec at fixed Z-ratio std error; learned ρ vs harmonic ρ",
        "value": ess_tot / el, "unit": "effective samples/s (harmonic rho, all temperatures)",
        "n_gpus": 1, "steps": K, "warmup": W, "ms_per_step": el / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.fp32 else "f64", "data": "synthetic (seeded model, device Philox samples)",
        "config": {"workload": f"A6_N12_T100-500: A={cfg['A']} N={cfg['N']}, 16 temperatures "
                               f"100-500 K, P_i = ceil(beta_i/tau0), tau0 = beta(100 K)/256 "
                               f"(P {int(Ps.min())}-{int(Ps.max())}), block={B} per temperature",
                   "samples_per_step": B * len(Ts), "schedule": tune["schedule"],
                   "value_from": "harmonic rho; block step (16 samplers, evaluation, 16 reductions) "
                                 "replayed as one hipGraph, evaluation schedule chosen by the autotuner"},
        "samples_beads_per_s": B * beads * K / el,
        "samples_per_s": B * len(Ts) * K / el,
        "batched_eval_ms": eval_ms,
        "tuned_batched_ms": tune["batched_ms"],
        "sequential_eval_ms": tune["sequential_ms"],
        "batched_speedup_vs_sequential": tune["sequential_ms"] / tune["batched_ms"],
        "temperatures": per_t,
        "fp32_error_vs_fp64": errors,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "peak_note": "78.6 TFLOP/s is also the non-packed FP32 VALU rate (packed FP32 "
                                  "is disabled in this build)",
                     "kernel": {"lanes": "k_eval_batched (persistent)", "wide": "k_eval_wide_batched (persistent)",
                                "sequential": "k_eval_fp64, one launch per temperature in the graph step"
                                }[tune["schedule"]["family"]],
                     "flops_per_launch": flops, "traffic_detail": "not profiled for this workload"},
        "cpu_baseline": None,
    }
    if cpu is not None:
        line["cpu_baseline"] = {
            "value": cpu["samples_beads_per_s"], "unit": "samples*beads/s", "cores": cpu["cores"],
            "kind": "port",
            "sample": f"oracle at the middle temperature only (T={cpu['T']:.1f} K, P={cpu['P']}), "
                      f"{cpu['samples']} samples in blocks of 500 over {cpu['seconds']:.1f} s, "
                      f"one process per core, numpy {cpu['numpy']}",
            "one_core_value": cpu["one_core"],
            "gpu_over_cpu": (B * beads * K / el) / cpu["samples_beads_per_s"],
        }
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--block", type=int, default=100_000)
    ap.add_argument("--lanes", type=int, default=0, help="lanes per sample (0 = automatic)")
    ap.add_argument("--pm", type=int, default=0, help="1: block_compute_pm (g at beta +- dbeta)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-cores", type=int, default=0, help="0 = min(16, os.cpu_count())")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--ring-bf16x3", action="store_true",
                    help="wide kernel with M and ring products as BF16x3 on the bf16 matrix cores "
                         "(PIB_EVAL_BF16X3); the error against FP64 on the same samples is reported")
    ap.add_argument("--wide", action="store_true",
                    help="harmonic leg: the one-wave-per-sample FP64 kernel (PIB_EVAL_WIDE) at any A <= 16")
    ap.add_argument("--fp32", action="store_true",
                    help="reduced-precision kernel (PIB_EVAL_FP32); the error against FP64 on the "
                         "same samples is reported")
    ap.add_argument("--workload", default="A4_N8_P64_T300")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--train-batch", type=int, default=2048)
    ap.add_argument("--flow-block", type=int, default=65536)
    ap.add_argument("--flow-layers", type=int, default=8)
    ap.add_argument("--flow-hidden", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--flow-gemm", default="bf16x3", choices=["f32", "bf16x3"],
                    help="precision of the flow's dense layers (FP32 MFMA or BF16x3 split)")
    ap.add_argument("--flow-fused", type=int, default=1,
                    help="0: layer-by-layer products in the draws instead of one launch per coupling layer")
    ap.add_argument("--flow-draw", default="bf16", choices=["bf16x3", "bf16"],
                    help="precision of the timed draws' one-launch layers (bf16: single bf16 products; "
                         "log q stays exact for the map applied); training is BF16x3 either way")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the timed region, copy a buffer of this many MiB 3 times "
                         "(calibrates profiler byte counters against a known transfer)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not "
                                   "supported (samples shard trivially; see DESIGN.md)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    import numpy as np
    from pibronic_amd import pimc, synthetic, _native
    name = args.workload
    cfg = synthetic.WORKLOADS[name]
    A, N, P, T = cfg["A"], cfg["N"], cfg["P"], cfg["T"]
    B = args.block
    pm = bool(args.pm)
    tmp = tempfile.mkdtemp(prefix="pib_bench_")
    FS = synthetic.write_workload(tmp, name)

    batched = P is None
    cpu = None
    if not args.no_cpu:
        cores = args.cpu_cores or min(16, os.cpu_count() or 1)
        if batched:
            Ts, Ps, _ = synthetic.c3_temperatures()
            T, P = float(Ts[8]), int(Ps[8])
        cpu = cpu_baseline(FS.path_vib_model, FS.path_rho_model, T, P, args.cpu_seconds, cores)
        cpu.update(cores=cores, T=T, P=P)
        log(f"[bench] cpu baseline {cpu['samples_beads_per_s']:.3e} samples*beads/s on {cores} cores")

    if batched:
        print(json.dumps(batched_leg(FS, cfg, args, cpu, log)))
        return

    import torch
    data = (pimc.BoxDataPM if pm else pimc.BoxData).from_FileStructure(FS)
    data.samples, data.beads, data.temperature = B * (args.steps + args.warmup), P, T
    data.block_size, data.blocks = B, args.steps + args.warmup
    data.seed = args.seed
    data.preprocess()
    ctx = data.context()
    if args.lanes:
        ctx.set_lanes_per_sample(args.lanes)
    lanes = ctx.lanes_per_sample(B)
    dev = data.device_buffers(B)
    stream = torch.cuda.current_stream()
    K, W = args.steps, args.warmup
    sums = torch.zeros((K + W, _native.NUM_SUMS), dtype=torch.float64, device="cuda")
    flags = ((_native.EVAL_PM if pm else 0) | (_native.EVAL_FP32 if args.fp32 else 0)
             | (_native.EVAL_BF16X3 if args.ring_bf16x3 else 0) | (_native.EVAL_WIDE if args.wide else 0))
    ev0 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    ev1 = [torch.cuda.Event(enable_timing=True) for _ in range(K)]

    def step(i, timed):
        ctx.sample_block(args.seed, i * B, B, dev["q"], None, stream)
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
    for i in range(W, W + K):
        step(i, True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    log(f"[bench] harmonic: {K} steps of {B} samples in {elapsed:.4f} s")

    if args.calib_copy_mib:
        nel = args.calib_copy_mib * (1 << 20) // 8
        src_t = torch.ones(nel, dtype=torch.float64, device="cuda")
        dst_t = torch.empty_like(src_t)
        for _ in range(3):
            dst_t.copy_(src_t)
        torch.cuda.synchronize()
        del src_t, dst_t

    harm = weight_summary(sums[W:].sum(dim=0).cpu().numpy(), elapsed)
    harm["ms_per_step"] = elapsed / K * 1e3
    n = harm["samples"]
    eval_ms = sum(a.elapsed_time(b) for a, b in zip(ev0, ev1)) / K
    nq = int(np.count_nonzero(np.abs(data.vib.quadratic).sum(axis=(2, 3))))
    fl = algorithmic_flops(A, N, P, nq, pm)
    achieved = fl["eval"] * B / (eval_ms * 1e-3) / 1e12
    traffic, tdetail = None, "profile only for the default workload"
    fp32_err = None
    if args.fp32 or args.ring_bf16x3:
        from pibronic_amd.precision import precision_report
        fp32_err = precision_report(data, dev["q"], B, pm=pm,
                                    variant="fp32" if args.fp32 else "bf16x3")
    if not pm and B == 100_000 and not (args.fp32 or args.ring_bf16x3):
        traffic, tdetail = profiled_traffic(
            os.path.join(ROOT, "profiles", f"k_eval_fp64_A{A}_N{N}_P{P}_B{B}.md"),
            f"void pib::k_eval_fp64<{A}, {A}, false, false>")

    learned = None
    if not args.no_learned and not pm and P % 64:
        # the flow's normal-mode whitening works in MFMA tiles of 64 beads
        log(f"[bench] learned leg skipped: P = {P} is not a multiple of 64")
    elif not args.no_learned and not pm:
        learned = learned_leg(data, args, log)
        log(f"[bench] learned: ess fraction {learned['ess_fraction']:.3e}, "
            f"{learned['ess_per_s']:.3e} ESS/s (harmonic {harm['ess_per_s']:.3e})")

    value = learned["ess_per_s"] if learned else harm["ess_per_s"]
    line = {
        "metric": "effective samples/sec at fixed Z-ratio std error; learned ρ vs harmonic ρ",
        "value": value,
        "unit": "effective samples/s" + ("" if learned else " (harmonic rho)"),
        "n_gpus": 1, "steps": K, "warmup": W,
        # the timed step of `value`: a learned draw + FP64 g + sums (or the harmonic step)
        "ms_per_step": (learned["seconds"] if learned else elapsed) / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.fp32 else ("f64 M, bf16x3/f32 ring" if args.ring_bf16x3 else "f64"),
        "data": "synthetic (seeded model, device Philox samples)",
        "config": {"workload": f"{name}: A={A} N={N} P={P} T={T:g}K block={B}"
                               + (" +-dbeta" if pm else ""),
                   "samples_per_step": learned["samples_per_step"] if learned else B,
                   "harmonic_block": B, "lanes_per_sample": lanes,
                   "value_from": ("learned proposal (ms_per_step: one draw of samples_per_step, FP64 g and "
                                  "sums); samples_beads_per_s, samples_per_s: the harmonic FP64 step")
                                 if learned else "harmonic rho",
                   "flow_draw": (f"{learned['draw_precision']} coupling-layer products in the timed draws "
                                 "(trained with BF16x3); g in FP64; log q exact for the map applied")
                                if learned else None},
        "samples_beads_per_s": n * P / elapsed,
        "samples_per_s": harm["samples_per_s"],
        "eval_kernel_ms": eval_ms,
        "harmonic": harm,
        "learned": learned,
        "learned_over_harmonic_ess_per_s": (learned["ess_per_s"] / harm["ess_per_s"]) if learned else None,
        "reduced_precision_error_vs_fp64": fp32_err,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "hbm_GBps": (traffic / (eval_ms * 1e-3) / 1e9) if (traffic and eval_ms) else None,
                     "hbm_frac": (traffic / (eval_ms * 1e-3) / HBM_PEAK_BPS) if (traffic and eval_ms) else None,
                     "kernel": ("k_eval_fp32" if args.fp32 else "k_eval_wide" if (A > 8 or args.ring_bf16x3 or args.wide)
                                else "k_eval_fp64"), "flops_per_sample": fl["eval"],
                     "algorithmic_bytes": B * N * P * 8 + B * (5 if pm else 3) * 8,
                     "traffic_detail": tdetail},
        "cpu_baseline": None,
    }
    if cpu is not None:
        line["cpu_baseline"] = {
            "value": cpu["samples_beads_per_s"], "unit": "samples*beads/s",
            "cores": cpu["cores"], "kind": "port",
            "sample": f"oracle (FP64 NumPy restatement, per-(sample,bead) dot loop) on the same "
                      f"model, {cpu['samples']} samples in blocks of 500 over {cpu['seconds']:.1f} s, "
                      f"one process per core, numpy {cpu['numpy']}",
            "one_core_value": cpu["one_core"],
            "effective_samples_per_s": cpu["samples_beads_per_s"] / P * cpu["ess_frac"],
            "gpu_over_cpu": (n * P / elapsed) / cpu["samples_beads_per_s"],
        }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
