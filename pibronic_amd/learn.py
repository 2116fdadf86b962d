"""Training and evaluation of the learned proposal against the harmonic rho.

On-policy adaptive importance sampling: each step draws fresh samples from the current flow
(which starts as exactly the harmonic rho), weights them by w = |g|/q with the FP64
evaluation kernel, and takes an Adam step on the self-normalised estimate of
KL(p || q) for p proportional to |g|. (Resampling a fixed weighted data set instead was
measured to overfit: the training likelihood rose while fresh-sample ESS collapsed.)

Evaluation draws fresh samples from both proposals and reports the Kish effective sample
fraction (sum w)^2 / (n sum w^2), the unbiased estimates of Z/Z_rho (mean signed weight) with
their standard errors, and the wall time per sample of each proposal on the device.
"""
import time

import numpy as np

from . import _native, flow


def _draw_rho(data, n, seed, offset, chunk):
    import torch
    Rs, lws, sgs = [], [], []
    for s in range(0, n, chunk):
        B = min(chunk, n - s)
        R = torch.empty((B, data.modes, data.beads), dtype=torch.float64, device="cuda")
        data.context().sample_block(seed, offset + s, B, R)
        lw, sg = flow.harmonic_log_weights(data, R)
        Rs.append(R), lws.append(lw), sgs.append(sg)
    return torch.cat(Rs), torch.cat(lws), torch.cat(sgs)


def _draw_flow(fl, n, seed, offset, chunk):
    import torch
    Rs, lws, sgs = [], [], []
    for s in range(0, n, chunk):
        B = min(chunk, n - s)
        R, lq = fl.draw(B, seed, offset + s)
        lw, sg = flow.log_weights(fl.data, R, lq)
        Rs.append(R), lws.append(lw), sgs.append(sg)
    return torch.cat(Rs), torch.cat(lws), torch.cat(sgs)


def weight_stats(logw, sign):
    """ESS fraction, Z estimate and its standard error from log|w| and sign(w)."""
    lw = logw.double().cpu().numpy()
    sg = sign.double().cpu().numpy()
    ok = np.isfinite(lw)
    lw, sg = lw[ok], sg[ok]
    n = lw.size
    M = lw.max()
    w = np.exp(lw - M)
    ess = w.sum() ** 2 / (w * w).sum()
    sw = sg * w
    z = np.exp(M) * sw.mean()
    err = np.exp(M) * sw.std() / np.sqrt(n - 1)
    return {"n": n, "ess_frac": ess / n, "z": z, "z_err": err, "neg_frac": float((sg < 0).mean()),
            "dropped": int((~ok).sum())}


def train(fl, steps, batch, lr, seed, alpha=1.0, warmup_steps=0, log=None, log_every=50):
    """On-policy training (adaptive importance sampling): every step draws a fresh batch from
    the current flow, weights it by w = |g|/q with the FP64 evaluation kernel, and takes an
    Adam step on -sum_b c_b log q(R_b) with c = w^alpha / sum w^alpha over the batch (the
    self-normalised gradient of KL(p || q)). Fresh samples every step leave no finite data
    set to overfit, and q = p is a fixed point (equal weights, E_q grad log q = 0). alpha
    ramps linearly from 0.5 to `alpha` over warmup_steps. Returns logged statistics."""
    import torch
    data = fl.data
    hist = []
    t0 = time.perf_counter()
    for step in range(steps):
        R, lq = fl.draw(batch, seed, step * batch)
        lw, sg = flow.log_weights(data, R, lq)
        a = alpha if step >= warmup_steps else 0.5 + (alpha - 0.5) * step / max(1, warmup_steps)
        lwf = torch.where(torch.isfinite(lw), lw, torch.full_like(lw, -torch.inf))
        c = torch.softmax(a * lwf, dim=0)
        loss = fl.loss_and_grad(R, c)
        fl.adam_step(lr)
        if step % log_every == 0 or step == steps - 1:
            st = weight_stats(lw, sg)
            st.update(step=step, alpha=a, loss=float(loss), seconds=time.perf_counter() - t0)
            hist.append(st)
            if log:
                log(st)
    return hist


def compare(fl, eval_samples, seed, chunk=16384):
    """Fresh samples from rho and from the flow: ESS fractions, Z estimates, time/sample."""
    import torch
    data = fl.data
    _draw_rho(data, chunk, seed + 1, 3 * 10 ** 9, chunk)      # warm-up of both paths
    _draw_flow(fl, chunk, seed + 2, 4 * 10 ** 9, chunk)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, lw, sg = _draw_rho(data, eval_samples, seed + 7919, 10 ** 9, chunk)
    torch.cuda.synchronize()
    t_rho = time.perf_counter() - t0
    rho = weight_stats(lw, sg)
    t0 = time.perf_counter()
    _, lw, sg = _draw_flow(fl, eval_samples, seed + 104729, 2 * 10 ** 9, chunk)
    torch.cuda.synchronize()
    t_flow = time.perf_counter() - t0
    fls = weight_stats(lw, sg)
    return {"ess_frac_rho": rho["ess_frac"], "ess_frac_flow": fls["ess_frac"],
            "z_rho": rho["z"], "z_rho_err": rho["z_err"], "z_flow": fls["z"], "z_flow_err": fls["z_err"],
            "neg_frac_rho": rho["neg_frac"], "neg_frac_flow": fls["neg_frac"],
            "samples_per_s_rho": eval_samples / t_rho, "samples_per_s_flow": eval_samples / t_flow,
            "ess_per_s_rho": rho["ess_frac"] * eval_samples / t_rho,
            "ess_per_s_flow": fls["ess_frac"] * eval_samples / t_flow}


def train_and_compare(data, layers=8, hidden=256, steps=1000, batch=2048, eval_samples=2 ** 18,
                      lr=1e-3, seed=0, smax=2.0, alpha=1.0, warmup_steps=0, log=None):
    fl = flow.Flow(data, layers=layers, hidden=hidden, smax=smax, seed=seed)
    hist = train(fl, steps, batch, lr, seed, alpha=alpha, warmup_steps=warmup_steps, log=log)
    out = compare(fl, eval_samples, seed)
    out["train"] = hist
    out["flow"] = fl
    return out
