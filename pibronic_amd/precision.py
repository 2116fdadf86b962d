"""Error of a reduced-precision kernel (FP32, or BF16x3 ring products) against the FP64
kernel, per result.

Both kernels run on the same bead coordinates; the report gives, for every per-sample output
and for the estimator weight r = g / rho, the largest and the median relative difference, the
largest difference relative to the mean |r| (what a block average sees), and the shift of the
block's Z estimate mean(r) in units of its standard error.
"""
import numpy as np

from . import _native


def _stats(a32, a64):
    a32, a64 = np.asarray(a32, dtype=np.float64), np.asarray(a64, dtype=np.float64)
    rel = np.abs(a32 - a64) / np.maximum(np.abs(a64), np.finfo(float).tiny)
    return {"max_rel": float(rel.max()), "median_rel": float(np.median(rel))}


def precision_report(data, q_dev, B, pm=None, variant="fp32"):
    """variant ("fp32": PIB_EVAL_FP32, "bf16x3": PIB_EVAL_BF16X3) vs the FP64 kernel of the
    same family on the device coordinates q_dev (B, N, P) of `data` (a BoxData)."""
    import torch
    if variant not in ("fp32", "bf16x3"):
        raise ValueError(variant)
    attr = "fp32" if variant == "fp32" else "ring_bf16x3"
    pm = data.pm if pm is None else pm
    dev = q_dev.device
    outs = {}
    saved = (data.fp32, data.ring_bf16x3, data.force_wide)
    try:
        data.fp32 = data.ring_bf16x3 = False
        if variant == "bf16x3":
            data.force_wide = True  # compare with the FP64 wide kernel
        for low in (False, True):
            setattr(data, attr, low)
            out = torch.empty((_native.NUM_OUT, B), dtype=torch.float64, device=dev)
            data.evaluate_device(q_dev, B, pm, out, B)
            outs[low] = out
        torch.cuda.synchronize(dev)
    finally:
        data.fp32, data.ring_bf16x3, data.force_wide = saved
    o64, o32 = outs[False].cpu().numpy(), outs[True].cpu().numpy()
    keys = [("s_rho", _native.OUT_RHO), ("s_g", _native.OUT_G)]
    if pm:
        keys += [("s_gP", _native.OUT_GP), ("s_gM", _native.OUT_GM)]
    rep = {k: _stats(o32[i], o64[i]) for k, i in keys}
    rep["log_scale_max_abs"] = float(np.max(np.abs(o32[_native.OUT_LOGS] - o64[_native.OUT_LOGS])))
    r64 = o64[_native.OUT_G] / o64[_native.OUT_RHO]
    r32 = o32[_native.OUT_G] / o32[_native.OUT_RHO]
    rep["r"] = _stats(r32, r64)
    scale = np.mean(np.abs(r64))
    rep["r"]["max_abs_over_mean_abs"] = float(np.max(np.abs(r32 - r64)) / scale)
    se = np.std(r64) / np.sqrt(max(B - 1, 1))
    rep["z_shift_sigma"] = float((np.mean(r32) - np.mean(r64)) / se) if se > 0 else 0.0
    return rep
