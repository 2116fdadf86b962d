"""Shared set-up for the learned-proposal tests."""
import numpy as np

from oracle.flow_oracle import FlowRef


def ref_for(data, layers, hidden, smax):
    rho = data.rho
    return FlowRef(data.circulant_eigvects, rho.standard_deviation[0],
                   rho.state_weight @ rho.state_shift, rho.const.cothAN[0], rho.const.cschAN[0],
                   rho.const.omatrix_prefactor, rho.state_shift, layers, hidden, smax)


def random_params(ref, seed, scale=0.05):
    """A flat parameter vector with every weight non-zero (so tests exercise all paths)."""
    rng = np.random.default_rng(seed)
    H, h, D = ref.H, ref.h, ref.D
    n = ref.L * (H * h + h + h * h + h + h * D + D)
    flat = rng.normal(0.0, scale, n)
    off = 0
    for _ in range(ref.L):
        flat[off:off + H * h] *= np.sqrt(2.0 / H) / scale
        off += H * h + h
        flat[off:off + h * h] *= np.sqrt(2.0 / h) / scale
        off += h * h + h + h * D + D
    return flat.astype(np.float32)
