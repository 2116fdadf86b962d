"""Seeded synthetic vibronic models for the benchmark workloads.

Draw order with numpy.random.default_rng(seed), all energies in eV:
  1. w_j ~ U[0.02, 0.3], j = 0..N-1
  2. E_aa ~ U[0, 0.5], a = 0..A-1
  3. E_ab ~ N(0, off) for a < b (row-major order), mirrored
  4. L^j_ab ~ N(0, 0.05) on the diagonal and N(0, off) for a < b, for j = 0..N-1 and
     (a, b) in row-major upper-triangle order (diagonal included), mirrored
  5. (quadratic=True) Q^jj_ab ~ N(0, 0.005) for a <= b, same order, mirrored; Q^jk = 0
     for j != k
with off = 0.05, or 0.1 for the near-degenerate large-surface workload. Then, per mode j,
the quadratic block is made stable: if the smallest eigenvalue of w_j I + Q^jj is below
w_j / 2, the eigenvalues of Q^jj are raised to -w_j / 2 (same eigenvectors). As drawn, the
A=8 and A=12 workloads have modes with w_j I + Q^jj indefinite, i.e. a potential unbounded
below and a divergent partition function; the clip leaves the A=2 and A=4 models unchanged.
The sampling model rho is the surface diagonal of the coupled model
(remove_coupling_from_model semantics, pibronic/vibronic/vibronic_model_io.py:749-761).

surface_diagonal=True draws the same numbers and then zeroes every coupling between
surfaces (E_ab, L^j_ab, Q^jj_ab for a != b) before the stability clip: the surfaces become
independent displaced, mode-softened oscillators, g(R) is a sum of Gaussians and Z_g/Z_rho has
a closed form, while the mode-diagonal Q draws keep rho's tails as light relative to g as in
the coupled workload.

WORKLOADS names each configuration by its sizes.
"""
import os

import numpy as np

from . import vibronic_model_io as vIO
from .vibronic_model_keys import VibronicModelKeys as VMK

WORKLOADS = {
    # A=2 surfaces, N=3 modes, P=12 beads, 300 K, linear couplings only; 1e5 samples / 1e4
    "A2_N3_P12_T300": dict(A=2, N=3, P=12, T=300.0, quadratic=False, off=0.05,
                           samples=100_000, block=10_000),
    # A=4, N=8, P=64, 300 K, linear + mode-diagonal quadratic; 1e6 samples in blocks of 1e5
    "A4_N8_P64_T300": dict(A=4, N=8, P=64, T=300.0, quadratic=True, off=0.05,
                           samples=1_000_000, block=100_000),
    # A=6, N=12, linear + diagonal quadratic; temperatures 100-500 K with uneven P
    "A6_N12_T100-500": dict(A=6, N=12, P=None, T=None, quadratic=True, off=0.05,
                            samples=1_000_000, block=100_000),
    # A=12 (padded to 16 on MFMA), N=24, P=128, 250 K, off-diagonal scale 0.1 eV
    "A12_N24_P128_T250": dict(A=12, N=24, P=128, T=250.0, quadratic=True, off=0.1,
                              samples=4_000_000, block=100_000),
    # A=8, N=16, P=64, 300 K (learned-proposal workload)
    "A8_N16_P64_T300": dict(A=8, N=16, P=64, T=300.0, quadratic=True, off=0.05,
                            samples=10_000_000, block=100_000),
    # the A4 and A8 workloads with every inter-surface coupling removed (exact Z_g/Z_rho)
    "A4_N8_P64_T300_diag": dict(A=4, N=8, P=64, T=300.0, quadratic=True, off=0.05,
                                samples=1_000_000, block=100_000, surface_diagonal=True),
    "A8_N16_P64_T300_diag": dict(A=8, N=16, P=64, T=300.0, quadratic=True, off=0.05,
                                 samples=10_000_000, block=100_000, surface_diagonal=True),
}


def generate(A, N, quadratic=True, off=0.05, seed=0, surface_diagonal=False):
    """Return the coupled model as a dict keyed by VibronicModelKeys members."""
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.02, 0.3, N)
    E = np.diag(rng.uniform(0.0, 0.5, A))
    for a in range(A):
        for b in range(a + 1, A):
            E[a, b] = E[b, a] = rng.normal(0.0, off)
    L = np.zeros((N, A, A))
    for j in range(N):
        for a in range(A):
            for b in range(a, A):
                L[j, a, b] = L[j, b, a] = rng.normal(0.0, 0.05 if a == b else off)
    Q = np.zeros((N, N, A, A))
    if quadratic:
        for j in range(N):
            for a in range(A):
                for b in range(a, A):
                    Q[j, j, a, b] = Q[j, j, b, a] = rng.normal(0.0, 0.005)
    if surface_diagonal:
        keep = np.eye(A, dtype=bool)
        E = np.where(keep, E, 0.0)
        L = np.where(keep, L, 0.0)
        Q = np.where(keep, Q, 0.0)
    if quadratic:
        for j in range(N):
            lam, vec = np.linalg.eigh(Q[j, j])
            if lam.min() < -0.5 * w[j]:
                Q[j, j] = (vec * np.maximum(lam, -0.5 * w[j])) @ vec.T
                Q[j, j] = 0.5 * (Q[j, j] + Q[j, j].T)
    return {VMK.N: N, VMK.A: A, VMK.E: E, VMK.w: w, VMK.G1: L, VMK.G2: Q}


def diagonal_of(model):
    """The sampling model (surface diagonal) of a coupled model dict."""
    d = {VMK.N: model[VMK.N], VMK.A: model[VMK.A], VMK.w: model[VMK.w].copy()}
    for k in (VMK.E, VMK.G1, VMK.G2):
        v = model[k]
        d[k] = np.diagonal(v, axis1=v.ndim - 2, axis2=v.ndim - 1).copy()
    return d


def write_workload(root, name, id_data=0, id_rho=0, seed=0):
    """Write coupled_model.json and sampling_model.json of a workload under
    root/data_set_{id_data}[/rho_{id_rho}]/parameters and return the FileStructure."""
    from .file_structure import FileStructure
    cfg = WORKLOADS[name]
    FS = FileStructure(root, id_data, id_rho)
    m = generate(cfg["A"], cfg["N"], cfg["quadratic"], cfg["off"], seed,
                 surface_diagonal=cfg.get("surface_diagonal", False))
    vIO.save_model_to_JSON(FS.path_vib_model, m)
    vIO.remove_coupling_from_model(FS.path_vib_model, FS.path_rho_model)
    FS.generate_model_hashes()
    return FS


def c3_temperatures(tau0=None):
    """Temperatures T_i = 100 * 5^(i/15) K (i = 0..15) and beads P_i = ceil(beta_i / tau0).

    A fixed tau0 cannot give P from 16 to 256 over 100-500 K (beta only varies 5x); tau0 is
    taken as beta(100 K)/256 so that the coldest point has P = 256 and the hottest P = 52.
    """
    from . import constants
    Ts = 100.0 * 5.0 ** (np.arange(16) / 15.0)
    betas = constants.beta(Ts)
    if tau0 is None:
        tau0 = constants.beta(100.0) / 256.0
    Ps = np.ceil(betas / tau0 - 1e-9).astype(int)
    return Ts, Ps, tau0


def model_arrays(path_vib, path_rho):
    """(E, w, L, Q) of a coupled file and (E, w, L) of a sampling file."""
    c = vIO.load_model_from_JSON(path_vib)
    A, N = vIO._extract_dimensions_from_dictionary(c)
    shp = vIO.model_shape_dict(A, N)
    L = c.get(VMK.G1, np.zeros(shp[VMK.G1]))
    Q = c.get(VMK.G2, np.zeros(shp[VMK.G2]))
    r = vIO.load_diagonal_model_from_JSON(path_rho)
    Ar, _ = vIO._extract_dimensions_from_dictionary(r)
    Lr = r.get(VMK.G1, np.zeros((N, Ar)))
    return (c[VMK.E], c[VMK.w], L, Q), (r[VMK.E], r[VMK.w], Lr)


