"""Training-data hooks on the GPU against the oracle's restatement (quirk Q6 included):
block_compute_rhoR_from_input_samples and block_compute_gR_from_raw_samples
(pibronic/pimc/pimc.py:1273-1338). Tolerance as for the FP64 path: 1e-10 relative."""
import os

import numpy as np
import pytest

from helpers import make_data, oracle_setup, random_model, rel_err, stage_dict_models
from oracle import pimc_oracle as po
from pibronic_amd import pimc
from pibronic_amd.vibronic_model_keys import VibronicModelKeys as VMK

pytestmark = pytest.mark.gpu


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _symmetric_shifted(A, N, seed):
    m = random_model(A, N, seed=seed, quad="full")
    m[VMK.G2] = 0.5 * (m[VMK.G2] + m[VMK.G2].transpose(1, 0, 2, 3))
    L = m[VMK.G1].copy()
    for j in range(N):
        L[j] = np.diag(np.diag(L[j]))
    m[VMK.G1] = L
    return m


@pytest.mark.parametrize("A", [3, 10])
def test_rho_hook(tmp_path, A):
    _announce(f"rho(R) hook A={A}")
    FS = stage_dict_models(tmp_path, random_model(A, 3, seed=A))
    P, X = 9, 300
    data = make_data(FS, 300.0, P, X, 100)
    res = pimc.BoxResult(data=data)
    R = np.random.default_rng(3).normal(0.0, 1.5, (X, 3, P))
    pimc.block_compute_rhoR_from_input_samples(data, res, R)
    want = po.rho_of_R_hook(oracle_setup(FS, 300.0, P), R)
    assert rel_err(res.scaled_rho, want) < 1e-10


@pytest.mark.parametrize("A", [3, 9])
def test_g_hook(tmp_path, A):
    _announce(f"g(R) hook A={A}")
    FS = stage_dict_models(tmp_path, _symmetric_shifted(A, 3, seed=A))
    P, X = 8, 256
    data = make_data(FS, 300.0, P, X, 128, seed=9)
    res = pimc.BoxResult(data=data)
    R = pimc.block_compute_gR_from_raw_samples(data, res, rng=np.random.default_rng(11))
    assert R.shape == (X, 3, P) and R.min() >= 0.0 and R.max() < 1.0
    want = po.g_of_R_hook(oracle_setup(FS, 300.0, P), R)
    assert rel_err(res.scaled_g, want) < 1e-10


def test_g_hook_rejects_nonsymmetric_V(tmp_path):
    _announce("g(R) hook asymmetric V")
    FS = stage_dict_models(tmp_path, random_model(3, 2, seed=5))
    data = make_data(FS, 300.0, 6, 32, 32)
    res = pimc.BoxResult(data=data)
    with pytest.raises(AssertionError, match="not symmetric"):
        pimc.block_compute_gR_from_raw_samples(data, res, rng=np.random.default_rng(0))


def test_g_hook_writes_training_files(tmp_path):
    _announce("g(R) hook files")
    FS = stage_dict_models(tmp_path, _symmetric_shifted(2, 2, seed=1))
    data = make_data(FS, 250.0, 6, 64, 64)
    res = pimc.BoxResult(data=data)
    res.path_root = str(tmp_path)
    R = pimc.block_compute_gR_from_raw_samples(data, res)
    g_file = os.path.join(tmp_path, "P6_T250.00_J0_training_data_g_output.npz")
    r_file = os.path.join(tmp_path, "P6_T250.00_J0_training_data_input.npz")
    with np.load(g_file) as d:
        assert int(d["number_of_samples"]) == 64
        assert np.array_equal(d["g"], res.scaled_g)
    with np.load(r_file) as d:
        assert np.array_equal(d["input_R_values"], R)


def test_rho_hook_other_surface_count_is_not_supported(tmp_path):
    _announce("rho(R) hook A_rho != A")
    m = random_model(2, 2, seed=2)
    rho = {VMK.N: 2, VMK.A: 3, VMK.w: m[VMK.w].copy(), VMK.E: np.array([0.1, 0.2, 0.3]),
           VMK.G1: np.zeros((2, 3))}
    FS = stage_dict_models(tmp_path, m, rho)
    data = make_data(FS, 300.0, 5, 8, 8)
    with pytest.raises(NotImplementedError, match="qTempTensor"):
        pimc.block_compute_rhoR_from_input_samples(data, pimc.BoxResult(data=data), np.zeros((8, 2, 5)))
