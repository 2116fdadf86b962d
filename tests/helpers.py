"""Shared builders for the tests: BoxData / oracle Setup pairs on the same model files."""
import os
import shutil

import numpy as np

from oracle import pimc_oracle as po
from pibronic_amd import pimc, synthetic
from pibronic_amd import vibronic_model_io as vIO
from pibronic_amd.file_structure import FileStructure
from pibronic_amd.vibronic_model_keys import VibronicModelKeys as VMK


def stage_models(root, vib_json, rho_json, id_data=0, id_rho=0):
    FS = FileStructure(str(root), id_data, id_rho)
    shutil.copy(vib_json, FS.path_vib_model)
    shutil.copy(rho_json, FS.path_rho_model)
    FS.generate_model_hashes()
    return FS


def stage_dict_models(root, coupled, sampling=None, id_data=0, id_rho=0):
    """Write a coupled model dict (and a sampling dict, default its diagonal)."""
    FS = FileStructure(str(root), id_data, id_rho)
    vIO.save_model_to_JSON(FS.path_vib_model, coupled)
    if sampling is None:
        vIO.remove_coupling_from_model(FS.path_vib_model, FS.path_rho_model)
    else:
        vIO.save_diagonal_model_to_JSON(FS.path_rho_model, sampling)
    FS.generate_model_hashes()
    return FS


def make_data(FS, T, P, samples, block, pm=False, seed=1234):
    data = pimc.BoxDataPM() if pm else pimc.BoxData()
    data = type(data).from_FileStructure(FS)
    data.samples, data.beads, data.temperature = samples, P, T
    data.block_size = block
    data.blocks = samples // block
    data.hash_vib, data.hash_rho = FS.hash_vib, FS.hash_rho
    data.seed = seed
    data.preprocess()
    return data


def oracle_setup(FS, T, P, beta=None):
    A, N, E, w, L, Q = po.read_model_json(FS.path_vib_model)
    Ar, Nr, Er, wr, Lr, _ = po.read_model_json(FS.path_rho_model, diagonal=True)
    return po.Setup((E, w, L, Q), (Er, wr, Lr), T, P, beta=beta)


def random_model(A, N, seed, quad="diag", off=0.05):
    """Small random coupled model dict; quad in {None, 'diag', 'full'}."""
    m = synthetic.generate(A, N, quadratic=(quad is not None), off=off, seed=seed)
    if quad == "full":
        rng = np.random.default_rng(seed + 1000)
        Q = rng.normal(0.0, 0.004, (N, N, A, A))
        Q = 0.5 * (Q + Q.transpose(0, 1, 3, 2))
        m[VMK.G2] = Q
    return m


def rel_err(a, b, floor=0.0):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))
