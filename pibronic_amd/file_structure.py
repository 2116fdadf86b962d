"""Directory layout and result-file names used by the sampling path.

Only what BoxData / BoxResult need: the data_set_{D}/rho_{R}/{parameters,results} layout
(pibronic/data/file_structure.py:32-138), model paths and hashes (:231-241) and the
result-file templates (pibronic/data/file_name.py:25-45, 98-135).
"""
import os
from os.path import join

from . import vibronic_model_io as vIO

coupled_model = "coupled_model.json"
harmonic_model = "harmonic_model.json"
sampling_model = "sampling_model.json"


def pimc(P="{P:d}", T="{T:.2f}", J="{J:d}"):
    """Result file of one job: P{P}_T{T:.2f}_J{J}_data_points.npz (file_name.py:31-45)."""
    return "P{:s}_T{:s}_J{:s}_data_points.npz".format(P, T, J)


def jackknife(P="{P:d}", T="{T:.2f}", X="{X:d}"):
    """Output of the jackknife analysis: P{P}_T{T:.2f}_X{X}_thermo (file_name.py:48-60)."""
    return "P{:s}_T{:s}_X{:s}_thermo".format(P, T, X)


def training_data_input(P="{P:d}", T="{T:.2f}", J="{J:d}"):
    """file_name.py:98-110"""
    return "P{:s}_T{:s}_J{:s}_training_data_input.npz".format(P, T, J)


def training_data_g_output(P="{P:d}", T="{T:.2f}", J="{J:d}"):
    """file_name.py:113-122"""
    return "P{:s}_T{:s}_J{:s}_training_data_g_output.npz".format(P, T, J)


def training_data_rho_output(P="{P:d}", T="{T:.2f}", J="{J:d}"):
    """file_name.py:125-135"""
    return "P{:s}_T{:s}_J{:s}_training_data_rho_output.npz".format(P, T, J)


class FileStructure:
    """Paths of one (data set, sampling distribution) pair under path_root."""

    template_vib = "data_set_{:d}/"
    template_rho = template_vib + "rho_{:d}/"

    @classmethod
    def from_boxdata(cls, path_root, data):
        return cls(path_root, data.id_data, data.id_rho)

    def __init__(self, path_root, id_data, id_rho=0, no_makedir=False):
        assert os.path.exists(path_root), f"the path_root ({path_root}) does not exist"
        assert os.path.isdir(path_root), f"the path_root ({path_root}) is not a directory"
        self.id_data, self.id_rho = id_data, id_rho
        self.path_root = os.path.abspath(path_root)
        vib = join(self.path_root, self.template_vib.format(id_data))
        rho = join(self.path_root, self.template_rho.format(id_data, id_rho))
        self.path_data, self.path_rho = vib, rho
        self.path_vib_params = join(vib, "parameters/")
        self.path_vib_results = join(vib, "results/")
        self.path_rho_params = join(rho, "parameters/")
        self.path_rho_results = join(rho, "results/")
        self.dir_list = ["path_vib_params", "path_vib_results", "path_rho_params",
                         "path_rho_results"]
        self.template_pimc = self.path_rho_results + pimc(J="{J:s}")
        self.template_jackknife = self.path_rho_results + jackknife()
        self.path_vib_model = join(self.path_vib_params, coupled_model)
        self.path_har_model = join(self.path_vib_params, harmonic_model)
        self.path_rho_model = join(self.path_rho_params, sampling_model)
        if not no_makedir:
            self.make_directories()

    def make_directories(self):
        for name in self.dir_list:
            os.makedirs(getattr(self, name), exist_ok=True)

    def generate_model_hashes(self, force_flag=False):
        """SHA-512 of both model files (file_structure.py:231-241)."""
        if hasattr(self, "hash_vib") and hasattr(self, "hash_rho") and not force_flag:
            return
        self.hash_vib = vIO.create_model_hash(FS=self)
        self.hash_rho = vIO.create_diagonal_model_hash(FS=self)
