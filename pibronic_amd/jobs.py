"""Job farm on one GPU.

The reference splits a run of X samples into jobs of at most MAX_SAMPLES_PER_JOB samples and
submits one SLURM job per (temperature, beads, job index) (PimcSubmissionClass,
pibronic/server/job_boss.py:493-611). Each job runs the job script
pibronic/server/nlogn_feynman/pimc.py:15-37: it rebuilds a BoxDataPM from the JSON parameter
string (',' written as ';'), runs block_compute_pm and saves
P{P}_T{T:.2f}_J{J}_data_points.npz, which the analysis later merges
(BoxResult.load_multiple_results).

Here the same split, parameter strings and job entry point run the jobs in this process, one
after another on the GPU (the scheduler, partitions and sbatch wrappers are not part of the
sampling path). Jobs are independent streams: with no seed every job takes its own seed
from OS entropy, as each reference job process reseeds numpy (pibronic/pimc/pimc.py:55);
with one explicit seed job J draws global samples [J X, (J + 1) X) of that seed's stream, so
the merged jobs are one run split at those offsets.
"""
import copy

from . import file_structure
from . import pimc
from . import vibronic_model_io as vIO


def run_job(input_parameters, path_results, id_job, pm=True, seed=None, device=0):
    """One job: the body of pibronic/server/nlogn_feynman/pimc.py:15-37. Returns the result."""
    data = (pimc.BoxDataPM if pm else pimc.BoxData).from_json_string(input_parameters)
    data.seed = seed
    data.device = device
    data.preprocess()
    result = (pimc.BoxResultPM if pm else pimc.BoxResult)(data=data)
    result.path_root = path_results
    result.id_job = id_job
    (pimc.block_compute_pm if pm else pimc.block_compute)(data, result)
    return result


class PimcSubmissionClass:
    """Split a run into jobs and run them (pibronic/server/job_boss.py:493-611)."""

    MAX_SAMPLES_PER_JOB = int(1e5)

    def __init__(self, FS, param_dict=None):
        self.FS = FS
        self.param_dict = {
            "temperature_list": [0.0], "bead_list": [0],
            "number_of_samples": 0, "number_of_samples_overall": 0,
            "number_of_samples_per_job": 0, "number_of_blocks": 0,
            "number_of_states": 0, "number_of_modes": 0, "number_of_beads": 0,
            "block_size": int(1e3), "id_data": FS.id_data, "id_rho": FS.id_rho,
        }
        if param_dict is not None:
            self.param_dict.update(param_dict)
        self.param_dict["path_root"] = FS.path_root
        if not self.param_dict["number_of_states"] or not self.param_dict["number_of_modes"]:
            A, N = vIO.extract_dimensions_of_model(FS)
            self.param_dict["number_of_states"], self.param_dict["number_of_modes"] = A, N
        self.n_jobs = 0

    def setup_blocks_and_jobs(self):
        """Samples and blocks per job and the number of jobs (job_boss.py:538-566)."""
        pd = self.param_dict
        n_samples, block_size = pd["number_of_samples"], pd["block_size"]
        assert isinstance(n_samples, int) and isinstance(block_size, int)
        assert block_size <= n_samples, \
            f"block size {block_size} must be less than or equal to the number of samples {n_samples}"
        assert block_size <= self.MAX_SAMPLES_PER_JOB, \
            f"block size {block_size} must be less than or equal to the maximum number of samples per job {self.MAX_SAMPLES_PER_JOB}"
        per_job = min(n_samples, self.MAX_SAMPLES_PER_JOB)
        pd["number_of_samples_per_job"] = per_job
        pd["number_of_samples"] = per_job
        pd["number_of_blocks"] = per_job // block_size
        if pd["number_of_samples_overall"] == 0:
            pd["number_of_samples_overall"] = n_samples
        self.n_jobs = max(1, pd["number_of_samples_overall"] // self.MAX_SAMPLES_PER_JOB)

    def job_parameters(self):
        """(T, P, job index, JSON parameter string) for every job (job_boss.py:568-611)."""
        self.setup_blocks_and_jobs()
        jobs = []
        for T in self.param_dict["temperature_list"]:
            for P in self.param_dict["bead_list"]:
                params = copy.deepcopy(self.param_dict)
                params["number_of_beads"] = P
                params["temperature"] = T
                keys = {k: params[k] for k in (
                    "number_of_samples", "number_of_blocks", "number_of_states", "number_of_beads",
                    "number_of_modes", "temperature", "block_size", "id_data", "id_rho", "path_root")}
                if "delta_beta" in params:
                    keys["delta_beta"] = params["delta_beta"]
                string = pimc.BoxData.json_serialize(keys)
                for J in range(self.n_jobs):
                    jobs.append((T, P, J, string))
        return jobs

    def submit_jobs(self, pm=True, seed=None, device=0):
        """Run every job in turn on `device`; returns the result file paths."""
        paths = []
        for T, P, J, string in self.job_parameters():
            res = run_job(string, self.FS.path_rho_results, J, pm=pm, seed=seed, device=device)
            paths.append(res.compute_path_to_file())
        return paths


def result_paths(FS, P, T):
    """Every job file of one (P, T) under FS's results directory, in job order."""
    import glob
    import os
    pattern = os.path.join(FS.path_rho_results, file_structure.pimc(P=str(P), T=f"{T:.2f}", J="*"))
    found = glob.glob(pattern)
    return sorted(found, key=lambda p: int(os.path.basename(p).split("_J")[1].split("_")[0]))
