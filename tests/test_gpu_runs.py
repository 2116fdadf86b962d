"""Whole runs on the GPU: independent job streams and the result-file -> analysis chain.

* Jobs (pibronic/server/job_boss.py:538-611 with the job script
  pibronic/server/nlogn_feynman/pimc.py:15-37): jobs that share an explicit seed draw
  disjoint slices of one stream, so merging them equals one run split at the same offsets;
  jobs without a seed take their own seeds from OS entropy (the reference reseeds numpy in
  every job process, pibronic/pimc/pimc.py:55) and write them to their result files.
* Device block_compute_pm over two job files -> P{P}_T{T}_J{J}_data_points.npz
  (pibronic/pimc/pimc.py:948-964) -> jackknife_analysis_of_pimc -> *_thermo
  (pibronic/stats/stats.py:336-377), against the oracle's jackknife on the same host-seeded
  samples.
"""
import json
import os

import numpy as np
import pytest

from helpers import oracle_setup
from oracle import pimc_oracle as po
from oracle import stats_oracle as so
from pibronic_amd import jobs, pimc, stats, synthetic

pytestmark = pytest.mark.gpu

T, P = 300.0, 12


def _announce(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _submission(tmp_path, samples, block):
    FS = synthetic.write_workload(str(tmp_path), "A2_N3_P12_T300")
    return FS, jobs.PimcSubmissionClass(FS, {"number_of_samples": samples, "block_size": block,
                                             "temperature_list": [T], "bead_list": [P]})


def test_jobs_with_one_seed_are_one_split_run(tmp_path):
    _announce("jobs: one seed, two jobs vs one run")
    FS, sub = _submission(tmp_path, 8_000, 2_000)
    sub.MAX_SAMPLES_PER_JOB = 4_000
    paths = sub.submit_jobs(pm=True, seed=99)
    assert len(paths) == 2
    parts = [np.load(p) for p in paths]
    assert all(int(d["seed"]) == 99 for d in parts)
    assert [int(d["sample_offset"]) for d in parts] == [0, 4_000]
    assert not np.any(np.isin(parts[0]["s_g"], parts[1]["s_g"]))
    data = pimc.BoxDataPM.from_FileStructure(FS)
    data.samples, data.beads, data.temperature, data.block_size, data.blocks = 8_000, P, T, 2_000, 4
    data.hash_vib, data.hash_rho, data.seed = FS.hash_vib, FS.hash_rho, 99
    data.preprocess()
    whole = pimc.block_compute_pm(data, pimc.BoxResultPM(data=data))
    merged = pimc.BoxResultPM()
    merged.load_multiple_results(paths)
    for a, b in (("scaled_g", "scaled_g"), ("scaled_rho", "scaled_rho"),
                 ("scaled_gofr_plus", "scaled_gofr_plus"), ("scaled_gofr_minus", "scaled_gofr_minus")):
        assert np.array_equal(getattr(merged, a), getattr(whole, b)), a


def test_jobs_without_seed_are_independent(tmp_path):
    _announce("jobs: no seed, two jobs")
    FS, sub = _submission(tmp_path, 4_000, 2_000)
    sub.MAX_SAMPLES_PER_JOB = 2_000
    paths = sub.submit_jobs(pm=False, seed=None)
    d0, d1 = np.load(paths[0]), np.load(paths[1])
    assert int(d0["seed"]) != int(d1["seed"])
    assert not np.any(np.isin(d0["s_g"], d1["s_g"]))


def test_results_to_thermo_file_matches_oracle(tmp_path):
    """Two jobs of host-seeded samples through the device, the npz files, the jackknife
    analysis and its *_thermo file; the same samples through the oracle."""
    _announce("end to end: block_compute_pm -> npz -> jackknife -> thermo")
    FS = synthetic.write_workload(str(tmp_path), "A2_N3_P12_T300")
    X, B, seed = 4_000, 1_000, 5
    string = pimc.BoxData.json_serialize({
        "number_of_samples": X, "number_of_blocks": X // B, "number_of_states": 2,
        "number_of_beads": P, "number_of_modes": 3, "temperature": T, "block_size": B,
        "id_data": FS.id_data, "id_rho": FS.id_rho, "path_root": FS.path_root})
    s = oracle_setup(FS, T, P)
    ref = {k: [] for k in ("s_rho", "s_g", "s_gP", "s_gM")}
    for J in range(2):
        data = pimc.BoxDataPM.from_json_string(string)
        data.sampler, data.seed = "host", seed
        data.preprocess()
        res = pimc.BoxResultPM(data=data)
        res.path_root, res.id_job = FS.path_rho_results, J
        pimc.block_compute_pm(data, res)
        for b in range(X // B):
            q, _ = data.draw_host_sample(np.random.default_rng([seed, J * X + b * B]), B)
            o = po.evaluate(s, q, pm=True)
            for k in ref:
                ref[k].append(o[k])
    written = stats.jackknife_analysis_of_pimc(FS, data.rho)
    assert len(written) == 1 and os.path.basename(written[0]) == f"P{P}_T{T:.2f}_X{2 * X}_thermo"
    with open(written[0]) as fh:
        thermo = json.load(fh)
    an = stats.analytic_sampling(data.rho, data.beta)
    # 1) the analysis of the files equals the oracle's jackknife of the same per-sample arrays
    merged = pimc.BoxResultPM()
    merged.load_multiple_results(jobs.result_paths(FS, P, T))
    mine = so.basic_jackknife(T, data.delta_beta, merged.scaled_rho, merged.scaled_g,
                              merged.scaled_gofr_plus, merged.scaled_gofr_minus, an["E"], an["Cv"])
    for k, v in mine.items():
        assert abs(thermo[k] - v) <= 1e-12 * abs(v) + 1e-300, (k, thermo[k], v)
    # 2) ... and the oracle's own evaluation of the same samples, up to the per-sample parity
    #    (1e-10) propagated through the finite differences of d1 and d2
    o = {k: np.concatenate(v) for k, v in ref.items()}
    full = so.basic_jackknife(T, data.delta_beta, o["s_rho"], o["s_g"], o["s_gP"], o["s_gM"],
                              an["E"], an["Cv"])
    tol = {"Z": 1e-10, "Z error": 1e-8, "E": 1e-6, "jk_E": 1e-6, "jk_E error": 1e-4,
           "Cv": 1e-2, "jk_Cv": 1e-2, "jk_Cv error": 1e-2}
    for k, t in tol.items():
        err = abs(thermo[k] - full[k]) / abs(full[k])
        print(f"{k}: device {thermo[k]:.12g} oracle {full[k]:.12g} rel {err:.2e}", flush=True)
        assert err < t, (k, thermo[k], full[k])
