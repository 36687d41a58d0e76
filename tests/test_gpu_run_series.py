"""The reference's ZrO run-params series end to end on the GPU:
states_from_run_params -> MonteCalculator.run_multistate with the fixture's
completion checks (at most 100 passes, potential energy to 0.001 eV at 95 %
confidence, checked every 10 passes) -> summary.json and completed_runs.json.

* summary.json passes a restatement of the reference's own layout checks
  (python/tests/conftest.py:142-259: keys of conditions, analysis,
  completion_check_results and statistics, one entry per run, is_converged of
  the requested quantity), semi-grand canonical and canonical;
* restart: a 5-state series is extended to the 11-state series in the same
  directory, which runs the 6 missing states only and keeps the first 5
  entries; a third call runs nothing;
* completion: every run stops at a check point (a multiple of 10 passes) and
  at most at the 100-pass cutoff; a run reported converged has a calculated
  precision of at most 0.001;
* results_io (write_observations, write_trajectory) with the
  clex.formation_energy.sparse_corr sampler: <label>/run.<i>/observations.json
  and trajectory.json per run (the files the reference's full-run tests
  expect, tests/unit/clexmonte/canonical_fullrun_test.cpp:348-357), one row
  per sample of every sampler; the sparse correlations are named by the ECI
  file's linear function indices of the non-zero ECIs, and the file ECIs
  dotted with them give each sample's formation energy (1e-9); every
  trajectory configuration's oracle energy equals its sample's formation
  energy within the FP32 accumulation bound of the table path.
"""
import copy
import json
import os

import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.calculator import MonteCalculator, System, run_series
from casmcode_clexmonte_amd.casm_io import clex_from_casm

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ZRO = json.load(open(os.path.join(ROOT, "tests", "golden", "casm_clex_fixtures.json")))["Clex_ZrO_Occ"]


def _print_case(name):
    print(f"[case] {name}", flush=True)


@pytest.fixture(scope="module")
def zro():
    return System(clex_from_casm(ZRO["prim"], ZRO["formation_energy_eci"]))


def _validate_summary_data(sub, keys, size):
    for x in keys:
        assert x in sub, x
        if "component_names" in sub[x]:
            for y in sub[x]["component_names"]:
                assert len(sub[x][y]) == size
        elif "value" in sub[x]:
            assert sub[x]["shape"] == []
            assert len(sub[x]["value"]) == size
        else:
            assert len(sub[x]) == size


def _validate_statistics(sub, keys, size):
    for x in keys:
        assert x in sub
        if "component_names" in sub[x]:
            for y in sub[x]["component_names"]:
                for z in ("mean", "calculated_precision"):
                    assert len(sub[x][y][z]) == size
        else:
            assert sub[x]["shape"] == []
            for z in ("mean", "calculated_precision"):
                assert len(sub[x]["value"][z]) == size


def _validate_summary_file(path, size, canonical):
    with open(path) as f:
        data = json.load(f)
    if canonical:
        cond_keys, an_keys = ["temperature", "param_composition", "mol_composition"], ["heat_capacity"]
    else:
        cond_keys = ["temperature", "param_chem_pot"]
        an_keys = ["heat_capacity", "mol_susc", "param_susc", "mol_thermochem_susc", "param_thermochem_susc"]
    _validate_summary_data(data["analysis"], an_keys, size)
    _validate_summary_data(data["completion_check_results"],
                           ["N_samples", "N_samples_for_statistics", "acceptance_rate", "count", "elapsed_clocktime",
                            "N_samples_for_all_to_equilibrate", "all_converged", "all_equilibrated"], size)
    _validate_summary_data(data["conditions"], cond_keys, size)
    _validate_statistics(data["statistics"],
                         ["clex.formation_energy", "mol_composition", "param_composition", "potential_energy"], size)
    assert "is_converged" in data["statistics"]["potential_energy"]["value"]
    return data


def _check_completion(data):
    cc = data["completion_check_results"]
    pe = data["statistics"]["potential_energy"]["value"]
    for n_s, conv, prec in zip(cc["N_samples"], cc["all_converged"], pe["calculated_precision"]):
        assert 10 <= n_s <= 100 and n_s % 10 == 0, n_s
        if conv:
            assert prec <= 0.001


def test_sgc_series_files_and_restart(zro, tmp_path):
    _print_case("ZrO semi-grand canonical series: summary.json, completed_runs.json, restart")
    rp = ZRO["run_params_sgc_complete"]
    rp5 = copy.deepcopy(rp)
    rp5["state_generation"]["kwargs"]["n_states"] = 5
    out = str(tmp_path)
    s5 = run_series(MonteCalculator("semigrand_canonical", zro, params={"seed": 1}), rp5, out)
    d5 = _validate_summary_file(os.path.join(out, "thermo", "summary.json"), 5, canonical=False)
    assert len(json.load(open(os.path.join(out, "completed_runs.json")))) == 5
    s11 = run_series(MonteCalculator("semigrand_canonical", zro, params={"seed": 2}), rp, out)
    d11 = _validate_summary_file(os.path.join(out, "thermo", "summary.json"), 11, canonical=False)
    runs = json.load(open(os.path.join(out, "completed_runs.json")))
    assert len(runs) == 11
    np.testing.assert_allclose([r["conditions"]["param_chem_pot"][0] for r in runs], -4.0 + 0.4 * np.arange(11))
    assert d11["conditions"]["param_chem_pot"]["a"][:5] == d5["conditions"]["param_chem_pot"]["a"]
    assert d11["statistics"]["potential_energy"]["value"]["mean"][:5] == \
        s5["statistics"]["potential_energy"]["value"]["mean"]
    again = run_series(MonteCalculator("semigrand_canonical", zro, params={"seed": 3}), rp, out)
    assert again == s11
    _check_completion(d11)
    # more O at higher chemical potential
    x = d11["statistics"]["param_composition"]["a"]["mean"]
    assert x[0] < 1e-3 and x[-1] > 0.4 and np.all(np.diff(x) > -0.02)


def test_canonical_series_files(zro, tmp_path):
    _print_case("ZrO canonical series: summary.json")
    out = str(tmp_path)
    run_series(MonteCalculator("canonical", zro, params={"seed": 1}), ZRO["run_params_complete"], out)
    d = _validate_summary_file(os.path.join(out, "thermo", "summary.json"), 11, canonical=True)
    np.testing.assert_allclose(d["conditions"]["mol_composition"]["O"], [22 / 216] * 11)
    np.testing.assert_allclose(d["conditions"]["temperature"]["value"], 300.0 + 10 * np.arange(11))
    _check_completion(d)


def test_series_observations_trajectory_and_sparse_corr(zro, tmp_path):
    _print_case("ZrO series: observations.json, trajectory.json, sparse_corr")
    rp = copy.deepcopy(ZRO["run_params_sgc_complete"])
    rp["state_generation"]["kwargs"]["n_states"] = 3
    fx = rp["sampling_fixtures"]["thermo"]
    fx["sampling"]["quantities"] = fx["sampling"]["quantities"] + ["clex.formation_energy.sparse_corr"]
    fx["results_io"]["kwargs"].update({"write_observations": True, "write_trajectory": True})
    out = str(tmp_path)
    run_series(MonteCalculator("semigrand_canonical", zro, params={"seed": 5}), rp, out)
    summ = json.load(open(os.path.join(out, "thermo", "summary.json")))
    n_samples = summ["completion_check_results"]["N_samples"]
    eci = {cf["linear_function_index"]: cf.get("eci", 0.0)
           for orb in ZRO["formation_energy_eci"]["orbits"] for cf in orb["cluster_functions"]}
    nonzero = sorted(k for k, v in eci.items() if v != 0.0)
    sp = summ["statistics"]["clex.formation_energy.sparse_corr"]
    assert sp["component_names"] == [str(k) for k in nonzero]
    osys = O.OracleSystem(zro.flat, (6, 6, 6))
    for i in range(3):
        run_dir = os.path.join(out, "thermo", f"run.{i + 1}")
        assert sorted(os.listdir(run_dir)) == ["observations.json", "trajectory.json"]
        obs = json.load(open(os.path.join(run_dir, "observations.json")))
        traj = json.load(open(os.path.join(run_dir, "trajectory.json")))
        ns = n_samples[i]
        assert len(obs["count"]) == ns and len(traj) == ns and obs["count"] == [t["count"] for t in traj]
        for k in ("clex.formation_energy", "potential_energy", "mol_composition", "param_composition",
                  "temperature", "param_chem_pot", "corr.formation_energy", "clex.formation_energy.sparse_corr"):
            assert len(obs[k]["value"]) == ns, k
        e = np.array(obs["clex.formation_energy"]["value"])[:, 0]
        spc = np.array(obs["clex.formation_energy.sparse_corr"]["value"])
        assert obs["clex.formation_energy.sparse_corr"]["component_names"] == [str(k) for k in nonzero]
        np.testing.assert_allclose(spc @ np.array([eci[k] for k in nonzero]), e, rtol=1e-9, atol=1e-12)
        for t, ei in zip(traj[::10], e[::10]):
            occ = np.array(t["dof"]["occ"], dtype=np.int8)
            assert abs(osys.energy(occ) / 216 - ei) <= 1e-5, (osys.energy(occ) / 216, ei)
