"""The C++ calculator layer (GpuCanonicalCalculator /
GpuSemiGrandCanonicalCalculator through the extern "C" factories), driven
like the reference's C++ calculator tests, checked against the oracle:
potential per_supercell (FP64 correlations, 1e-9 relative), per-proposal
occ_delta_per_supercell (FP32, 1e-5 of the batch scale), composition
conservation and sampler / analysis output layout."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.system import FlatCE, Supercell, save_flat
from casmcode_clexmonte_amd.workloads import fcc_binary_ce, random_occupation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "calculator_driver")


def _run(tmp_path, method, dims, occ, T, mu, n_equil, n_passes, probes=None, env=None):
    flat_path = tmp_path / "system.flat"
    occ_path = tmp_path / "occ.txt"
    probe_path = tmp_path / "probe.txt"
    occ_path.write_text(" ".join(str(int(v)) for v in occ))
    args = [DRIVER, str(flat_path), method, *map(str, dims), str(occ_path), str(T), "-" if mu is None else str(mu),
            str(n_equil), str(n_passes), "12345"]
    if probes:
        probe_path.write_text("\n".join(f"{len(s)} {' '.join(map(str, s))} {' '.join(map(str, o))}" for s, o in probes))
        args.append(str(probe_path))
    out = subprocess.run(args, capture_output=True, text=True, timeout=300, env=env)
    return json.loads(out.stdout.strip().splitlines()[-1])


@pytest.fixture(scope="module")
def flat():
    return FlatCE(fcc_binary_ce())


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["canonical", "semigrand_canonical"])
def test_cpp_calculator_vs_oracle(tmp_path, flat, method):
    print(f"[case] cpp calculator {method}", flush=True)
    save_flat(flat, str(tmp_path / "system.flat"))
    dims = (8, 8, 8)
    sc = Supercell(flat.ce.prim, dims)
    occ = random_occupation(sc.n_sites, 0.5, 21)
    mu = None if method == "canonical" else -0.02
    rng = np.random.default_rng(1)
    probes = []
    for _ in range(20):
        s = int(rng.integers(sc.n_sites))
        probes.append(([s], [1 - int(occ[s])]))
    for _ in range(10):
        a, b = rng.choice(sc.n_sites, 2, replace=False)
        probes.append(([int(a), int(b)], [int(occ[b]), int(occ[a])]))
    r = _run(tmp_path, method, dims, occ, 800.0, mu, 50, 200, probes)
    assert "error" not in r, r
    osys = O.OracleSystem(flat, dims)
    E = osys.energy(occ)
    if method == "canonical":
        assert abs(r["per_supercell"] - E) <= 1e-9 * max(1, abs(E))
    else:
        assert abs(r["per_supercell"] - (E - mu * int(occ.sum()))) <= 1e-9 * max(1, abs(E))
    ref = []
    for s, o in probes:
        d = osys.delta(occ, s, o)
        if method != "canonical":
            d -= mu * sum(int(v) - int(occ[x]) for x, v in zip(s, o))
        ref.append(d)
    ref = np.array(ref)
    assert np.all(np.abs(np.array(r["deltas"]) - ref) <= 1e-5 * np.abs(ref).max() + 1e-6)
    assert r["n_samples"] == 200
    assert r["mol_size"] == sc.n_sites
    if method == "canonical":
        assert r["final_occupation_sum"] == int(occ.sum())
        assert set(r["analysis"]) == {"heat_capacity"}
    else:
        assert {"heat_capacity", "mol_susc", "param_susc", "mol_thermochem_susc", "param_thermochem_susc"} <= \
            set(r["analysis"])
    assert "potential_energy(0)" in r["statistics"] and "mol_composition(1)" in r["statistics"]
    assert r["n_accept"] + r["n_reject"] > 0


def test_cpp_calculator_fails_loudly_without_gpu(tmp_path, flat):
    save_flat(flat, str(tmp_path / "system.flat"))
    dims = (4, 4, 4)
    occ = random_occupation(64, 0.5, 1)
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = ""
    env["ROCR_VISIBLE_DEVICES"] = ""
    r = _run(tmp_path, "canonical", dims, occ, 800.0, None, 1, 1, env=env)
    assert "error" in r and "clexmc_create" in r["error"]


def test_cpp_calculator_validates_conditions(tmp_path, flat):
    """Missing param_chem_pot is an invalid semi-grand canonical state
    (SemiGrandCanonicalCalculator.cc validate_conditions) - checked before
    any device call."""
    save_flat(flat, str(tmp_path / "system.flat"))
    occ = random_occupation(64, 0.5, 1)
    r = _run(tmp_path, "semigrand_canonical", (4, 4, 4), occ, 800.0, None, 1, 1)
    assert "error" in r and "param_chem_pot" in r["error"]
