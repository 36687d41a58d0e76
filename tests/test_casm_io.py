"""CPU tests of the CASM file reader (casm_io) on the reference's fixtures
Clex_ZrO_Occ and FCC_binary_vacancy (data in tests/golden/casm_clex_fixtures.json,
written by tests/golden/make_casm_clex_fixtures.py).

Pins:
* orbit structure: every CASM orbit matched once, with its size,
  multiplicity and lengths (the reference enumerated these itself);
* inclusion-exclusion: for every orbit of the file and every decoration of
  its prototype, the oracle's supercell energies over the subsets of the
  decorated sites (other sites holding their first occupant, whose
  occupation-basis value is 0) combine to sum_f eci_f * c_f(decoration) /
  mult, computed from the file's own latex, ECIs and multiplicities.  A
  wrong ECI assignment, normalisation, decoration order or site-basis sign
  fails it;
* the composition axes equal the fixture's system.json axes;
* run-parameter state generation: temperatures, chemical potentials and the
  composition the reference's enforce_composition reaches;
* planted mistakes in the file data are refused.
"""
import copy
import itertools
import json
import os

import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.calculator import System
from casmcode_clexmonte_amd.casm_io import clex_from_casm, states_from_run_params
from casmcode_clexmonte_amd.system import FlatCE, Supercell, check_supercell, parse_casm_function_latex

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = json.load(open(os.path.join(ROOT, "tests", "golden", "casm_clex_fixtures.json")))
ZRO = FIX["Clex_ZrO_Occ"]
FCC = FIX["FCC_binary_vacancy"]


@pytest.fixture(scope="module")
def zro_ce():
    return clex_from_casm(ZRO["prim"], ZRO["formation_energy_eci"])


def test_zro_orbits_and_ecis(zro_ce):
    bs = zro_ce.basis_set
    sizes = {}
    for o in bs.orbits:
        sizes[o.size] = sizes.get(o.size, 0) + 1
    assert sizes == {0: 1, 1: 1, 2: 16, 3: 42, 4: 14}
    assert bs.n_functions == 74
    file_eci = {cf["linear_function_index"]: cf["eci"] for o in ZRO["formation_energy_eci"]["orbits"]
                for cf in o["cluster_functions"] if "eci" in cf}
    assert len(file_eci) == 33 and np.count_nonzero(zro_ce.eci) == 33
    for lfi, v in file_eci.items():
        assert zro_ce.eci[zro_ce.casm_to_local[lfi]] == v
    assert sorted(zro_ce.casm_to_local.values()) == list(range(74))


def test_fcc_multi_function_orbits_and_sparse():
    ce = clex_from_casm(FCC["prim"], FCC["formation_energy_eci"])
    np.testing.assert_array_equal(ce.eci, [0, -0.1, 0.3, 0.1, 0.1, 0.5, 0, 0, 0])
    sp = clex_from_casm(FCC["prim"], FCC["formation_energy_eci"], FCC["formation_energy_sparse_eci"])
    np.testing.assert_array_equal(sp.eci, ce.eci)


def _inclusion_exclusion(fixture, dims, max_orbits=None):
    """max |V_oracle - V_file| over orbits and decorations."""
    ce = clex_from_casm(fixture["prim"], fixture["formation_energy_eci"])
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, dims)
    osys = O.OracleSystem(flat, dims)
    prim = ce.prim
    worst = 0.0
    n_checked = 0
    for co in fixture["formation_energy_eci"]["orbits"][:max_orbits]:
        sites = [tuple(s) for s in co["prototype"]["sites"]]
        if not sites:
            continue
        funcs = []
        for cf in co["cluster_functions"]:
            latex = [v for k, v in cf.items() if k.startswith("\\Phi")][0]
            funcs.append((float(cf.get("eci", 0.0)), parse_casm_function_latex(latex)))
        # every decoration: site i holds occupant deco[i] + 1 (occupation basis function index deco[i])
        ranges = [range(prim.n_occ(b) - 1) for (b, _, _, _) in sites]
        for deco in itertools.product(*ranges):
            expect = 0.0
            for eci, terms in funcs:
                for c, d in terms:
                    if all(d.get(i) == deco[i] for i in range(len(sites))):
                        expect += eci * c
            expect /= co["mult"]
            V = 0.0
            for r in range(len(sites) + 1):
                for S in itertools.combinations(range(len(sites)), r):
                    occ = np.zeros(sc.n_sites, np.int8)
                    for q in S:
                        b, i, j, k = sites[q]
                        occ[sc.linear_index(b, i % dims[0], j % dims[1], k % dims[2])] = deco[q] + 1
                    V += (-1) ** (len(sites) - r) * osys.energy(occ)
            worst = max(worst, abs(V - expect))
            n_checked += 1
    return worst, n_checked


def test_zro_inclusion_exclusion():
    # 6x6x4: no cluster (pairs <= 9 A, a = 3.234 A, c = 5.169 A) meets its own periodic image
    worst, n = _inclusion_exclusion(ZRO, (6, 6, 4))
    assert n == 73
    assert worst < 1e-12, worst


def test_fcc_ternary_inclusion_exclusion():
    worst, n = _inclusion_exclusion(FCC, (4, 4, 4))
    assert n == 2 + 4 + 4
    assert worst < 1e-12, worst


def test_zro_composition_axes_match_fixture(zro_ce):
    axes = System(zro_ce).composition_axes
    ref = ZRO["composition_axes"]
    assert axes.components == ref["components"]
    np.testing.assert_allclose(axes.origin, np.array(ref["origin"]).ravel())
    end_a = np.array(ref["a"]).ravel()
    np.testing.assert_allclose(axes.R[:, 0], end_a - axes.origin)
    # param_formula a(0.5-0.25Va+0.25O)
    mol = np.array([2.0, 1.4, 0.6])
    assert abs(axes.param_composition(mol)[0] - (0.5 - 0.25 * 1.4 + 0.25 * 0.6)) < 1e-12


def test_zro_6x6x6_fits_the_gpu_layout(zro_ce):
    check_supercell(FlatCE(zro_ce), Supercell(zro_ce.prim, (6, 6, 6)))


def test_run_params_states(zro_ce):
    system = System(zro_ce)
    can = states_from_run_params(ZRO["run_params_complete"], system, "canonical", seed=3)
    assert [s.conditions["temperature"] for s in can] == [300.0 + 10 * i for i in range(11)]
    n = 216
    for s in can:
        occ = np.asarray(s.occupation).reshape(4, n)
        assert s.supercell == (6, 6, 6)
        assert not occ[:2].any()
        # 0.1 O per unit cell of 216: the closest reachable count is 22
        assert int(occ[2:].sum()) == 22
        np.testing.assert_allclose(s.conditions["mol_composition"], [2.0, 2 - 22 / n, 22 / n])
    assert not np.array_equal(can[0].occupation, can[1].occupation)
    sgc = states_from_run_params(ZRO["run_params_sgc_complete"], system, "semigrand_canonical")
    assert [s.conditions["temperature"] for s in sgc] == [1000.0] * 11
    np.testing.assert_allclose([s.conditions["param_chem_pot"][0] for s in sgc], -4.0 + 0.4 * np.arange(11))
    nf = states_from_run_params(ZRO["run_params_sgc_complete"], system, "nfold")  # same conditions
    assert [s.conditions["temperature"] for s in nf] == [1000.0] * 11
    np.testing.assert_allclose([s.conditions["param_chem_pot"][0] for s in nf], -4.0 + 0.4 * np.arange(11))


def test_planted_mistakes_are_refused():
    bad = copy.deepcopy(ZRO["formation_energy_eci"])
    bad["orbits"][3]["mult"] = 5
    with pytest.raises(ValueError, match="multiplicity"):
        clex_from_casm(ZRO["prim"], bad)
    bad = copy.deepcopy(ZRO["formation_energy_eci"])
    bad["orbits"][3]["prototype"]["sites"][1] = [2, 0, 2, 0]
    with pytest.raises(ValueError):
        clex_from_casm(ZRO["prim"], bad)
    bad = copy.deepcopy(FCC["formation_energy_eci"])
    f = bad["orbits"][2]["cluster_functions"][0]
    key = [k for k in f if k.startswith("\\Phi")][0]
    f[key] = "\\phi_{0,0}(s_{0})\\phi_{0,1}(s_{1})"   # an unsymmetrised decoration
    with pytest.raises(ValueError, match="matches no function"):
        clex_from_casm(FCC["prim"], bad)
    with pytest.raises(ValueError, match="not a function"):
        clex_from_casm(FCC["prim"], FCC["formation_energy_eci"], [[42, 1.0]])


def test_zro_asymmetric_units_match_fixture(zro_ce):
    """The two Zr sublattices and the two interstitial sublattices are each
    one asymmetric unit (the file's site_functions asym_unit column):
    canonical swaps may exchange O between the two interstitial sublattices."""
    ref = [e["asym_unit"] for e in sorted(ZRO["formation_energy_eci"]["site_functions"], key=lambda e: e["sublat"])]
    assert list(FlatCE(zro_ce).asym) == ref == [0, 0, 1, 1]


def test_enforce_composition_reference_values(zro_ce):
    """The reference's own composition results for the ZrO system
    (python/tests/functions/test_enforce_composition.py:49,59): a supercell of
    1029 unit cells holding one O per three cells ([2, 5/3, 1/3]) driven to
    the target [2, 1.8, 0.2] ends at [2, 1852/1029, 206/1029], the closest
    reachable composition.  Here the 1029 cells form the diagonal 21x7x7
    supercell (the composition walk depends only on counts)."""
    from casmcode_clexmonte_amd.casm_io import enforce_composition

    system = System(zro_ce)
    axes = system.composition_axes
    dims = (21, 7, 7)
    n = 21 * 7 * 7
    sc = Supercell(zro_ce.prim, dims)
    occ = np.zeros(sc.n_sites, np.int8)
    for k in range(7):
        for j in range(7):
            for i in range(0, 21, 3):
                occ[sc.linear_index(2, i, j, k)] = 1
    np.testing.assert_allclose(axes.mol_composition(occ, n), [2.0, 5.0 / 3.0, 1.0 / 3.0])
    out = enforce_composition(occ, axes, zro_ce.prim, n, [2.0, 1.8, 0.2], np.random.default_rng(1))
    np.testing.assert_allclose(axes.mol_composition(out, n), [2.0, 1852 / 1029, 206 / 1029])
    assert np.all(out[:2 * n] == 0)


def test_canonical_state_from_param_composition(zro_ce):
    """param_composition a = 0.5 from the default (all Va) configuration gives
    mol_composition [2, 1, 1] (python/tests/canonical/test_CanonicalCalculator.py:131-142)."""
    system = System(zro_ce)
    rp = {"state_generation": {"method": "incremental", "kwargs": {
        "initial_configuration": {"method": "fixed", "kwargs": {
            "transformation_matrix_to_supercell": [[10, 0, 0], [0, 10, 0], [0, 0, 10]]}},
        "initial_conditions": {"temperature": 300.0, "param_composition": {"a": 0.5}},
        "conditions_increment": {"temperature": 0.0}, "n_states": 1, "dependent_runs": False}}}
    (st,) = states_from_run_params(rp, system, "canonical", seed=2)
    np.testing.assert_allclose(st.conditions["mol_composition"], [2.0, 1.0, 1.0])
    np.testing.assert_allclose(st.conditions["param_composition"], [0.5])


def test_flat_file_cli_matches_save_flat(tmp_path):
    """The command-line route from CASM files to the C++ calculators' flat
    file writes what save_flat writes for the same expansion."""
    import subprocess
    import sys
    from casmcode_clexmonte_amd.system import save_flat
    prim = tmp_path / "prim.json"
    eci = tmp_path / "formation_energy_eci.json"
    prim.write_text(json.dumps(ZRO["prim"]))
    eci.write_text(json.dumps(ZRO["formation_energy_eci"]))
    out = tmp_path / "system.flat"
    r = subprocess.run([sys.executable, "-m", "casmcode_clexmonte_amd.casm_io", "--prim", str(prim),
                        "--eci", str(eci), "--out", str(out)], cwd=ROOT, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    ref = tmp_path / "ref.flat"
    save_flat(FlatCE(clex_from_casm(ZRO["prim"], ZRO["formation_energy_eci"])), str(ref))
    assert out.read_text() == ref.read_text()
    assert "n_sublat" in out.read_text() and "flower_tables" in out.read_text()
