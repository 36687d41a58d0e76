"""The Python calculator face (MonteCalculator / MontePotential /
run_fixture) on the GPU: potential values against the oracle and the
summary layout the reference's python tests check
(python/tests/conftest.py:140-259, validate_summary_file)."""
import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.calculator import (MonteCalculator, MonteCarloState, MontePotential, System,
                                               merge_summaries)
from casmcode_clexmonte_amd.workloads import fcc_binary_ce, random_occupation


def validate_summary_data(subdata, expected_keys, expected_size):
    # restates the reference's conftest.py validate_summary_data (:140-154)
    for x in expected_keys:
        assert x in subdata
        if "component_names" in subdata[x]:
            for y in subdata[x]["component_names"]:
                assert len(subdata[x][y]) == expected_size
        elif "value" in subdata[x]:
            assert subdata[x]["shape"] == []
            assert len(subdata[x]["value"]) == expected_size
        else:
            assert len(subdata[x]) == expected_size


def validate_statistics_data(subdata, expected_keys, expected_size):
    for x in expected_keys:
        assert x in subdata
        if "component_names" in subdata[x]:
            for y in subdata[x]["component_names"]:
                for z in ["mean", "calculated_precision"]:
                    assert len(subdata[x][y][z]) == expected_size
        else:
            assert subdata[x]["shape"] == []
            for z in ["mean", "calculated_precision"]:
                assert len(subdata[x]["value"][z]) == expected_size


def validate_summary(data, expected_size, is_canonical):
    if is_canonical:
        cond_keys, an_keys = ["temperature", "param_composition", "mol_composition"], ["heat_capacity"]
    else:
        cond_keys = ["temperature", "param_chem_pot"]
        an_keys = ["heat_capacity", "mol_susc", "param_susc", "mol_thermochem_susc", "param_thermochem_susc"]
    validate_summary_data(data["analysis"], an_keys, expected_size)
    validate_summary_data(data["completion_check_results"],
                          ["N_samples", "N_samples_for_statistics", "acceptance_rate", "count", "elapsed_clocktime",
                           "N_samples_for_all_to_equilibrate", "all_converged", "all_equilibrated"], expected_size)
    validate_summary_data(data["conditions"], cond_keys, expected_size)
    validate_statistics_data(data["statistics"], ["clex.formation_energy", "mol_composition", "param_composition",
                                                  "potential_energy"], expected_size)


@pytest.fixture(scope="module")
def system():
    return System(fcc_binary_ce())


@pytest.mark.gpu
def test_python_potential_and_canonical_series(system):
    print("[case] python calculator canonical", flush=True)
    calc = MonteCalculator("canonical", system)
    dims = (6, 6, 6)
    occ = random_occupation(216, 0.5, 2)
    state = MonteCarloState(dims, occ.copy(), {"temperature": 900.0})
    pot = MontePotential(calc, state)
    osys = O.OracleSystem(system.flat, dims)
    assert abs(pot.per_supercell() - osys.energy(occ)) < 1e-9
    d = pot.occ_delta_per_supercell([3, 17], [int(occ[17]), int(occ[3])])
    ref = osys.delta(occ, [3, 17], [int(occ[17]), int(occ[3])])
    assert abs(d - ref) <= 1e-5 * max(abs(ref), 0.01) + 1e-6
    summaries = []
    for T in (1200.0, 900.0, 600.0):
        state.conditions["temperature"] = T
        p = calc.make_default_sampling_fixture_params(n_equilibration_passes=20, n_passes=100)
        summaries.append(calc.run_fixture(state, p))
        assert int(state.occupation.sum()) == int(occ.sum())
    validate_summary(merge_summaries(summaries), 3, True)


@pytest.mark.gpu
def test_python_sgc_multistate(system):
    print("[case] python calculator sgc multistate", flush=True)
    calc = MonteCalculator("semigrand_canonical", system)
    states = [MonteCarloState((8, 8, 8), random_occupation(512, 0.5, 10 + i, exact=False),
                              {"temperature": 1000.0, "param_chem_pot": [mu]})
              for i, mu in enumerate(np.linspace(-0.1, 0.1, 8))]
    p = calc.make_default_sampling_fixture_params(n_equilibration_passes=50, n_passes=200)
    out = calc.run_multistate(states, p)
    validate_summary(merge_summaries(out), 8, False)
    x = [o["statistics"]["param_composition"]["a"]["mean"][0] for o in out]
    assert np.all(np.diff(x) > 0)  # composition increases with the B chemical potential


def test_calculator_rejects_bad_conditions(system):
    """Condition validation happens before any device call."""
    with pytest.raises(Exception):
        MonteCalculator("grand_canonical_typo", system)


def _moment_blocks(w, E, N, n_blocks):
    """Residence-weighted moment records (clexmc_nfold_run_moments layout)
    of states (w, E, N[4]) cut into n_blocks consecutive blocks."""
    from casmcode_clexmonte_amd.engine import NFOLD_MOMENTS, NFOLD_PAIRS
    blk = np.zeros((n_blocks, NFOLD_MOMENTS))
    for b, s in enumerate(np.array_split(np.arange(len(w)), n_blocks)):
        ww, ee, nn = w[s].astype(float), E[s], N[s].astype(float)
        blk[b, 0], blk[b, 1], blk[b, 2] = ww.sum(), (ww * ee).sum(), (ww * ee * ee).sum()
        blk[b, 3:7] = (ww[:, None] * nn).sum(0)
        for t, (o, q) in enumerate(NFOLD_PAIRS):
            blk[b, 7 + t] = (ww * nn[:, o] * nn[:, q]).sum()
        blk[b, 17:21] = (ww[:, None] * ee[:, None] * nn).sum(0)
    return blk


def _compare_nfold_summary(system, st, w, E, N, comps, axes_names):
    nf = object.__new__(MonteCalculator)
    nf.method, nf.system = "nfold", system
    sg = object.__new__(MonteCalculator)
    sg.method, sg.system = "semigrand_canonical", system
    got = nf._nfold_summary(st, _moment_blocks(w, E, N, 30), 3000, 0.0, None)
    Er, Nr = np.repeat(E, w), np.repeat(N, w, axis=0)
    ref = sg._summary(st, Er, Nr, 1, 1, 0.0)
    for q in ("clex.formation_energy", "potential_energy"):
        assert got["statistics"][q]["value"]["mean"][0] == pytest.approx(ref["statistics"][q]["value"]["mean"][0],
                                                                         rel=1e-12, abs=1e-12)
    for k in comps:
        assert got["statistics"]["mol_composition"][k]["mean"][0] == pytest.approx(
            ref["statistics"]["mol_composition"][k]["mean"][0], rel=1e-12, abs=1e-14)
    for k in axes_names:
        assert got["statistics"]["param_composition"][k]["mean"][0] == pytest.approx(
            ref["statistics"]["param_composition"][k]["mean"][0], rel=1e-12, abs=1e-14)
    assert got["analysis"]["heat_capacity"]["value"][0] == pytest.approx(ref["analysis"]["heat_capacity"]["value"][0],
                                                                        rel=1e-8)
    for key, names in (("mol_susc", [f"{i},{j}" for i in comps for j in comps]),
                       ("param_susc", [f"{i},{j}" for i in axes_names for j in axes_names]),
                       ("mol_thermochem_susc", [f"S,{j}" for j in comps]),
                       ("param_thermochem_susc", [f"S,{j}" for j in axes_names])):
        for c in names:
            g, r = got["analysis"][key][c][0], ref["analysis"][key][c][0]
            assert g == pytest.approx(r, rel=1e-8, abs=1e-10), (key, c, g, r)
    return got


def test_nfold_summary_equals_replicated_samples(system):
    """The N-fold summary is formed from residence-weighted moments (sums of
    w, wE, wE^2, wN_o, wN_oN_q, wEN_o per block).  With integer weights, the
    same states repeated w times are ordinary samples: their plain summary
    (the semi-grand canonical one) must give the same means, heat capacity
    and susceptibilities.  Binary FCC and ternary HCP (two independent
    composition axes).  No device work: the calculators are built without an
    engine."""
    rng = np.random.default_rng(5)
    n = 512
    E = rng.normal(-20.0, 3.0, 300)
    N1 = rng.integers(150, 360, 300)
    N = np.zeros((300, 4), dtype=np.int64)
    N[:, 0], N[:, 1] = n - N1, N1
    w = rng.integers(1, 6, 300)
    st = MonteCarloState((8, 8, 8), np.zeros(n, np.int8), {"temperature": 700.0, "param_chem_pot": [0.07]})
    got = _compare_nfold_summary(system, st, w, E, N, ("A", "B"), ("a",))
    assert got["conditions"]["param_chem_pot"]["a"] == [0.07]
    # ternary HCP: 2 sublattices x 4^3 cells, occupants A, B, C on both
    from casmcode_clexmonte_amd.workloads import hcp_ternary_ce
    sys3 = System(hcp_ternary_ce())
    n3 = 2 * 64
    N3 = np.zeros((300, 4), dtype=np.int64)
    N3[:, 1] = rng.integers(20, 60, 300)
    N3[:, 2] = rng.integers(10, 50, 300)
    N3[:, 0] = n3 - N3[:, 1] - N3[:, 2]
    E3 = rng.normal(-5.0, 1.5, 300)
    st3 = MonteCarloState((4, 4, 4), np.zeros(n3, np.int8), {"temperature": 900.0, "param_chem_pot": [0.05, -0.03]})
    axes = sys3.composition_axes
    _compare_nfold_summary(sys3, st3, w, E3, N3, tuple(axes.components),
                           tuple(chr(ord("a") + i) for i in range(axes.n_independent)))


@pytest.mark.gpu
def test_python_nfold_matches_sgc(system):
    """MonteCalculator("nfold") (residence-time-weighted rejection-free
    events) and the semi-grand canonical calculator sample the same
    ensemble: formation energy and composition means agree within 4 combined
    standard errors at every chemical potential."""
    print("[case] python calculator nfold vs sgc, fcc binary 8^3, T 900 K", flush=True)
    mus = [-0.06, 0.0, 0.06]

    def states():
        return [MonteCarloState((8, 8, 8), random_occupation(512, 0.5, 40 + i, exact=False),
                                {"temperature": 900.0, "param_chem_pot": [mu]}) for i, mu in enumerate(mus)]
    sg = MonteCalculator("semigrand_canonical", system)
    p = sg.make_default_sampling_fixture_params(n_equilibration_passes=200, n_passes=3000)
    ref = sg.run_multistate(states(), p)
    nf = MonteCalculator("nfold", system)
    got = nf.run_multistate(states(), nf.make_default_sampling_fixture_params(n_equilibration_passes=200,
                                                                               n_passes=1500))
    validate_summary(merge_summaries(got), len(mus), False)
    for r, g in zip(ref, got):
        for sec, key in (("clex.formation_energy", "value"), ("param_composition", "a")):
            a, b = r["statistics"][sec][key], g["statistics"][sec][key]
            se = np.hypot(a["calculated_precision"][0], b["calculated_precision"][0]) / 2.0
            assert abs(a["mean"][0] - b["mean"][0]) < 4 * se + 1e-6, (sec, a, b)
        assert g["completion_check_results"]["count"][0] == 1500 * 512


@pytest.mark.gpu
def test_python_nfold_ternary_matches_sgc():
    """The nfold method for a ternary expansion (HCP, 2 sublattices, A/B/C;
    the records carry the moments of every occupant count) against the
    semi-grand canonical calculator: formation energy and both composition
    axes agree within 4 combined standard errors at two chemical
    potentials; heat capacity and the a,a / b,b susceptibilities agree
    within 25 % (their errors are not estimated)."""
    from casmcode_clexmonte_amd.workloads import hcp_ternary_ce
    print("[case] python calculator nfold vs sgc, hcp ternary 6x6x4, T 1500 K", flush=True)
    sys3 = System(hcp_ternary_ce())
    mus = [[0.0, 0.0], [0.04, -0.03]]
    n = 2 * 6 * 6 * 4

    def states():
        rng = np.random.default_rng(3)
        return [MonteCarloState((6, 6, 4), rng.integers(0, 3, n).astype(np.int8),
                                {"temperature": 1500.0, "param_chem_pot": list(mu)}) for mu in mus]
    sg = MonteCalculator("semigrand_canonical", sys3)
    ref = sg.run_multistate(states(), sg.make_default_sampling_fixture_params(n_equilibration_passes=300,
                                                                              n_passes=4000))
    nf = MonteCalculator("nfold", sys3)
    got = nf.run_multistate(states(), nf.make_default_sampling_fixture_params(n_equilibration_passes=300,
                                                                               n_passes=2000))
    validate_summary(merge_summaries(got), len(mus), False)
    for r, g in zip(ref, got):
        for sec, key in (("clex.formation_energy", "value"), ("param_composition", "a"), ("param_composition", "b")):
            a, b = r["statistics"][sec][key], g["statistics"][sec][key]
            se = np.hypot(a["calculated_precision"][0], b["calculated_precision"][0]) / 2.0
            print(sec, key, a["mean"][0], b["mean"][0], se)
            assert abs(a["mean"][0] - b["mean"][0]) < 4 * se + 1e-6, (sec, key, a, b)
        for key, comp in (("heat_capacity", None), ("param_susc", "a,a"), ("param_susc", "b,b")):
            a = r["analysis"][key]["value" if comp is None else comp][0]
            b = g["analysis"][key]["value" if comp is None else comp][0]
            print(key, comp, a, b)
            assert abs(a - b) <= 0.25 * abs(a), (key, comp, a, b)
        assert g["completion_check_results"]["count"][0] == 2000 * n


@pytest.mark.gpu
def test_python_nfold_time_sampling_matches_sgc(system):
    """MonteCalculator("nfold") sampled by time (sample_by "time", one sample
    per 0.5 passes of N-fold time): formation energy and composition means
    agree with the semi-grand canonical calculator within 4 combined standard
    errors, and the sample count is n_passes / period."""
    print("[case] python calculator nfold sampled by time vs sgc, fcc binary 8^3", flush=True)
    from casmcode_clexmonte_amd.calculator import SamplingFixtureParams
    mus = [-0.04, 0.05]

    def states():
        return [MonteCarloState((8, 8, 8), random_occupation(512, 0.5, 60 + i, exact=False),
                                {"temperature": 900.0, "param_chem_pot": [mu]}) for i, mu in enumerate(mus)]
    sg = MonteCalculator("semigrand_canonical", system)
    ref = sg.run_multistate(states(), sg.make_default_sampling_fixture_params(n_equilibration_passes=200,
                                                                              n_passes=3000))
    nf = MonteCalculator("nfold", system)
    p = SamplingFixtureParams(n_equilibration_passes=200, n_passes=1500, sample_period=0.5, sample_by="time")
    got = nf.run_multistate(states(), p)
    validate_summary(merge_summaries(got), len(mus), False)
    for r, g in zip(ref, got):
        assert g["completion_check_results"]["N_samples"] == [3000]
        for sec, key in (("clex.formation_energy", "value"), ("param_composition", "a")):
            a, b = r["statistics"][sec][key], g["statistics"][sec][key]
            se = np.hypot(a["calculated_precision"][0], b["calculated_precision"][0]) / 2.0
            print(sec, a["mean"][0], b["mean"][0], se)
            assert abs(a["mean"][0] - b["mean"][0]) < 4 * se + 1e-6, (sec, a, b)


def test_fixture_sample_by_time():
    """CASM sampling fixtures may sample by time (the nfold method)."""
    from casmcode_clexmonte_amd.calculator import SamplingFixtureParams
    p = SamplingFixtureParams.from_casm("thermo", {"sampling": {"sample_by": "time", "period": 0.25,
                                                                "quantities": ["formation_energy"]}})
    assert p.sample_by == "time" and p.sample_period == 0.25
    with pytest.raises(ValueError):
        SamplingFixtureParams.from_casm("thermo", {"sampling": {"sample_by": "step"}})


def test_fixture_legacy_quantity_names():
    """CASM run params name the samplers formation_energy / formation_energy_corr."""
    from casmcode_clexmonte_amd.calculator import SamplingFixtureParams
    p = SamplingFixtureParams.from_casm("thermo", {"sampling": {"period": 2, "quantities": [
        "formation_energy", "formation_energy_corr", "potential_energy"]}})
    assert p.sampler_names == ["clex.formation_energy", "corr.formation_energy", "potential_energy"]
    assert p.sample_period == 2


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["canonical", "semigrand_canonical"])
def test_python_corr_sampler(system, method):
    """The corr.formation_energy sampler: all correlations per unit cell at
    every sample.  The energy is linear in them, so ECI . <corr> equals the
    mean formation energy of the same samples (1e-9) only if the correlation
    reads line up with the kernel's energy samples; the empty-cluster
    correlation is 1."""
    print(f"[case] python calculator corr sampler {method}", flush=True)
    calc = MonteCalculator(method, system)
    cond = {"temperature": 1000.0}
    if method != "canonical":
        cond["param_chem_pot"] = [0.02]
    states = [MonteCarloState((6, 6, 6), random_occupation(216, 0.5, 70 + i), dict(cond)) for i in range(2)]
    p = calc.make_default_sampling_fixture_params(n_equilibration_passes=20, n_passes=60, sample_period=3)
    p.sampler_names = p.sampler_names + ["corr.formation_energy"]
    out = calc.run_multistate(states, p)
    eci = np.asarray(system.flat.eci)
    for o in out:
        d = o["statistics"]["corr.formation_energy"]
        assert d["shape"] == [len(eci)] and o["completion_check_results"]["N_samples"] == [20]
        mc = np.array([d[k]["mean"][0] for k in d["component_names"]])
        assert mc[0] == pytest.approx(1.0, abs=1e-12)
        e = o["statistics"]["clex.formation_energy"]["value"]["mean"][0]
        assert float(eci @ mc) == pytest.approx(e, rel=1e-9, abs=1e-12)


@pytest.mark.gpu
def test_python_sparse_corr_sampler():
    """clex.formation_energy.sparse_corr (sampling_functions.cc:185-219): the
    correlations of the non-zero ECIs only.  The ordering CE has zero empty
    and point ECIs, so the sampler holds the 6 pair and triplet correlations,
    named by their indices, and the non-zero ECIs dotted with their means
    equal the mean formation energy (1e-9)."""
    print("[case] python calculator sparse_corr sampler", flush=True)
    from casmcode_clexmonte_amd.workloads import fcc_ordering_ce
    sysm = System(fcc_ordering_ce())
    calc = MonteCalculator("canonical", sysm)
    st = MonteCarloState((6, 6, 6), random_occupation(216, 0.25, 3), {"temperature": 600.0})
    p = calc.make_default_sampling_fixture_params(n_equilibration_passes=20, n_passes=40, sample_period=2)
    p.sampler_names = p.sampler_names + ["clex.formation_energy.sparse_corr"]
    o = calc.run_fixture(st, p)
    eci = np.asarray(sysm.flat.eci)
    nz = np.flatnonzero(eci)
    d = o["statistics"]["clex.formation_energy.sparse_corr"]
    assert d["component_names"] == [str(i) for i in nz] and d["shape"] == [len(nz)] and len(nz) == 6
    mc = np.array([d[k]["mean"][0] for k in d["component_names"]])
    e = o["statistics"]["clex.formation_energy"]["value"]["mean"][0]
    assert float(eci[nz] @ mc) == pytest.approx(e, rel=1e-9, abs=1e-12)


def test_nfold_refuses_corr_sampler(system):
    """The N-fold method's block sums hold energies and compositions only: a
    fixture asking for correlations is refused before any device work."""
    nf = object.__new__(MonteCalculator)
    nf.method, nf.system = "nfold", system
    p = nf.make_default_sampling_fixture_params(n_passes=10)
    p.sampler_names = p.sampler_names + ["corr.formation_energy"]
    st = MonteCarloState((4, 4, 4), np.zeros(64, np.int8), {"temperature": 800.0, "param_chem_pot": [0.0]})
    with pytest.raises(ValueError, match="corr"):
        nf.run_multistate([st], p)
