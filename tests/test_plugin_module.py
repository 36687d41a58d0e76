"""The pybind11 face of the C++ calculator layer (plugin.py /
_clexmc_plugin), checked like the reference's Python calculator tests
(python/tests/canonical/test_CanonicalCalculator.py): construction from the
extern "C" factories with clone semantics, names and defaults on the CPU;
potential, per-proposal deltas, runs and multi-state runs on the GPU against
the oracle.

Tolerances: per_supercell (FP64 correlations) 1e-9 relative; occ_delta (FP32)
1e-5 of the batch scale + 1e-6 eV; composition exact.
"""
import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd import plugin as P
from casmcode_clexmonte_amd.system import FlatCE, Supercell
from casmcode_clexmonte_amd.workloads import fcc_binary_ce, random_occupation


@pytest.fixture(scope="module")
def flat():
    return FlatCE(fcc_binary_ce())


def test_factories_and_clone_on_cpu(flat, tmp_path):
    system = P.system_from_flat(flat, str(tmp_path / "system.flat"))
    assert system.n_sublat == 1 and system.n_corr == flat.n_corr
    np.testing.assert_allclose(system.eci, flat.eci)
    method = P.make_canonical_calculator()
    assert method.calculator_name == "GpuCanonicalCalculator" and not method.is_multistate_method
    names = method.standard_sampling_function_names()
    for k in ("temperature", "mol_composition", "param_composition", "potential_energy"):
        assert k in names
    assert "heat_capacity" in method.standard_analysis_function_names()
    calc = P.MonteCalculator(method, system, {"mol_composition_tol": 1e-6})
    calc2 = P.MonteCalculator(method, system)       # the instance is cloned, not consumed
    assert calc.calculator_name == calc2.calculator_name == "GpuCanonicalCalculator"
    sgc = P.MonteCalculator("semigrand_canonical", system)
    assert sgc.calculator_name == "GpuSemiGrandCanonicalCalculator"
    nf = P.MonteCalculator("nfold", system)
    assert nf.calculator_name == "GpuNfoldCalculator"
    assert P.make_nfold_calculator().calculator_name == "GpuNfoldCalculator"
    with pytest.raises(ValueError):
        P.MonteCalculator("kinetic", system)
    st = P.make_state((4, 4, 4), np.zeros(64, np.int8), {"temperature": 600.0, "param_chem_pot": [0.1]})
    assert st.configuration.L == [4, 4, 4] and st.conditions.vector_values == {"param_chem_pot": [0.1]}
    assert system.mol_composition(st.configuration) == [1.0, 0.0]


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["canonical", "semigrand_canonical"])
def test_plugin_calculator_vs_oracle(flat, tmp_path, method):
    print(f"[case] plugin module {method}", flush=True)
    system = P.system_from_flat(flat, str(tmp_path / "system.flat"))
    calc = P.MonteCalculator(method, system)
    dims = (8, 8, 8)
    sc = Supercell(flat.ce.prim, dims)
    occ = random_occupation(sc.n_sites, 0.5, 31)
    cond = {"temperature": 800.0}
    if method != "canonical":
        cond["param_chem_pot"] = [-0.02]
    state = P.make_state(dims, occ, cond)
    loc = P.OccLocation()
    calc.set_state_and_potential(state, loc)
    assert loc.mol_size() == sc.n_sites
    osys = O.OracleSystem(flat, dims)
    pot = calc.potential
    E = osys.energy(occ)
    # semi-grand canonical potential (SemiGrandCanonicalCalculator.cc:171-213):
    # Omega = E - n_cells mu x with x = n_B / n_cells, so Omega = E - mu n_B and
    # a change of n_B by dn adds -mu dn
    mu = cond["param_chem_pot"][0] if method != "canonical" else 0.0
    Omega = E - mu * float(occ.sum())
    assert abs(pot.per_supercell() - Omega) <= 1e-9 * max(1.0, abs(E))
    rng = np.random.default_rng(2)
    dE, ref = [], []
    for _ in range(20):
        s = int(rng.integers(sc.n_sites))
        dE.append(pot.occ_delta_per_supercell([s], [1 - int(occ[s])]))
        dn = 1 if occ[s] == 0 else -1
        ref.append(osys.delta(occ, [s], [1 - int(occ[s])]) - mu * dn)
    dE, ref = np.array(dE), np.array(ref)
    assert np.all(np.abs(dE - ref) <= 1e-5 * np.abs(ref).max() + 1e-6), (dE, ref)
    rm = P.RunManager()
    rm.seed = 7
    prm = calc.potential  # noqa: F841 (potential stays valid across runs)
    params = P.SamplingFixtureParams()
    params.n_equilibration_passes = 20
    params.n_passes = 100
    rm.params = params
    calc.run(state, loc, rm)
    res = rm.results
    assert res.n_accept > 0 and len(res.samplers["potential_energy"]) == 100
    final = np.array(state.configuration.occupation)
    if method == "canonical":
        assert final.sum() == occ.sum()
        assert abs(calc.potential.per_supercell() - osys.energy(final)) <= 1e-9 * max(1.0, abs(E))
    # multi-state run: independent chains of one launch
    states = [P.make_state(dims, random_occupation(sc.n_sites, 0.5, 40 + i), dict(cond, temperature=600.0 + 200 * i))
              for i in range(3)]
    locs = [P.OccLocation() for _ in states]
    rm2 = P.RunManager()
    rm2.seed = 9
    rm2.params = params
    out = calc.run_multistate(states, locs, rm2)
    assert len(out) == 3 and len(rm2.multistate_results) == 3
    for s0, s1 in zip(states, out):
        if method == "canonical":
            assert sum(s1.configuration.occupation) == sum(s0.configuration.occupation)


@pytest.mark.gpu
def test_plugin_nfold_matches_sgc(flat, tmp_path):
    """The C++ N-fold calculator (rejection-free events, residence-time
    weighted statistics) against the C++ semi-grand canonical calculator:
    formation energy and composition means within 4 combined standard errors
    at three chemical potentials (FCC 8^3, T 900 K); one sample per pass."""
    print("[case] plugin module nfold vs sgc", flush=True)
    system = P.system_from_flat(flat, str(tmp_path / "system.flat"))
    dims = (8, 8, 8)
    mus = [-0.06, 0.0, 0.06]

    def run(method, n_passes, seed):
        calc = P.MonteCalculator(method, system)
        states = [P.make_state(dims, random_occupation(512, 0.5, 60 + i, exact=False),
                               {"temperature": 900.0, "param_chem_pot": [mu]}) for i, mu in enumerate(mus)]
        rm = P.RunManager()
        rm.seed = seed
        params = P.SamplingFixtureParams()
        params.n_equilibration_passes = 200
        params.n_passes = n_passes
        rm.params = params
        calc.run_multistate(states, [P.OccLocation() for _ in states], rm)
        return rm.multistate_results
    ref = run("semigrand_canonical", 3000, 11)
    got = run("nfold", 1500, 12)
    for r, g in zip(ref, got):
        assert len(g.samplers["potential_energy"]) == 1500 and g.n_reject == 0 and g.n_accept == 1500 * 512
        for key in ("clex.formation_energy(0)", "param_composition(0)"):
            (ma, ea), (mb, eb) = r.statistics[key], g.statistics[key]
            assert abs(ma - mb) < 4 * np.hypot(ea, eb) + 1e-6, (key, ma, ea, mb, eb)
        for key in ("heat_capacity", "param_susc"):
            assert np.isfinite(g.analysis[key][0]) and g.analysis[key][0] > 0


@pytest.mark.gpu
def test_plugin_nfold_equals_python_nfold(flat, tmp_path):
    """Same state, same chain seed (0) and no equilibration: the C++ and the
    Python N-fold calculators launch the same deterministic events, so their
    residence-weighted means and analysis values agree to rounding.  The
    Python summary is checked against replicated samples on the CPU; this
    ties the C++ one to it."""
    print("[case] plugin nfold vs python nfold, same seed", flush=True)
    from casmcode_clexmonte_amd.calculator import MonteCalculator as PyCalc, MonteCarloState, System
    dims = (8, 8, 8)
    occ = random_occupation(512, 0.5, 77, exact=False)
    cond = {"temperature": 800.0, "param_chem_pot": [0.03]}
    pc = PyCalc("nfold", System(flat.ce))
    py = pc.run_multistate([MonteCarloState(dims, occ.copy(), dict(cond))],
                           pc.make_default_sampling_fixture_params(n_equilibration_passes=0, n_passes=400))[0]
    system = P.system_from_flat(flat, str(tmp_path / "system.flat"))
    calc = P.MonteCalculator("nfold", system)
    state = P.make_state(dims, occ.copy(), cond)
    rm = P.RunManager()
    rm.seed = 0
    params = P.SamplingFixtureParams()
    params.n_equilibration_passes = 0
    params.n_passes = 400
    rm.params = params
    calc.run(state, P.OccLocation(), rm)
    res = rm.results
    st = py["statistics"]
    pairs = [(res.statistics["clex.formation_energy(0)"][0], st["clex.formation_energy"]["value"]["mean"][0]),
             (res.statistics["potential_energy(0)"][0], st["potential_energy"]["value"]["mean"][0]),
             (res.statistics["param_composition(0)"][0], st["param_composition"]["a"]["mean"][0]),
             (res.statistics["mol_composition(1)"][0], st["mol_composition"]["B"]["mean"][0]),
             (res.analysis["heat_capacity"][0], py["analysis"]["heat_capacity"]["value"][0]),
             (res.analysis["param_susc"][0], py["analysis"]["param_susc"]["a,a"][0]),
             (res.analysis["mol_susc"][3], py["analysis"]["mol_susc"]["B,B"][0]),
             (res.analysis["mol_thermochem_susc"][1], py["analysis"]["mol_thermochem_susc"]["S,B"][0]),
             (res.analysis["param_thermochem_susc"][0], py["analysis"]["param_thermochem_susc"]["S,a"][0])]
    for c, p in pairs:
        assert c == pytest.approx(p, rel=1e-9, abs=1e-12), pairs
    assert np.array_equal(np.array(state.configuration.occupation), pc.engine.read_state(0)["occ"])


@pytest.mark.gpu
def test_plugin_nfold_ternary_equals_python_nfold(tmp_path):
    """As above for a ternary expansion (HCP, two composition axes): the C++
    and the Python N-fold calculators, fed the moments of every occupant
    count, agree on every mean and analysis value to rounding."""
    print("[case] plugin nfold vs python nfold, hcp ternary, same seed", flush=True)
    from casmcode_clexmonte_amd.calculator import MonteCalculator as PyCalc, MonteCarloState, System
    from casmcode_clexmonte_amd.system import FlatCE
    from casmcode_clexmonte_amd.workloads import hcp_ternary_ce
    ce = hcp_ternary_ce()
    flat3 = FlatCE(ce)
    dims = (6, 6, 4)
    occ = np.random.default_rng(8).integers(0, 3, 2 * 6 * 6 * 4).astype(np.int8)
    cond = {"temperature": 1400.0, "param_chem_pot": [0.03, -0.02]}
    pc = PyCalc("nfold", System(ce))
    py = pc.run_multistate([MonteCarloState(dims, occ.copy(), dict(cond))],
                           pc.make_default_sampling_fixture_params(n_equilibration_passes=0, n_passes=300))[0]
    system = P.system_from_flat(flat3, str(tmp_path / "system3.flat"))
    calc = P.MonteCalculator("nfold", system)
    state = P.make_state(dims, occ.copy(), cond)
    rm = P.RunManager()
    rm.seed = 0
    params = P.SamplingFixtureParams()
    params.n_equilibration_passes = 0
    params.n_passes = 300
    rm.params = params
    calc.run(state, P.OccLocation(), rm)
    res = rm.results
    st, an = py["statistics"], py["analysis"]
    comps = st["mol_composition"]["component_names"]
    axes = st["param_composition"]["component_names"]
    assert len(axes) == 2 and len(comps) == 3
    pairs = [(res.statistics["clex.formation_energy(0)"][0], st["clex.formation_energy"]["value"]["mean"][0]),
             (res.statistics["potential_energy(0)"][0], st["potential_energy"]["value"]["mean"][0]),
             (res.analysis["heat_capacity"][0], an["heat_capacity"]["value"][0])]
    for i, k in enumerate(axes):
        pairs.append((res.statistics[f"param_composition({i})"][0], st["param_composition"][k]["mean"][0]))
        pairs.append((res.analysis["param_thermochem_susc"][i], an["param_thermochem_susc"][f"S,{k}"][0]))
        for j, q in enumerate(axes):
            pairs.append((res.analysis["param_susc"][j * 2 + i], an["param_susc"][f"{k},{q}"][0]))
    for i, k in enumerate(comps):
        pairs.append((res.statistics[f"mol_composition({i})"][0], st["mol_composition"][k]["mean"][0]))
        for j, q in enumerate(comps):
            pairs.append((res.analysis["mol_susc"][j * 3 + i], an["mol_susc"][f"{k},{q}"][0]))
    for c, p in pairs:
        assert c == pytest.approx(p, rel=1e-9, abs=1e-12), pairs
    assert np.array_equal(np.array(state.configuration.occupation), pc.engine.read_state(0)["occ"])


@pytest.mark.gpu
@pytest.mark.parametrize("method", ["canonical", "semigrand_canonical"])
def test_plugin_corr_sampler(flat, tmp_path, method):
    """corr.formation_energy in the C++ calculators: one row of correlations
    per unit cell at every sample; ECI . corr equals that sample's formation
    energy (1e-9), which holds only when the reads line up with the kernel's
    samples (period 3 after 20 equilibration passes: the first sampled pass
    is the first after equilibration)."""
    print(f"[case] plugin corr sampler {method}", flush=True)
    system = P.system_from_flat(flat, str(tmp_path / "system.flat"))
    calc = P.MonteCalculator(method, system)
    cond = {"temperature": 1000.0}
    if method != "canonical":
        cond["param_chem_pot"] = [0.02]
    state = P.make_state((6, 6, 6), random_occupation(216, 0.5, 5), cond)
    rm = P.RunManager()
    rm.seed = 4
    params = P.SamplingFixtureParams()
    params.sampler_names = ["clex.formation_energy", "potential_energy", "mol_composition", "param_composition",
                            "corr.formation_energy"]
    params.n_equilibration_passes = 20
    params.n_passes = 60
    params.sample_period = 3
    rm.params = params
    calc.run(state, P.OccLocation(), rm)
    res = rm.results
    corr = np.array(res.samplers["corr.formation_energy"])
    e = np.array(res.samplers["clex.formation_energy"])[:, 0]
    assert corr.shape == (20, flat.n_corr) and np.allclose(corr[:, 0], 1.0, atol=1e-12)
    np.testing.assert_allclose(corr @ np.asarray(flat.eci), e, rtol=1e-9, atol=1e-12)
    assert "corr.formation_energy(1)" in res.statistics
    assert "clex.formation_energy.sparse_corr" in P._FACTORIES[method]().standard_sampling_function_names()
    # the sparse sampler: the correlations of the non-zero ECIs (all of them here)
    params.sampler_names = params.sampler_names + ["clex.formation_energy.sparse_corr"]
    rm.params = params
    calc.run(state, P.OccLocation(), rm)
    res = rm.results
    nz = np.flatnonzero(np.asarray(flat.eci))
    sp = np.array(res.samplers["clex.formation_energy.sparse_corr"])
    assert sp.shape == (20, len(nz))
    np.testing.assert_allclose(sp @ np.asarray(flat.eci)[nz], np.array(res.samplers["clex.formation_energy"])[:, 0],
                               rtol=1e-9, atol=1e-12)
    if method == "semigrand_canonical":  # the N-fold method refuses the correlation sampler
        nf = P.MonteCalculator("nfold", system)
        with pytest.raises(RuntimeError, match="energies and compositions only"):
            nf.run(P.make_state((6, 6, 6), random_occupation(216, 0.5, 5), cond), P.OccLocation(), rm)


@pytest.mark.gpu
def test_plugin_multiswaps(flat, tmp_path):
    """The C++ semi-grand canonical calculator with the parameter
    semigrand_canonical_multiswaps = 2 (every pair of single changes of the
    binary expansion: {A->B x2}, {A->B, B->A}, {B->A x2}): composition changes
    only by multiswaps, so the number of B sites keeps its parity while it
    moves toward the chemical potential's side; the running energy equals the
    oracle's energy of the final state (exact counts)."""
    print("[case] plugin multiswaps", flush=True)
    system = P.system_from_flat(flat, str(tmp_path / "system.flat"))
    calc = P.MonteCalculator("semigrand_canonical", system,
                             {"semigrand_canonical_multiswaps": 2, "multiswaps_per_pass": 64})
    occ = random_occupation(216, 0.5, 8)
    n0 = int(occ.sum())
    state = P.make_state((6, 6, 6), occ, {"temperature": 1000.0, "param_chem_pot": [0.3]})
    rm = P.RunManager()
    rm.seed = 3
    params = P.SamplingFixtureParams()
    params.n_equilibration_passes = 0
    params.n_passes = 200
    rm.params = params
    calc.run(state, P.OccLocation(), rm)
    mol = np.array(rm.results.samplers["mol_composition"])
    nB = np.rint(mol[:, 1] * 216).astype(int)
    assert np.all(nB % 2 == n0 % 2) and nB.max() > n0 + 20, (n0, nB.min(), nB.max())
    final = np.array(state.configuration.occupation)
    assert int(final.sum()) % 2 == n0 % 2
    osys = O.OracleSystem(flat, (6, 6, 6))
    E = osys.energy(final)
    assert abs(calc.potential.per_supercell() - (E - 0.3 * float(final.sum()))) <= 1e-9 * max(1.0, abs(E))
    with pytest.raises(RuntimeError, match="single-site"):
        nf = P.MonteCalculator("nfold", system, {"semigrand_canonical_multiswaps": 2})
        nf.run(P.make_state((6, 6, 6), occ, {"temperature": 1000.0, "param_chem_pot": [0.3]}), P.OccLocation(), rm)


def test_no_device_fails_loudly(flat, tmp_path):
    """Without a GPU the C++ calculators raise (no CPU fallback); with one,
    this test has nothing to check."""
    system = P.system_from_flat(flat, str(tmp_path / "system.flat"))
    calc = P.MonteCalculator("canonical", system)
    st = P.make_state((4, 4, 4), random_occupation(64, 0.5, 1), {"temperature": 600.0})
    try:
        calc.set_state_and_potential(st, P.OccLocation())
    except RuntimeError as e:
        assert "clexmc_create failed" in str(e)
        return
    pytest.skip("a HIP device is present")
