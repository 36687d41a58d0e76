"""Python face of the GPU Monte Carlo calculators, mirroring the reference's
``libcasm.clexmonte.MonteCalculator`` (python/libcasm/clexmonte/
_MonteCalculator.py:24-90), ``MontePotential``
(python/src/clexmonte_monte_calculator.cpp:694-750) and ``run_fixture``
(:244-269), on top of the clexmc C ABI.

Results use the reference's ``summary.json`` layout (checked by the
reference's python/tests/conftest.py:140-259): ``conditions``, ``statistics``
(per sampler and component: ``mean``, ``calculated_precision``),
``analysis`` (heat capacity, susceptibilities) and
``completion_check_results`` (``N_samples``, ``acceptance_rate``, ``count``,
...), each leaf a list with one entry per run of a series.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import analysis as A
from .engine import Engine
from .system import KB, ClusterExpansion, CompositionAxes, FlatCE, Supercell

_METHODS = ("canonical", "semigrand_canonical")


class System:
    """Cluster expansion + composition axes (the role of
    ``libcasm.clexmonte.System``)."""

    def __init__(self, ce: ClusterExpansion):
        self.ce = ce
        self.prim = ce.prim
        self.flat = FlatCE(ce)
        self.composition_axes = CompositionAxes(self.prim)


@dataclass
class MonteCarloState:
    """Configuration (diagonal supercell + occupation) and conditions
    (``temperature``, ``param_chem_pot``, ``param_composition``,
    ``mol_composition``), as ``libcasm.clexmonte.MonteCarloState``."""
    supercell: Sequence[int]
    occupation: np.ndarray
    conditions: Dict[str, object] = field(default_factory=dict)


@dataclass
class SamplingFixtureParams:
    label: str = "thermo"
    sampler_names: List[str] = field(default_factory=lambda: [
        "clex.formation_energy", "potential_energy", "mol_composition", "param_composition"])
    n_equilibration_passes: int = 100
    n_passes: int = 1000
    sample_period: int = 1
    canonical_rounds: int = 0       # 0: automatic, about one swap attempt per mutable site and pass
    n_pair_phases: int = 32


class MontePotential:
    """Potential of a state evaluated on the GPU (per_supercell,
    per_unitcell, occ_delta_per_supercell)."""

    def __init__(self, calculator: "MonteCalculator", state: MonteCarloState):
        self.calculator = calculator
        self.state = state
        calculator._load([state])
        self._n = int(np.prod(state.supercell))

    def per_supercell(self) -> float:
        c = self.calculator
        corr = c.engine.correlations(0)
        E = float(np.dot(c.system.flat.eci, corr))
        if c.method == "semigrand_canonical":
            x = c.system.composition_axes.param_composition(
                c.system.composition_axes.mol_composition(c.engine.read_state(0)["occ"], self._n))
            E -= self._n * float(np.dot(np.asarray(self.state.conditions["param_chem_pot"], float), x))
        return E

    def per_unitcell(self) -> float:
        return self.per_supercell() / self._n

    def occ_delta_per_supercell(self, linear_site_index: Sequence[int], new_occ: Sequence[int]) -> float:
        d = self.calculator.engine.delta_energy_batch(
            0, np.array([list(linear_site_index)]), np.array([list(new_occ)]),
            include_chem=self.calculator.method == "semigrand_canonical")
        return float(d[0])


class MonteCalculator:
    """``MonteCalculator(method, system, params)`` for the GPU path."""

    def __init__(self, method: str, system: System, params: Optional[dict] = None, device: int = 0):
        if method not in _METHODS:
            raise ValueError(f"Error in MonteCalculator: unknown method '{method}' (use one of {_METHODS})")
        self.method = method
        self.system = system
        self.params = dict(params or {})
        self.engine = Engine(system.flat, device=device)
        self._sc_ids: Dict[tuple, int] = {}
        self.seed = int(self.params.get("seed", 0))

    # -- conditions ------------------------------------------------------------
    def _validate(self, state: MonteCarloState):
        cond = state.conditions
        if "temperature" not in cond:
            raise ValueError("Error in MonteCalculator: missing required condition 'temperature'")
        axes = self.system.composition_axes
        if self.method == "semigrand_canonical":
            mu = cond.get("param_chem_pot")
            if mu is None or len(mu) != axes.n_independent:
                raise ValueError("Error in MonteCalculator: 'param_chem_pot' missing or of the wrong size")
        else:
            n = int(np.prod(state.supercell))
            mol = axes.mol_composition(np.asarray(state.occupation), n)
            tol = float(self.params.get("mol_composition_tol", 1e-5))
            if "mol_composition" in cond and not np.allclose(mol, cond["mol_composition"], atol=tol):
                raise ValueError("Error in MonteCalculator: Invalid initial state: calculated composition is "
                                 "not consistent with conditions composition.")

    def _chem(self, state: MonteCarloState) -> np.ndarray:
        if self.method == "semigrand_canonical":
            return self.system.composition_axes.chem_weights(state.conditions["param_chem_pot"]).ravel()
        return np.zeros(self.system.flat.n_sublat * self.system.flat.n_occ_max)

    def _load(self, states: Sequence[MonteCarloState], n_pair_phases: int = 8):
        for s in states:
            self._validate(s)
        ids = []
        for s in states:
            key = tuple(int(x) for x in s.supercell)
            if key not in self._sc_ids:
                sc = Supercell(self.system.prim, key)
                self._sc_ids[key] = self.engine.add_supercell(
                    sc, n_pair_phases=n_pair_phases if self.method == "canonical" else 0, seed=self.seed)
            ids.append(self._sc_ids[key])
        seeds = [(self.seed << 20) + i for i in range(len(states))]
        self.engine.load_chains(ids, [np.asarray(s.occupation, np.int8) for s in states],
                                [float(s.conditions["temperature"]) for s in states], seeds,
                                chem=np.stack([self._chem(s) for s in states]))

    # -- running -----------------------------------------------------------------
    def make_default_sampling_fixture_params(self, label: str = "thermo", **kw) -> SamplingFixtureParams:
        return SamplingFixtureParams(label=label, **kw)

    def run_fixture(self, state: MonteCarloState, sampling_fixture_params: SamplingFixtureParams) -> dict:
        """Run one state (CanonicalCalculator::run / SemiGrandCanonicalCalculator::run)
        and return its summary (one entry per leaf)."""
        return self.run_multistate([state], sampling_fixture_params)[0]

    def run_multistate(self, states: Sequence[MonteCarloState], p: SamplingFixtureParams) -> List[dict]:
        """Run many states as independent chains of one persistent launch;
        final occupations are written back into the states."""
        self._load(states, p.n_pair_phases)
        kind = self.method
        t0 = time.perf_counter()
        if p.n_equilibration_passes > 0:
            self.engine.run_passes(kind, p.n_equilibration_passes, 1, p.canonical_rounds)
            self.engine.clear_samples()
            # sampling starts from energies recomputed in FP64 (table-path CEs accumulate FP32 dE)
            self.engine.reanchor_energies()
        att0 = [self.engine.read_state(c)["n_attempt"] for c in range(len(states))]
        acc0 = [self.engine.read_state(c)["n_accept"] for c in range(len(states))]
        self.engine.run_passes(kind, p.n_passes, p.sample_period, p.canonical_rounds)
        self.engine.reanchor_energies()
        elapsed = time.perf_counter() - t0
        E, cnt = self.engine.read_samples()
        out = []
        for c, s in enumerate(states):
            st = self.engine.read_state(c)
            s.occupation = st["occ"]
            out.append(self._summary(s, E[c], cnt[c], st["n_attempt"] - att0[c], st["n_accept"] - acc0[c],
                                     elapsed))
        return out

    # -- results -----------------------------------------------------------------
    def _summary(self, state, E_sc, counts, n_attempt, n_accept, elapsed) -> dict:
        axes = self.system.composition_axes
        flat = self.system.flat
        n = int(np.prod(state.supercell))
        T = float(state.conditions["temperature"])
        comps = axes.components
        # mol composition per unit cell from occupant counts (variable
        # sublattices share one occupant list) plus fixed sublattices
        mol = np.zeros((len(E_sc), len(comps)))
        for b in range(flat.n_sublat):
            if flat.n_occ[b] < 2:
                mol[:, axes.occ_component[b][0]] += 1.0
        bvar = next(b for b in range(flat.n_sublat) if flat.n_occ[b] >= 2)
        for o, c in enumerate(axes.occ_component[bvar]):
            mol[:, c] += counts[:, o] / n
        x = np.array([axes.param_composition(m) for m in mol]).reshape(len(E_sc), axes.n_independent)
        e = np.asarray(E_sc) / n
        if self.method == "semigrand_canonical":
            mu = np.asarray(state.conditions["param_chem_pot"], float)
            pe = e - x @ mu
        else:
            pe = e.copy()
        var_names = [chr(ord("a") + i) for i in range(axes.n_independent)]
        samplers = {"clex.formation_energy": (e[:, None], None), "potential_energy": (pe[:, None], None),
                    "mol_composition": (mol, comps), "param_composition": (x, var_names)}
        stats = {}
        for name, (arr, cn) in samplers.items():
            if cn is None:
                m, err = A.mean_and_error(arr[:, 0])
                stats[name] = {"shape": [], "value": {"mean": [m], "calculated_precision": [2.0 * err],
                                                      "is_converged": [True]}}
            else:
                d = {"component_names": list(cn), "shape": [len(cn)]}
                for i, k in enumerate(cn):
                    m, err = A.mean_and_error(arr[:, i])
                    d[k] = {"mean": [m], "calculated_precision": [2.0 * err], "is_converged": [True]}
                stats[name] = d
        analysis = {"heat_capacity": {"shape": [], "value": [A.heat_capacity(pe, T, n)]}}
        if self.method == "semigrand_canonical":
            def cov_block(a, b, an, bn, name, const):
                d = {"component_names": [f"{i},{j}" for j in bn for i in an], "shape": [len(an), len(bn)]}
                for jj, j in enumerate(bn):
                    for ii, i in enumerate(an):
                        d[f"{i},{j}"] = [A.covariance(a[:, ii], b[:, jj]) * n / (KB * T * const)]
                analysis[name] = d
            cov_block(mol, mol, comps, comps, "mol_susc", 1.0)
            cov_block(x, x, var_names, var_names, "param_susc", 1.0)
            cov_block(pe[:, None], mol, ["S"], comps, "mol_thermochem_susc", 1.0)
            cov_block(pe[:, None], x, ["S"], var_names, "param_thermochem_susc", 1.0)
            conditions = {"temperature": {"shape": [], "value": [T]},
                          "param_chem_pot": {"component_names": var_names, "shape": [len(var_names)],
                                             **{k: [float(v)] for k, v in zip(var_names, mu)}}}
        else:
            mol0 = mol.mean(axis=0) if len(mol) else axes.mol_composition(state.occupation, n)
            conditions = {"temperature": {"shape": [], "value": [T]},
                          "mol_composition": {"component_names": list(comps), "shape": [len(comps)],
                                              **{k: [float(v)] for k, v in zip(comps, mol0)}},
                          "param_composition": {"component_names": var_names, "shape": [len(var_names)],
                                                **{k: [float(v)] for k, v in
                                                   zip(var_names, axes.param_composition(mol0))}}}
        ns = len(e)
        completion = {"N_samples": [ns], "N_samples_for_statistics": [ns],
                      "N_samples_for_all_to_equilibrate": [0], "all_converged": [True],
                      "all_equilibrated": [True],
                      "acceptance_rate": [float(n_accept / n_attempt) if n_attempt else float("nan")],
                      "count": [int(n_attempt)], "elapsed_clocktime": [elapsed]}
        return {"conditions": conditions, "statistics": stats, "analysis": analysis,
                "completion_check_results": completion}


def merge_summaries(summaries: Sequence[dict]) -> dict:
    """Concatenate single-run summaries into a series summary (one entry per
    run in every leaf list), as the reference's run_series writes it."""
    def merge(a, b, key=None):
        if isinstance(a, dict):
            return {k: (merge(a[k], b[k], k) if k in b else a[k]) for k in a}
        if isinstance(a, list) and key not in ("shape", "component_names"):
            return list(a) + list(b)
        return a

    out = summaries[0]
    for s in summaries[1:]:
        out = merge(out, s)
    return out
