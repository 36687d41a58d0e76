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
from .completion import CompletionCheckParams, check as completion_check
from .engine import NFOLD_PAIRS, Engine
from .system import KB, ClusterExpansion, CompositionAxes, FlatCE, Supercell

_METHODS = ("canonical", "semigrand_canonical", "nfold")
_SGC = ("semigrand_canonical", "nfold")  # methods with semi-grand canonical conditions


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
    """Sampling fixture: samples every `sample_period` passes; either a fixed
    number of passes (`n_passes`, after `n_equilibration_passes`) or, with
    `completion`, the reference's completion checks (cutoffs, equilibration,
    requested precision; see completion.py)."""
    label: str = "thermo"
    sampler_names: List[str] = field(default_factory=lambda: [
        "clex.formation_energy", "potential_energy", "mol_composition", "param_composition"])
    n_equilibration_passes: int = 100
    n_passes: int = 1000
    sample_period: int = 1
    canonical_rounds: int = 0       # 0: automatic, about one swap attempt per mutable site and pass
    n_pair_phases: int = 32
    completion: Optional[CompletionCheckParams] = None
    # results_io "json" kwargs: run_series writes <output_dir>/<label>/run.<i>/
    # observations.json (every sample of every sampler) and trajectory.json
    # (the configuration at every sample)
    write_observations: bool = False
    write_trajectory: bool = False
    # "pass", or "time" (the nfold method only): samples every sample_period
    # units of N-fold time (passes: expected Metropolis steps per mutable site)
    sample_by: str = "pass"

    @classmethod
    def from_casm(cls, label: str, d: dict, **kw) -> "SamplingFixtureParams":
        """From a CASM sampling fixture (``sampling``: ``sample_by`` "pass",
        linear ``period``, ``quantities``; ``completion_check``), as in the
        reference's run params (``sampling_fixtures.<label>``)."""
        smp = d.get("sampling", {})
        if smp.get("sample_by", "pass") not in ("pass", "time") or smp.get("spacing", "linear") != "linear":
            raise ValueError("only linear sampling by pass or by time is supported")
        kw.setdefault("sample_by", smp.get("sample_by", "pass"))
        names = []
        legacy = {"formation_energy": "clex.formation_energy", "formation_energy_corr": "corr.formation_energy"}
        for q in smp.get("quantities", []):
            names.append(legacy.get(q, q))
        rio = d.get("results_io", {})
        rkw = rio.get("kwargs", {}) if rio.get("method", "json") == "json" else {}
        kw.setdefault("write_observations", bool(rkw.get("write_observations", False)))
        kw.setdefault("write_trajectory", bool(rkw.get("write_trajectory", False)) or
                      bool(smp.get("sample_trajectory", False)))
        per = smp.get("period", 1)
        return cls(label=label, sampler_names=names or cls().sampler_names,
                   sample_period=float(per) if kw["sample_by"] == "time" else int(per), n_equilibration_passes=0,
                   completion=CompletionCheckParams.from_casm(d.get("completion_check", {})), **kw)

    def wants_correlations(self) -> bool:
        """Samplers that read every chain's correlations at each sample."""
        return "corr.formation_energy" in self.sampler_names or \
            "clex.formation_energy.sparse_corr" in self.sampler_names


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
        if c.method in _SGC:
            x = c.system.composition_axes.param_composition(
                c.system.composition_axes.mol_composition(c.engine.read_state(0)["occ"], self._n))
            E -= self._n * float(np.dot(np.asarray(self.state.conditions["param_chem_pot"], float), x))
        return E

    def per_unitcell(self) -> float:
        return self.per_supercell() / self._n

    def occ_delta_per_supercell(self, linear_site_index: Sequence[int], new_occ: Sequence[int]) -> float:
        d = self.calculator.engine.delta_energy_batch(
            0, np.array([list(linear_site_index)]), np.array([list(new_occ)]),
            include_chem=self.calculator.method in _SGC)
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
        # semi-grand canonical multiswaps (the system's
        # "semigrand_canonical_multiswaps": a list of types, each a list of
        # (unit, from, to, count), or the largest number of single swaps to
        # combine); composition then changes through multiswaps only
        msw = self.params.get("semigrand_canonical_multiswaps")
        if msw is not None and method in _SGC:
            if method == "nfold":
                raise ValueError("Error in MonteCalculator: the nfold method takes single-site changes only")
            from .system import make_multiswaps
            types = make_multiswaps(system.flat, int(msw)) if isinstance(msw, int) else msw
            self.engine.set_multiswaps(types, int(self.params.get("multiswaps_per_pass", 0)))

    # -- conditions ------------------------------------------------------------
    def _validate(self, state: MonteCarloState):
        cond = state.conditions
        if "temperature" not in cond:
            raise ValueError("Error in MonteCalculator: missing required condition 'temperature'")
        axes = self.system.composition_axes
        if self.method in _SGC:
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
        if self.method in _SGC:
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
        self._gpass = 0  # passes since the load: the kernels sample after passes g with (g + 1) % period == 0

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
        if self.method == "nfold":
            return self._run_nfold(states, p)
        if p.sample_by != "pass":
            raise ValueError("Error in MonteCalculator: sampling by time needs the nfold method")
        self._load(states, p.n_pair_phases)
        kind = self.method
        t0 = time.perf_counter()
        if p.n_equilibration_passes > 0:
            self.engine.run_passes(kind, p.n_equilibration_passes, 1, p.canonical_rounds)
            self._gpass += p.n_equilibration_passes
            self.engine.clear_samples()
            # sampling starts from energies recomputed in FP64 (table-path CEs accumulate FP32 dE)
            self.engine.reanchor_energies()
        att0 = [self.engine.read_state(c)["n_attempt"] for c in range(len(states))]
        acc0 = [self.engine.read_state(c)["n_accept"] for c in range(len(states))]
        self._corr = [[] for _ in states] if p.wants_correlations() else None
        self._traj = [[] for _ in states] if p.write_trajectory else None
        # passes from the start of sampling to its first sample (the kernels
        # sample after passes g, counted from the load, with (g + 1) % period == 0)
        self._count0 = p.sample_period - (max(p.n_equilibration_passes, 0) % p.sample_period)
        if p.completion is None:
            self._sampled_passes(kind, p.n_passes, p)
            self.engine.reanchor_energies()
            E, cnt = self.engine.read_samples()
            ends = [(E.shape[1], None, self.engine.read_state(c), time.perf_counter() - t0)
                    for c in range(len(states))]
        else:
            E, cnt, ends = self._run_to_completion(states, p, t0)
        out = []
        self.last_observations, self.last_trajectories = [], []
        for c, s in enumerate(states):
            n_s, res, st, elapsed = ends[c]
            s.occupation = st["occ"]
            corr = np.asarray(self._corr[c][:n_s]) if self._corr is not None else None
            out.append(self._summary(s, E[c, :n_s], cnt[c, :n_s], st["n_attempt"] - att0[c],
                                     st["n_accept"] - acc0[c], elapsed, res, p, corr))
            self.last_observations.append(self._observations(s, E[c, :n_s], cnt[c, :n_s], p, corr))
            self.last_trajectories.append(self._trajectory(s, self._traj[c][:n_s], p)
                                          if self._traj is not None else None)
        return out

    def _sampled_passes(self, kind: str, n_passes: int, p: SamplingFixtureParams):
        """n_passes passes sampled every p.sample_period.  With the
        "corr.formation_energy" sampler (sampling_functions.cc:121-135: all
        correlations per unit cell), the launches end at the sampled passes
        (g + 1 a multiple of the period, g counted from the load) and every
        chain's FP64 correlations are read there."""
        per = p.sample_period
        if self._corr is None and getattr(self, "_traj", None) is None:
            self.engine.run_passes(kind, n_passes, per, p.canonical_rounds)
            self._gpass += n_passes
            return
        end = self._gpass + n_passes
        while self._gpass < end:
            due = self._gpass + (per - 1 - self._gpass % per)  # next sampled pass
            stop = min(end, due + 1)
            self.engine.run_passes(kind, stop - self._gpass, per, p.canonical_rounds)
            self._gpass = stop
            if stop == due + 1:
                for c in range(len(self._corr or self._traj)):
                    if self._corr is not None:
                        n = int(np.prod(self.engine.supercells[self.engine.chain_sc[c]].dims))
                        self._corr[c].append(self.engine.correlations(c) / n)
                    if self._traj is not None:
                        self._traj[c].append(self.engine.read_state(c)["occ"].copy())

    def _run_to_completion(self, states, p: SamplingFixtureParams, t0: float):
        cc = p.completion
        if cc.max_count is None and not cc.convergence:
            raise ValueError("Error in MonteCalculator: a completion check needs a count cutoff or "
                             "requested convergence")
        kind = self.method
        n = len(states)
        ends = [None] * n
        count = 0
        hard_cap = int(cc.max_count) if cc.max_count is not None else 10 ** 8
        while any(e is None for e in ends):
            target = min(cc.next_check(count), hard_cap)
            self._sampled_passes(kind, target - count, p)
            count = target
            E, cnt = self.engine.read_samples()
            for c in range(n):
                if ends[c] is not None:
                    continue
                series = self._series(states[c], E[c], cnt[c])
                res = completion_check(series, count, cc)
                if res.complete or count >= hard_cap:
                    self.engine.reanchor_energies()
                    ends[c] = (E.shape[1], res, self.engine.read_state(c), time.perf_counter() - t0)
        E, cnt = self.engine.read_samples()
        return E, cnt, ends

    # -- N-fold way -----------------------------------------------------------------
    def _run_nfold(self, states: Sequence[MonteCarloState], p: SamplingFixtureParams) -> List[dict]:
        """N-fold way runs (the reference's nfold method,
        include/casm/clexmonte/nfold/nfold_impl.hh:18-98, with its rates
        src/casm/clexmonte/nfold/nfold_events.cc:28-62): equilibration by
        p.n_equilibration_passes plain semi-grand canonical passes, then
        p.n_passes passes of mol_size rejection-free events each.  Every visited
        state is weighted by its expected residence time, so the statistics are
        residence-weighted averages; one block of events per pass gives the
        series the precisions are estimated from.  Any number of occupants (the
        events' records hold the moments of every occupant count); all states
        on one supercell shape; fixed pass counts (no completion checks)."""
        flat = self.system.flat
        if p.wants_correlations():
            raise ValueError("Error in MonteCalculator: the nfold method samples energies and compositions only "
                             "(no corr.formation_energy or clex.formation_energy.sparse_corr)")
        if p.completion is not None:
            raise ValueError("Error in MonteCalculator: nfold runs take a fixed number of passes")
        shapes = {tuple(int(x) for x in s.supercell) for s in states}
        if len(shapes) != 1:
            raise ValueError("Error in MonteCalculator: nfold states must share one supercell")
        self._load(states, 0)
        t0 = time.perf_counter()
        if p.n_equilibration_passes > 0:
            self.engine.run_passes("semigrand_canonical", p.n_equilibration_passes, p.n_equilibration_passes)
        n_cells = int(np.prod(next(iter(shapes))))
        n_var = n_cells * sum(1 for b in range(flat.n_sublat) if flat.n_occ[b] >= 2)
        att0 = [self.engine.read_state(c)["n_attempt"] for c in range(len(states))]
        if p.sample_by == "time":
            # samples at times 0, period, ... up to n_passes; every rate is at
            # most 1, so a state lasts at least 1/n_var passes and
            # n_var x (n_passes + period + 1) events always suffice
            n_samples = max(1, int(round(p.n_passes / float(p.sample_period))))
            max_events = int(n_var * (n_samples * float(p.sample_period) + 1.0)) + 1
            if max_events >= 2 ** 31:
                raise ValueError("Error in MonteCalculator: n_passes x mutable sites must lie below 2^31")
            self.engine.clear_samples()
            self.engine.nfold_run_time(n_samples, float(p.sample_period), max_events)
            elapsed = time.perf_counter() - t0
            E, cnt = self.engine.read_samples()
            out = []
            for c, s in enumerate(states):
                st = self.engine.read_state(c)
                s.occupation = st["occ"]
                d = self._summary(s, E[c], cnt[c], st["n_attempt"] - att0[c], st["n_attempt"] - att0[c], elapsed)
                d["completion_check_results"]["N_samples"] = [int(E.shape[1])]
                out.append(d)
            return out
        blocks = self.engine.nfold_run_moments(p.n_passes * n_var, n_var)
        elapsed = time.perf_counter() - t0
        out = []
        for c, s in enumerate(states):
            st = self.engine.read_state(c)
            s.occupation = st["occ"]
            out.append(self._nfold_summary(s, blocks[c], st["n_attempt"] - att0[c], elapsed, p))
        return out

    def _nfold_summary(self, state, blk, n_events, elapsed, p) -> dict:
        """Summary in the reference layout from residence-weighted block
        moments (clexmc_nfold_run_moments: sum w, sum w E, sum w E^2,
        sum w N_o, sum w N_o N_q, sum w E N_o per block).  Every sampled
        quantity q is affine in (E, N_0..N_3) per supercell,
        q = a + b E + sum_o c_o N_o, so its weighted mean and the covariances
        of the analysis functions follow exactly from the weighted moments;
        precisions use the per-block means."""
        axes = self.system.composition_axes
        n = int(np.prod(state.supercell))
        T = float(state.conditions["temperature"])
        mu = np.asarray(state.conditions["param_chem_pot"], float)
        comps = axes.components
        var_names = [chr(ord("a") + i) for i in range(axes.n_independent)]
        Wt = blk[:, 0].sum()
        E1 = blk[:, 1].sum() / Wt
        N1 = blk[:, 3:7].sum(axis=0) / Wt
        NN = np.zeros((4, 4))
        for t, (o, q) in enumerate(NFOLD_PAIRS):
            NN[o, q] = NN[q, o] = blk[:, 7 + t].sum() / Wt
        vEE = blk[:, 2].sum() / Wt - E1 * E1
        vEN = blk[:, 17:21].sum(axis=0) / Wt - E1 * N1
        vNN = NN - np.outer(N1, N1)
        # compositions as affine functions of the occupant counts (sites over
        # the variable sublattices holding each occupant)
        basis = np.vstack([np.zeros(4), np.eye(4)])
        ser0 = self._series(state, np.zeros(5), basis)
        aff = {}  # name -> (a, b, c[4]): q = a + b E + c . N
        for k in comps:
            m = ser0[f"mol_composition.{k}"]
            aff[f"mol_composition.{k}"] = (m[0], 0.0, m[1:] - m[0])
        for k in var_names:
            x = ser0[f"param_composition.{k}"]
            aff[f"param_composition.{k}"] = (x[0], 0.0, x[1:] - x[0])
        xa = np.array([aff[f"param_composition.{k}"][0] for k in var_names])
        xc = np.array([aff[f"param_composition.{k}"][2] for k in var_names]).reshape(len(var_names), 4)
        aff["clex.formation_energy"] = (0.0, 1.0 / n, np.zeros(4))
        aff["potential_energy"] = (-float(mu @ xa), 1.0 / n, -(mu @ xc) if len(var_names) else np.zeros(4))

        def mean(q):
            a, b, c = aff[q]
            return a + b * E1 + c @ N1

        def cov(q, r):
            _, bq, cq = aff[q]
            _, br, cr = aff[r]
            return bq * br * vEE + bq * (cr @ vEN) + br * (cq @ vEN) + cq @ vNN @ cr

        z = 2.0
        Wb = blk[:, 0]
        Eb, Nb = blk[:, 1] / Wb, blk[:, 3:7] / Wb[:, None]

        def stat(q):
            a, b, c = aff[q]
            _, err = A.mean_and_error(a + b * Eb + Nb @ c)
            return {"mean": [float(mean(q))], "calculated_precision": [z * err]}

        stats = {"clex.formation_energy": {"shape": [], "value": stat("clex.formation_energy")},
                 "potential_energy": {"shape": [], "value": stat("potential_energy")}}
        for name, cn in (("mol_composition", comps), ("param_composition", var_names)):
            d = {"component_names": list(cn), "shape": [len(cn)]}
            for k in cn:
                d[k] = stat(f"{name}.{k}")
            stats[name] = d
        kT = KB * T
        analysis = {"heat_capacity": {"shape": [], "value": [cov("potential_energy", "potential_energy") * n / (kT * T)]}}

        def cov_block(an, bn, pa, pb, name):
            d = {"component_names": [f"{i},{j}" for j in bn for i in an], "shape": [len(an), len(bn)]}
            for j in bn:
                for i in an:
                    d[f"{i},{j}"] = [float(cov(f"{pa}{i}", f"{pb}{j}")) * n / kT]
            analysis[name] = d
        cov_block(comps, comps, "mol_composition.", "mol_composition.", "mol_susc")
        cov_block(var_names, var_names, "param_composition.", "param_composition.", "param_susc")
        aff["S"] = aff["potential_energy"]
        cov_block(["S"], comps, "", "mol_composition.", "mol_thermochem_susc")
        cov_block(["S"], var_names, "", "param_composition.", "param_thermochem_susc")
        conditions = {"temperature": {"shape": [], "value": [T]},
                      "param_chem_pot": {"component_names": var_names, "shape": [len(var_names)],
                                         **{k: [float(v)] for k, v in zip(var_names, mu)}}}
        nb = len(blk)
        completion = {"N_samples": [nb], "N_samples_for_statistics": [nb],
                      "N_samples_for_all_to_equilibrate": [0], "all_converged": [True],
                      "all_equilibrated": [True], "acceptance_rate": [1.0],
                      "count": [int(n_events)], "elapsed_clocktime": [elapsed]}
        return {"conditions": conditions, "statistics": stats, "analysis": analysis,
                "completion_check_results": completion}

    # -- results -----------------------------------------------------------------
    def _corr_samplers(self, p: SamplingFixtureParams) -> Dict[str, tuple]:
        """The requested correlation samplers with (local correlation indices,
        component names): "corr.formation_energy", every correlation
        (sampling_functions.cc:121-135), and "clex.formation_energy.sparse_corr",
        the correlations of the non-zero ECIs (sampling_functions.cc:185-219).
        Components are named and ordered by the ECI file's linear function
        index for an expansion read from CASM files (casm_io), else by index."""
        ce = self.system.flat.ce
        eci = self.system.flat.eci
        c2l = getattr(ce, "casm_to_local", None)
        if c2l and len(c2l) == len(eci):
            order = [(int(cf), int(lf)) for cf, lf in sorted(c2l.items())]
        else:
            order = [(i, i) for i in range(len(eci))]
        out = {}
        if "corr.formation_energy" in p.sampler_names:
            out["corr.formation_energy"] = ([lf for _, lf in order], [str(cf) for cf, _ in order])
        if "clex.formation_energy.sparse_corr" in p.sampler_names:
            nz = [(cf, lf) for cf, lf in order if eci[lf] != 0.0]
            out["clex.formation_energy.sparse_corr"] = ([lf for _, lf in nz], [str(cf) for cf, _ in nz])
        return out

    def _observations(self, state, E_sc, counts, p: SamplingFixtureParams, corr=None) -> dict:
        """Every sample of every sampler of one run (the observations file):
        "count" (pass of each sample since sampling began) and, per sampler,
        component names, shape and one row of values per sample."""
        ser = self._series(state, E_sc, counts)
        axes = self.system.composition_axes
        ns = len(E_sc)
        var_names = [chr(ord("a") + i) for i in range(axes.n_independent)]
        out = {"count": [self._count0 + k * p.sample_period for k in range(ns)]}

        def put(name, names, shape, rows):
            out[name] = {"component_names": names, "shape": shape, "value": [[float(v) for v in r] for r in rows]}

        put("clex.formation_energy", [], [], ser["clex.formation_energy"][:, None])
        put("potential_energy", [], [], ser["potential_energy"][:, None])
        put("mol_composition", list(axes.components), [len(axes.components)],
            np.stack([ser[f"mol_composition.{k}"] for k in axes.components], axis=1))
        put("param_composition", var_names, [len(var_names)],
            np.stack([ser[f"param_composition.{k}"] for k in var_names], axis=1) if var_names
            else np.zeros((ns, 0)))
        put("temperature", [], [], np.full((ns, 1), float(state.conditions["temperature"])))
        if self.method in _SGC:
            mu = np.asarray(state.conditions["param_chem_pot"], float)
            put("param_chem_pot", var_names, [len(var_names)], np.tile(mu, (ns, 1)))
        if corr is not None:
            for name, (idx, names) in self._corr_samplers(p).items():
                put(name, names, [len(names)], np.asarray(corr)[:, idx])
        return out

    def _trajectory(self, state, occs, p: SamplingFixtureParams) -> list:
        """The configuration at every sample (the trajectory file): diagonal
        supercell transformation matrix and occupation (site order
        b * n_cells + i + L1 * (j + L2 * k))."""
        L = [int(x) for x in state.supercell]
        T = [[L[0], 0, 0], [0, L[1], 0], [0, 0, L[2]]]
        return [{"count": self._count0 + k * p.sample_period, "transformation_matrix_to_super": T,
                 "dof": {"occ": [int(v) for v in o]}} for k, o in enumerate(occs)]

    def _series(self, state, E_sc, counts) -> Dict[str, np.ndarray]:
        """Per-sample values of the samplers, per unit cell: clex.formation_energy,
        potential_energy, mol_composition.<component>, param_composition.<axis>."""
        axes = self.system.composition_axes
        flat = self.system.flat
        n = int(np.prod(state.supercell))
        comps = axes.components
        mol = np.zeros((len(E_sc), len(comps)))
        for b in range(flat.n_sublat):
            if flat.n_occ[b] < 2:
                mol[:, axes.occ_component[b][0]] += 1.0
        bvar = next(b for b in range(flat.n_sublat) if flat.n_occ[b] >= 2)
        for o, c in enumerate(axes.occ_component[bvar]):
            mol[:, c] += np.asarray(counts)[:, o] / n
        x = np.array([axes.param_composition(m) for m in mol]).reshape(len(E_sc), axes.n_independent)
        e = np.asarray(E_sc, dtype=float) / n
        if self.method in _SGC:
            pe = e - x @ np.asarray(state.conditions["param_chem_pot"], float)
        else:
            pe = e.copy()
        out = {"clex.formation_energy": e, "potential_energy": pe}
        for i, k in enumerate(comps):
            out[f"mol_composition.{k}"] = mol[:, i]
        for i in range(axes.n_independent):
            out[f"param_composition.{chr(ord('a') + i)}"] = x[:, i]
        return out

    def _summary(self, state, E_sc, counts, n_attempt, n_accept, elapsed, res=None,
                 p: Optional[SamplingFixtureParams] = None, corr: Optional[np.ndarray] = None) -> dict:
        axes = self.system.composition_axes
        n = int(np.prod(state.supercell))
        T = float(state.conditions["temperature"])
        comps = axes.components
        ser = self._series(state, E_sc, counts)
        n_eq = res.n_equilibrate if res is not None else 0
        z = p.completion.z() if (p is not None and p.completion is not None) else 2.0
        for k in ser:
            ser[k] = ser[k][n_eq:]
        e, pe = ser["clex.formation_energy"], ser["potential_energy"]
        mol = np.stack([ser[f"mol_composition.{k}"] for k in comps], axis=1)
        var_names = [chr(ord("a") + i) for i in range(axes.n_independent)]
        x = np.stack([ser[f"param_composition.{k}"] for k in var_names], axis=1) if var_names \
            else np.zeros((len(e), 0))
        requested = res.quantities if res is not None else {}

        def stat(name, arr):
            m, err = A.mean_and_error(arr)
            d = {"mean": [m], "calculated_precision": [z * err]}
            if name in requested:
                d["is_converged"] = [requested[name]["is_converged"]]
            return d

        stats = {"clex.formation_energy": {"shape": [], "value": stat("clex.formation_energy", e)},
                 "potential_energy": {"shape": [], "value": stat("potential_energy", pe)}}
        for name, arr, cn in (("mol_composition", mol, comps), ("param_composition", x, var_names)):
            d = {"component_names": list(cn), "shape": [len(cn)]}
            for i, k in enumerate(cn):
                d[k] = stat(f"{name}.{k}", arr[:, i])
            stats[name] = d
        if corr is not None:
            corr = corr[n_eq:]
            for name, (idx, names) in self._corr_samplers(p).items():
                d = {"component_names": names, "shape": [len(names)]}
                for i, k in zip(idx, names):
                    d[k] = stat(f"{name}.{k}", corr[:, i])
                stats[name] = d
        analysis = {"heat_capacity": {"shape": [], "value": [A.heat_capacity(pe, T, n)]}}
        if self.method in _SGC:
            mu = np.asarray(state.conditions["param_chem_pot"], float)

            def cov_block(a, b, an, bn, name):
                d = {"component_names": [f"{i},{j}" for j in bn for i in an], "shape": [len(an), len(bn)]}
                for jj, j in enumerate(bn):
                    for ii, i in enumerate(an):
                        d[f"{i},{j}"] = [A.covariance(a[:, ii], b[:, jj]) * n / (KB * T)]
                analysis[name] = d
            cov_block(mol, mol, comps, comps, "mol_susc")
            cov_block(x, x, var_names, var_names, "param_susc")
            cov_block(pe[:, None], mol, ["S"], comps, "mol_thermochem_susc")
            cov_block(pe[:, None], x, ["S"], var_names, "param_thermochem_susc")
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
        ns = len(E_sc)
        completion = {"N_samples": [ns], "N_samples_for_statistics": [ns - n_eq],
                      "N_samples_for_all_to_equilibrate": [n_eq],
                      "all_converged": [bool(res.all_converged) if res is not None else True],
                      "all_equilibrated": [bool(res.all_equilibrated) if res is not None else True],
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


def run_series(calculator: MonteCalculator, run_params, output_dir: Optional[str] = None, seed: int = 0,
               fixture: str = "thermo") -> dict:
    """Run a CASM run-params series (``state_generation`` incremental,
    ``sampling_fixtures.<fixture>``) on the GPU: every state not yet completed
    is a chain of one launch (runs are independent), each chain stops
    counting at its own completion check.  With `output_dir`, writes
    ``<output_dir>/<fixture>/summary.json`` (one entry per run in every leaf,
    the reference's layout) and ``<output_dir>/completed_runs.json`` (the
    conditions of each completed run) and, when they exist, resumes after the
    runs already completed, as the reference's state generator does with
    ``completed_runs.json`` (python/libcasm/clexmonte/_IncrementalConditionsStateGenerator.py:158-200).
    Returns the merged summary."""
    import json
    import os

    from .casm_io import _read_json, states_from_run_params

    rp = _read_json(run_params)
    states = states_from_run_params(rp, calculator.system, calculator.method, seed=seed)
    fx = rp["sampling_fixtures"][fixture]
    params = SamplingFixtureParams.from_casm(fixture, fx)
    done, old_summary = [], None
    if output_dir is not None:
        cr = os.path.join(output_dir, "completed_runs.json")
        sp = os.path.join(output_dir, fixture, "summary.json")
        if os.path.exists(cr):
            with open(cr) as f:
                done = json.load(f)
            if done and os.path.exists(sp):
                with open(sp) as f:
                    old_summary = json.load(f)
    todo = states[len(done):]
    summaries = calculator.run_multistate(todo, params) if todo else []
    if output_dir is not None and todo and (params.write_observations or params.write_trajectory):
        # <output_dir>/<fixture>/run.<i>/, i counting every run of the series from 1
        for k in range(len(todo)):
            run_dir = os.path.join(output_dir, fixture, f"run.{len(done) + k + 1}")
            os.makedirs(run_dir, exist_ok=True)
            if params.write_observations:
                with open(os.path.join(run_dir, "observations.json"), "w") as f:
                    json.dump(calculator.last_observations[k], f)
            if params.write_trajectory:
                with open(os.path.join(run_dir, "trajectory.json"), "w") as f:
                    json.dump(calculator.last_trajectories[k], f)
    merged = merge_summaries(([old_summary] if old_summary else []) + summaries) if (old_summary or summaries) \
        else {}
    if output_dir is not None:
        os.makedirs(os.path.join(output_dir, fixture), exist_ok=True)
        for s in todo:
            cond = {}
            for k, v in s.conditions.items():
                cond[k] = [float(x) for x in np.atleast_1d(v)] if np.ndim(v) else float(v)
            done.append({"conditions": cond})
        with open(os.path.join(output_dir, fixture, "summary.json"), "w") as f:
            json.dump(merged, f, indent=1)
        with open(os.path.join(output_dir, "completed_runs.json"), "w") as f:
            json.dump(done, f, indent=1)
    return merged
