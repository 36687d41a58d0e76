"""Cluster expansions and run parameters from CASM project files.

Reads a CASM ``prim.json`` with a CASM ECI file (``*_eci.json``: ``bspecs``,
``orbits`` with prototypes, multiplicities and cluster functions carrying
``linear_function_index`` and ``eci``), optionally with a sparse ECI list
(``[[linear_function_index, eci], ...]``), and builds this package's
:class:`ClusterExpansion` for the GPU path:

* the cluster specs of ``bspecs`` (``periodic_max_length`` with
  ``orbit_branch_specs``; site basis ``OCCUPATION`` or ``CHEBYCHEV``) are
  enumerated again by :class:`BasisSet`;
* every CASM orbit is matched to exactly one enumerated orbit: an element of
  that orbit is the CASM prototype up to a translation, and sizes,
  multiplicities and lengths must agree;
* every CASM cluster function is matched to exactly one function of the
  matched orbit by its terms (``\\phi_{b,f}(s_{i})`` products and the
  ``\\sqrt{1/n}`` prefactor), carried from the CASM prototype's site order to
  the enumerated element's;
* each function's ECI is stored under the matched function.

The result keeps ``casm_to_local`` (CASM linear function index -> index
here) so sparse ECI lists and correlation outputs can be translated.

Run parameters: :func:`states_from_run_params` reads the ``incremental``
state generation of a CASM run-params file (``initial_configuration``
``fixed`` with a diagonal transformation matrix, ``initial_conditions``,
``conditions_increment``, ``n_states``) into :class:`MonteCarloState` objects,
fixing the composition of canonical states as the reference's
``enforce_composition`` does.

Reference files followed: the fixtures
``tests/unit/clexmonte/data/Clex_ZrO_Occ/{prim,formation_energy_eci,
run_params_complete,run_params_sgc_complete}.json`` and
``tests/unit/clexmonte/data/FCC_binary_vacancy/formation_energy{,_sparse}_eci.json``;
ECI layout as read by ``src/casm/clexmonte/system/io/json/System_json_io.cc:453-1006``;
state generation ``include/casm/clexmonte/run/IncrementalConditionsStateGenerator.hh:54-120``
(parsed by ``include/casm/clexmonte/run/io/json/StateGenerator_json_io_impl.hh:19-200``)
and composition fixing ``include/casm/clexmonte/state/enforce_composition.hh:47-110``.
"""
from __future__ import annotations

import json
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np

from .system import BasisSet, ClusterExpansion, CompositionAxes, Prim, canon, parse_casm_function_latex

_BASIS_KINDS = ("occupation", "chebychev")

Site = Tuple[int, int, int, int]


def _read_json(p):
    if isinstance(p, (dict, list)):
        return p
    with open(p) as f:
        return json.load(f)


def _prim(p) -> Prim:
    return p if isinstance(p, Prim) else Prim.from_dict(_read_json(p))


def branch_specs(bspecs: dict) -> Tuple[Dict[int, float], str]:
    """(max length per cluster size >= 2, site basis kind) of a CASM bspecs."""
    cs = bspecs["cluster_specs"]
    if cs.get("method", "periodic_max_length") != "periodic_max_length":
        raise ValueError(f"cluster_specs method {cs.get('method')!r} is not supported")
    branches = {}
    for k, v in cs["params"]["orbit_branch_specs"].items():
        n = int(k)
        if n >= 2:
            branches[n] = float(v["max_length"])
    kind = bspecs["basis_function_specs"]["dof_specs"]["occ"]["site_basis_functions"].lower()
    if kind not in _BASIS_KINDS:
        raise ValueError(f"site basis functions {kind!r} are not supported")
    return branches, kind


def _translate_onto(elem: Sequence[Site], target: Sequence[Site]) -> Optional[List[int]]:
    """perm with elem translated = target[perm[j]] at site j, or None."""
    if len(elem) != len(target):
        return None
    if not elem:
        return []
    e0, t0 = min(elem), min(target)
    d = (t0[1] - e0[1], t0[2] - e0[2], t0[3] - e0[3])
    tr = [(b, i + d[0], j + d[1], k + d[2]) for (b, i, j, k) in elem]
    tgt = [tuple(s) for s in target]
    if sorted(tr) != sorted(tgt):
        return None
    return [tgt.index(s) for s in tr]


def _term_key(terms) -> Tuple:
    return tuple(sorted((round(c, 8), d) for c, d in terms))


def match_basis(bs: BasisSet, casm_orbits: List[dict], tol: float = 1e-4) -> Dict[int, int]:
    """CASM linear function index -> enumerated function index."""
    where = {}
    for oi, orb in enumerate(bs.orbits):
        for ei, el in enumerate(orb.elements):
            where[canon(el)] = (oi, ei)
    used_orbits = set()
    out: Dict[int, int] = {}
    for co in casm_orbits:
        tag = co.get("linear_orbit_index")
        sites = [tuple(int(x) for x in s) for s in co["prototype"]["sites"]]
        key = canon(sites)
        if key not in where:
            raise ValueError(f"CASM orbit {tag} {sites} matches no enumerated orbit")
        oi, ei = where[key]
        orb = bs.orbits[oi]
        if orb.size != len(sites) or orb.mult != int(co["mult"]):
            raise ValueError(f"CASM orbit {tag}: size/multiplicity {len(sites)}/{co['mult']} "
                             f"!= {orb.size}/{orb.mult}")
        p = co["prototype"]
        if len(sites) >= 2 and (abs(orb.max_length - float(p["max_length"])) > tol or
                                abs(orb.min_length - float(p["min_length"])) > tol):
            raise ValueError(f"CASM orbit {tag}: lengths differ")
        if oi in used_orbits:
            raise ValueError(f"CASM orbit {tag}: enumerated orbit {oi} matched twice")
        used_orbits.add(oi)
        # element ei keeps the prototype's site order: our site j sits at CASM site perm[j]
        perm = _translate_onto(orb.elements[ei], sites)
        ours = {_term_key(bs.functions[fi].terms): fi for fi in orb.functions}
        if len(co["cluster_functions"]) != len(orb.functions):
            raise ValueError(f"CASM orbit {tag}: {len(co['cluster_functions'])} functions, "
                             f"{len(orb.functions)} enumerated")
        for cf in co["cluster_functions"]:
            lfi = int(cf["linear_function_index"])
            latex = [v for k, v in cf.items() if k.startswith("\\Phi")][0]
            terms = []
            for c, d in parse_casm_function_latex(latex):
                deco = tuple(d[perm[j]] for j in range(orb.size)) if orb.size else tuple()
                terms.append((c, deco))
            k = _term_key(terms)
            if k not in ours:
                raise ValueError(f"CASM function {lfi} ({latex}) matches no function of orbit {oi}")
            out[lfi] = ours.pop(k)
    if len(used_orbits) != len(bs.orbits):
        raise ValueError(f"{len(bs.orbits) - len(used_orbits)} enumerated orbits have no CASM orbit")
    return out


def clex_from_casm(prim, eci_file, sparse_eci=None) -> ClusterExpansion:
    """ClusterExpansion of a CASM prim and ECI file (dicts or paths); a sparse
    ECI list, if given, replaces the file's ECIs."""
    prim = _prim(prim)
    d = _read_json(eci_file)
    branches, kind = branch_specs(d["bspecs"])
    bs = BasisSet(prim, branches, kind)
    _check_site_functions(bs, d.get("site_functions"))
    mapping = match_basis(bs, d["orbits"])
    eci = np.zeros(bs.n_functions)
    if sparse_eci is None:
        for co in d["orbits"]:
            for cf in co["cluster_functions"]:
                if "eci" in cf:
                    eci[mapping[int(cf["linear_function_index"])]] = float(cf["eci"])
    else:
        for lfi, v in _read_json(sparse_eci):
            if int(lfi) not in mapping:
                raise ValueError(f"sparse ECI index {lfi} is not a function of the basis set")
            eci[mapping[int(lfi)]] = float(v)
    ce = ClusterExpansion(bs, eci)
    ce.casm_to_local = dict(mapping)
    return ce


def _check_site_functions(bs: BasisSet, site_functions) -> None:
    """The file's site basis values must equal the enumerated ones."""
    if not site_functions:
        return
    prim = bs.prim
    for entry in site_functions:
        b = int(entry["sublat"])
        basis = (entry.get("occ") or {}).get("basis")
        if not basis:
            if bs.phi[b].shape[0] != 0:
                raise ValueError(f"sublattice {b}: file has no site functions, {bs.phi[b].shape[0]} enumerated")
            continue
        for name, vals in basis.items():
            f = int(name.split(",")[1].rstrip("}"))
            for occ, v in vals.items():
                o = prim.occupants[b].index(occ)
                if abs(bs.phi[b][f][o] - float(v)) > 1e-8:
                    raise ValueError(f"site function {name} on sublattice {b}: {occ} -> {v}, "
                                     f"enumerated {bs.phi[b][f][o]}")


def load_casm_clex(prim_path: str, eci_path: str, sparse_eci_path: Optional[str] = None) -> ClusterExpansion:
    return clex_from_casm(prim_path, eci_path, sparse_eci_path)


# ---------------------------------------------------------------------------
# run parameters
# ---------------------------------------------------------------------------
def enforce_composition(occ: np.ndarray, axes: CompositionAxes, prim: Prim, n_cells: int,
                        target_mol, rng: np.random.Generator) -> np.ndarray:
    """Change single occupants, one at a time, while a change brings the mean
    number of each component per unit cell closer to ``target_mol`` (by more
    than 1e-3 / n_cells); among the equally best kinds of change, pick one
    with probability proportional to the number of sites it can act on, then
    a uniformly random such site (enforce_composition.hh:47-110)."""
    occ = np.array(occ, dtype=np.int8, copy=True)
    target = np.asarray(target_mol, float)
    dn = 1.0 / n_cells
    tol = dn * 1e-3
    comp = axes.occ_component
    nb = prim.n_sublat
    # kinds of change: (sublattice, from occupant, to occupant)
    kinds = [(b, a, c) for b in range(nb) for a in range(prim.n_occ(b)) for c in range(prim.n_occ(b)) if a != c]
    for _ in range(occ.size + 1):
        mol = axes.mol_composition(occ, n_cells)
        d0 = float(np.linalg.norm(mol - target))
        best, choices = d0, []
        for (b, a, c) in kinds:
            sub = occ[b * n_cells:(b + 1) * n_cells]
            n_a = int(np.count_nonzero(sub == a))
            if n_a == 0:
                continue
            t = mol.copy()
            t[comp[b][a]] -= dn
            t[comp[b][c]] += dn
            dist = float(np.linalg.norm(t - target))
            if dist > d0 - tol:
                continue
            if dist < best - tol:
                best, choices = dist, [(b, a, c, n_a)]
            elif dist < best + tol:
                choices.append((b, a, c, n_a))
        if not choices:
            return occ
        w = np.array([x[3] for x in choices], float)
        b, a, c, _ = choices[int(rng.choice(len(choices), p=w / w.sum()))]
        sites = np.flatnonzero(occ[b * n_cells:(b + 1) * n_cells] == a)
        occ[b * n_cells + int(rng.choice(sites))] = c
    raise RuntimeError("enforce_composition did not converge")


def mol_composition_of_param(axes: CompositionAxes, param_composition) -> np.ndarray:
    """Mean number of each component per unit cell at a parametric
    composition: origin + R x."""
    x = np.asarray(param_composition, float).reshape(-1)
    if x.size != axes.n_independent:
        raise ValueError(f"param_composition has {x.size} values, the axes {axes.n_independent}")
    return axes.origin + (axes.R @ x if axes.n_independent else 0.0)


def _axis_vector(values: Union[dict, Sequence[float]], names: Sequence[str]) -> np.ndarray:
    if isinstance(values, dict):
        return np.array([float(values[n]) for n in names])
    return np.asarray(values, float)


def states_from_run_params(run_params, system, method: str, seed: int = 0):
    """MonteCarloState list of a CASM ``incremental`` state generation.

    Conditions per state i: initial + i * increment of ``temperature`` and of
    ``mol_composition`` (canonical; a dict by component name) or
    ``param_composition`` (canonical; a dict by axis name ``a``, ``b``, ...)
    or ``param_chem_pot`` (semi-grand canonical; by axis name).  ``dependent_runs`` true is refused: every state starts from
    the fixed initial configuration (all sites holding their first occupant),
    composition-fixed for canonical states."""
    from .calculator import MonteCarloState

    rp = _read_json(run_params)
    sg = rp["state_generation"]
    if sg.get("method") != "incremental":
        raise ValueError(f"state generation method {sg.get('method')!r} is not supported")
    kw = sg["kwargs"]
    if kw.get("dependent_runs", False):
        raise ValueError("dependent_runs: true is not supported (states are run as independent chains)")
    if kw.get("modifiers"):
        raise ValueError("state modifiers are not supported")
    ic = kw["initial_configuration"]
    if ic.get("method") != "fixed":
        raise ValueError(f"initial configuration method {ic.get('method')!r} is not supported")
    T = np.asarray(ic["kwargs"]["transformation_matrix_to_supercell"], int)
    if np.count_nonzero(T - np.diag(np.diag(T))):
        raise ValueError("only diagonal transformation matrices are supported")
    dims = tuple(int(x) for x in np.diag(T))
    n_cells = int(np.prod(dims))
    prim = system.prim
    axes = system.composition_axes
    comps = axes.components
    axis_names = [chr(ord("a") + i) for i in range(axes.n_independent)]
    c0, dc = kw["initial_conditions"], kw.get("conditions_increment", {})
    rng = np.random.default_rng(seed)
    states = []
    for i in range(int(kw["n_states"])):
        cond = {"temperature": float(c0["temperature"]) + i * float(dc.get("temperature", 0.0))}
        occ = np.zeros(prim.n_sublat * n_cells, dtype=np.int8)
        if method == "canonical":
            if "mol_composition" in c0:
                mol = _axis_vector(c0["mol_composition"], comps)
                if "mol_composition" in dc:
                    mol = mol + i * _axis_vector(dc["mol_composition"], comps)
            elif "param_composition" in c0:
                x = _axis_vector(c0["param_composition"], axis_names)
                if "param_composition" in dc:
                    x = x + i * _axis_vector(dc["param_composition"], axis_names)
                mol = mol_composition_of_param(axes, x)
            else:
                raise ValueError("canonical run parameters need mol_composition or param_composition")
            occ = enforce_composition(occ, axes, prim, n_cells, mol, rng)
            cond["mol_composition"] = axes.mol_composition(occ, n_cells)
            cond["param_composition"] = axes.param_composition(cond["mol_composition"])
        elif method in ("semigrand_canonical", "nfold"):  # the N-fold way takes semi-grand canonical conditions
            mu = _axis_vector(c0["param_chem_pot"], axis_names)
            if "param_chem_pot" in dc:
                mu = mu + i * _axis_vector(dc["param_chem_pot"], axis_names)
            cond["param_chem_pot"] = mu
        else:
            raise ValueError(f"unknown method {method!r}")
        states.append(MonteCarloState(dims, occ, cond))
    return states


def write_flat_from_casm(prim_path, eci_path, out_path: str, sparse_eci_path=None) -> None:
    """The flattened tables of a CASM project's cluster expansion, written
    for the C++ calculators (``CASM::clexmonte::gpu::System::from_flat_file``):
    the route a C++ host takes from ``prim.json`` and ``*_eci.json`` to the
    GPU path (the reference builds its System from the same files,
    src/casm/clexmonte/system/io/json/System_json_io.cc:453-539)."""
    from .system import FlatCE, save_flat
    save_flat(FlatCE(load_casm_clex(prim_path, eci_path, sparse_eci_path)), out_path)


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(
        prog="python -m casmcode_clexmonte_amd.casm_io",
        description="Write the flat cluster-expansion file the C++ calculators read "
                    "(System::from_flat_file) from a CASM prim.json and ECI file.")
    ap.add_argument("--prim", required=True, help="CASM prim.json")
    ap.add_argument("--eci", required=True, help="CASM ECI file (bspecs and orbits with cluster functions)")
    ap.add_argument("--sparse-eci", default=None, help="optional [[linear_function_index, eci], ...] list")
    ap.add_argument("--out", required=True, help="flat file to write")
    a = ap.parse_args(argv)
    write_flat_from_casm(a.prim, a.eci, a.out, a.sparse_eci)
    print(f"wrote {a.out}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
