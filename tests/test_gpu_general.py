"""General (table) path on more than two occupants and more than one
sublattice, the 2-bit packed LDS image, and uneven supercells in one launch.

* Exact enumeration (3 sigma): ternary HCP (2 sublattices, 2x2x1 = 8 sites,
  3^8 states) and quaternary BCC (2x2x2 = 8 sites, 4^8 states), semi-grand
  canonical and canonical, int8 and packed images.
* Uneven ternary HCP supercells in one persistent launch: per chain, the
  FP64 running energy of FP32 dEs versus the oracle's energy of the final
  occupation (drift tolerance 4e-6 eV x sqrt(accepted moves): FP32 rounding
  of dE sums of up to a few hundred table values of O(0.1-1) eV, accumulated
  as a random walk), and after re-anchoring on a full FP64 recompute within
  1e-9 relative; occupant counts conserved.
* The largest shapes of the measured workloads (48x48x32 HCP = 147,456
  sites; 64^3 BCC = 262,144 sites) run with the automatically chosen packed
  image, with the same energy and composition checks.
"""
import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.engine import Engine
from casmcode_clexmonte_amd.system import BasisSet, ClusterExpansion, FlatCE, Supercell, bcc_prim, hcp_prim
from casmcode_clexmonte_amd.workloads import (bcc_quaternary_ce, composition_counts, hcp_ternary_ce,
                                              random_occupation_counts)

pytestmark = pytest.mark.gpu

NSIG = 3.0


def _print_case(name):
    print(f"[case] {name}", flush=True)


def hcp_small_flat(seed=0):
    prim = hcp_prim(1.0, occupants=("A", "B", "C"))
    bs = BasisSet(prim, {2: 1.01, 3: 1.01}, "chebychev")
    eci = np.random.default_rng(seed).normal(0, 0.04, bs.n_functions)
    return FlatCE(ClusterExpansion(bs, eci))


def bcc_small_flat(seed=1):
    prim = bcc_prim(1.0, ("A", "B", "C", "D"))
    bs = BasisSet(prim, {2: 0.87, 3: 1.01}, "chebychev")
    eci = np.random.default_rng(seed).normal(0, 0.04, bs.n_functions)
    return FlatCE(ClusterExpansion(bs, eci))


_EXACT = {}


def _exact(name, flat, dims, kind, T, **kw):
    key = (name, dims, kind, T, repr(sorted(kw.items())))
    if key not in _EXACT:
        _EXACT[key] = O.OracleSystem(flat, dims).enumerate(kind, T, **kw)
    return _EXACT[key]


def _check(E, cnt, exact, keys, burn):
    e = E[:, burn:]
    n = cnt[:, burn:, 1].astype(float)
    per = {"E": e.mean(1), "E2": (e * e).mean(1), "n1": n.mean(1), "n12": (n * n).mean(1)}
    for k in keys:
        m, se = per[k].mean(), per[k].std(ddof=1) / np.sqrt(len(per[k]))
        print(f"{k}: gpu {m:.6f} +- {se:.6f} exact {exact[k]:.6f}")
        assert abs(m - exact[k]) <= NSIG * se + 1e-12, (k, m, exact[k], se)


CASES = {
    "hcp_ternary": (hcp_small_flat, (2, 2, 1), 3, [0.0, 0.02, -0.03], [3, 3, 2]),
    "bcc_quaternary": (bcc_small_flat, (2, 2, 2), 4, [0.0, 0.03, -0.02, 0.01], [2, 2, 2, 2]),
}


@pytest.mark.parametrize("packing", ["int8", "packed2"])
@pytest.mark.parametrize("system", list(CASES))
def test_sgc_exact_enumeration(system, packing):
    _print_case(f"{system} sgc exact enumeration {packing}")
    make, dims, nocc, chem_b, _ = CASES[system]
    flat = make()
    sc = Supercell(flat.ce.prim, dims)
    T = 1000.0
    chem = list(chem_b) * flat.n_sublat
    exact = _exact(system, flat, dims, "sgc", T, chem=chem)
    nch = 128
    eng = Engine(flat)
    eng.set_image_packing(packing)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    rng = np.random.default_rng(5)
    occs = [rng.integers(0, nocc, sc.n_sites).astype(np.int8) for _ in range(nch)]
    eng.load_chains([sid] * nch, occs, [T] * nch, 100 + np.arange(nch), chem=[chem] * nch)
    eng.run_passes("semigrand_canonical", 4000, 1)
    E, cnt = eng.read_samples()
    _check(E, cnt, exact, ["E", "E2", "n1", "n12"], 200)
    eng.close()


@pytest.mark.parametrize("packing", ["int8", "packed2"])
@pytest.mark.parametrize("system", list(CASES))
def test_canonical_exact_enumeration(system, packing):
    _print_case(f"{system} canonical exact enumeration {packing}")
    make, dims, nocc, _, counts = CASES[system]
    flat = make()
    sc = Supercell(flat.ce.prim, dims)
    T = 1000.0
    exact = _exact(system, flat, dims, "canonical", T, fixed_counts=counts)
    nch = 128
    eng = Engine(flat)
    eng.set_image_packing(packing)
    eng.set_global_swaps(4)   # too small for many non-interacting pairs: serial random-pair swaps mix
    sid = eng.add_supercell(sc, n_pair_phases=64, seed=3)
    occs = [random_occupation_counts(counts, 300 + c) for c in range(nch)]
    eng.load_chains([sid] * nch, occs, [T] * nch, 700 + np.arange(nch))
    eng.run_passes("canonical", 5000, 1, canonical_rounds=2)
    E, cnt = eng.read_samples()
    assert np.all(cnt[:, :, :nocc] == np.array(counts))
    _check(E, cnt, exact, ["E", "E2"], 300)
    eng.close()


def _energy_tracking(eng, flat, dims_of_chain, occ0, n_chains):
    refs = []
    for c in range(n_chains):
        st = eng.read_state(c)
        Eref = O.OracleSystem(flat, dims_of_chain[c]).energy(st["occ"])
        refs.append(Eref)
        n_acc = max(1, st["n_accept"])
        tol = 4e-6 * np.sqrt(n_acc) + 1e-9 * abs(Eref)
        print(f"chain {c} {dims_of_chain[c]}: E {st['E']:.9f} oracle {Eref:.9f} accepted {n_acc}")
        assert abs(st["E"] - Eref) <= tol, (c, st["E"], Eref, tol)
        assert np.array_equal(np.bincount(st["occ"], minlength=4), np.bincount(occ0[c], minlength=4))
        assert st["n_accept"] > 0
    drift = eng.reanchor_energies()
    print(f"largest drift before re-anchoring: {drift:.3e} eV")
    for c in range(n_chains):
        E = eng.read_state(c)["E"]
        assert abs(E - refs[c]) <= 1e-9 * max(1.0, abs(refs[c])), (c, E, refs[c])


@pytest.mark.parametrize("packing", ["int8", "packed2"])
def test_uneven_hcp_supercells_one_launch(packing):
    _print_case(f"uneven hcp supercells one launch {packing}")
    ce = hcp_ternary_ce()
    flat = FlatCE(ce)
    shapes = [(6, 6, 4), (8, 8, 5), (10, 10, 6), (12, 12, 8)]
    eng = Engine(flat)
    eng.set_image_packing(packing)
    sids = [eng.add_supercell(Supercell(ce.prim, d), n_pair_phases=4, seed=i) for i, d in enumerate(shapes)]
    dims, occs, sid_of = [], [], []
    for c in range(8):
        d = shapes[c % len(shapes)]
        n = 2 * d[0] * d[1] * d[2]
        dims.append(d)
        sid_of.append(sids[c % len(shapes)])
        occs.append(random_occupation_counts(composition_counts(n, (0.4, 0.3, 0.3)), 50 + c))
    temps = np.linspace(600.0, 1800.0, 8)
    eng.load_chains(sid_of, occs, temps, 900 + np.arange(8))
    eng.run_passes("canonical", 12, 1, canonical_rounds=2)
    _energy_tracking(eng, flat, dims, occs, 8)
    eng.close()


def test_largest_shapes_packed_automatically():
    _print_case("largest shapes: hcp 48x48x32 + bcc 64^3")
    ce = hcp_ternary_ce()
    flat = FlatCE(ce)
    eng = Engine(flat)
    shapes = [(48, 48, 32), (16, 16, 10)]
    sids = [eng.add_supercell(Supercell(ce.prim, d), n_pair_phases=2, seed=1) for d in shapes]
    occs = [random_occupation_counts(composition_counts(2 * d[0] * d[1] * d[2], (0.4, 0.3, 0.3)), 7 + i)
            for i, d in enumerate(shapes)]
    eng.load_chains(sids, occs, [900.0, 1500.0], [1, 2])
    eng.run_passes("canonical", 3, 1, canonical_rounds=2)
    assert eng.last_run_slabs() > 1  # the large chain runs as slabs (automatic)
    _energy_tracking(eng, flat, shapes, occs, 2)
    eng.close()

    ce5 = bcc_quaternary_ce()
    flat5 = FlatCE(ce5)
    eng = Engine(flat5)
    sc = Supercell(ce5.prim, (64, 64, 64))
    sid = eng.add_supercell(sc, n_pair_phases=2, seed=2)
    counts = composition_counts(sc.n_sites, (0.25, 0.25, 0.25, 0.25))
    occs = [random_occupation_counts(counts, 11 + i) for i in range(2)]
    eng.load_chains([sid, sid], occs, [600.0, 2000.0], [3, 4])
    eng.run_passes("canonical", 2, 1, canonical_rounds=2)
    assert eng.last_run_slabs() > 1
    _energy_tracking(eng, flat5, [sc.dims] * 2, occs, 2)
    eng.close()
