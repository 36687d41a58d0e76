"""CPU tests: pin the oracle and the host-side system builder.

Pins used (each chosen so that a plausible mistake fails one of them):
* the reference's FCC_binary_vacancy fixture (orbits, multiplicities, cluster
  function terms incl. the sqrt(1/2) symmetrised decorations, corr_size 9);
* neighbour-shell counts of FCC / BCC / HCP (closed forms);
* closed-form correlations of pure A / pure B for the Chebyshev basis;
* dE == E(after) - E(before) by full recompute (flip, swap, neighbour swap);
* translation and point-group invariance of the correlations;
* exact enumeration: infinite-temperature limits and a detailed-balance
  chi-square of the oracle's Metropolis histogram;
* Philox4x32-10 and mt19937_64 known-answer vectors;
* the analysis functions on series with known variance / autocorrelation.
"""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd import analysis as A
from casmcode_clexmonte_amd import engine as ENG
from casmcode_clexmonte_amd.system import (BasisSet, ClusterExpansion, FlatCE, Prim, Supercell, bcc_prim,
                                           check_supercell, color_classes, decode_tau, factor_group, fcc_prim,
                                           hcp_prim, interaction_offsets, site_basis_functions)
from casmcode_clexmonte_amd.workloads import fcc_binary_basis, fcc_binary_ce, random_occupation

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


# ---------------------------------------------------------------------------
# reference fixture
# ---------------------------------------------------------------------------
def test_fcc_binary_vacancy_fixture_functions():
    g = json.load(open(os.path.join(GOLDEN, "fcc_binary_vacancy_basis.json")))
    prim = Prim.from_dict(g["prim"])
    bs = BasisSet(prim, {int(k): v for k, v in g["max_length"].items()}, g["site_basis"])
    assert bs.n_functions == g["corr_size"] == 9
    np.testing.assert_allclose(bs.phi[0], np.array(g["site_functions"]))
    assert len(bs.orbits) == len(g["orbits"])
    for orb, ref in zip(bs.orbits, g["orbits"]):
        assert orb.mult == ref["mult"]
        assert orb.size == len(ref["prototype_sites"])
        assert abs(orb.max_length - ref["max_length"]) < 1e-6
        assert len(orb.functions) == len(ref["functions"])
        for fi, rf in zip(orb.functions, ref["functions"]):
            assert fi == rf["linear_function_index"]
            mine = sorted((round(c, 6), tuple(d)) for c, d in bs.functions[fi].terms)
            theirs = sorted((round(c, 6), tuple(f for _, f in sorted(t))) for c, t in rf["terms"])
            assert mine == theirs, (fi, mine, theirs)


def test_fixture_flower_counts():
    """Nearest-neighbour pair orbit of the fixture: 12 pairs in the site
    flower with weight 1/6 (FCC_binary_vacancy_Clexulator_default.cc:923-930)."""
    g = json.load(open(os.path.join(GOLDEN, "fcc_binary_vacancy_basis.json")))
    prim = Prim.from_dict(g["prim"])
    bs = BasisSet(prim, {2: 4.1}, "occupation")
    eci = np.zeros(bs.n_functions)
    for k, v in g["eci"].items():
        eci[int(k)] = v
    flat = FlatCE(ClusterExpansion(bs, eci))
    counts = np.bincount(flat.entry_orbit, minlength=flat.n_orbits)
    assert counts[2] == 12 and counts[3] == 6 and counts[1] == 1
    # neighbourhood: the fixture's nlist has 19 sites (self + 12 + 6)
    assert len(interaction_offsets(flat)[0]) + 1 == g["nlist_size"]


# ---------------------------------------------------------------------------
# lattices and shells
# ---------------------------------------------------------------------------
def _shell_counts(prim, rmax):
    """Neighbours of a sublattice-0 site per distance shell (orbits of equal
    length merged: in HCP the in-plane and out-of-plane nearest neighbours
    are distinct orbits at the same distance)."""
    bs = BasisSet(prim, {2: rmax}, "chebychev")
    shells = {}
    for o in bs.orbits:
        if o.size == 2:
            n = sum(1 for e in o.elements for s in e if s[0] == 0)
            key = round(o.max_length, 5)
            shells[key] = shells.get(key, 0) + n
    return [shells[k] for k in sorted(shells)]


def test_shell_counts_fcc_bcc_hcp():
    assert _shell_counts(fcc_prim(), 1.5) == [12, 6, 24, 12]
    assert _shell_counts(bcc_prim(), 2.3) == [8, 6, 12, 24, 8, 6, 24, 24]
    assert _shell_counts(hcp_prim(), 2.05) == [12, 6, 2, 18, 12, 6]
    assert len(factor_group(fcc_prim())) == 48
    assert len(factor_group(bcc_prim())) == 48
    assert len(factor_group(hcp_prim())) == 24


def test_workload_ce_structure():
    bs = fcc_binary_basis()
    flat = FlatCE(ClusterExpansion(bs, np.ones(bs.n_functions)))
    assert bs.n_functions == 8
    assert flat.n_entries == 115  # 1 + 54 pairs + 24 + 36 triplets
    assert len(interaction_offsets(flat)[0]) == 54
    assert flat.binary


def test_chebychev_site_basis_orthonormal():
    for n in (2, 3, 4):
        phi = site_basis_functions(n, "chebychev")
        G = np.vstack([np.ones(n), phi])
        np.testing.assert_allclose(G @ G.T / n, np.eye(n), atol=1e-12)
    np.testing.assert_allclose(site_basis_functions(2, "chebychev"), [[-1.0, 1.0]])


# ---------------------------------------------------------------------------
# oracle pins
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def fcc():
    ce = fcc_binary_ce()
    return FlatCE(ce)


def ternary_flat(seed=3):
    prim = fcc_prim(1.0, ("A", "B", "C"))
    bs = BasisSet(prim, {2: 1.05, 3: 0.75}, "chebychev")
    return FlatCE(ClusterExpansion(bs, np.random.default_rng(seed).normal(0, 0.05, bs.n_functions)))


def test_pure_phase_correlations(fcc):
    osys = O.OracleSystem(fcc, (4, 4, 4))
    nc = 64
    k = fcc.orbit_k
    cA = osys.corr(np.zeros(osys.n_sites, np.int8)) / nc
    cB = osys.corr(np.ones(osys.n_sites, np.int8)) / nc
    np.testing.assert_allclose(cA, (-1.0) ** k, atol=1e-12)
    np.testing.assert_allclose(cB, np.ones_like(cB), atol=1e-12)


@pytest.mark.parametrize("which", ["binary", "ternary"])
def test_delta_equals_energy_difference(fcc, which):
    flat = fcc if which == "binary" else ternary_flat()
    nocc = int(flat.n_occ[0])
    osys = O.OracleSystem(flat, (5, 5, 6))
    rng = np.random.default_rng(2)
    occ = rng.integers(0, nocc, osys.n_sites).astype(np.int8)
    E0 = osys.energy(occ)
    for _ in range(20):
        n = int(rng.integers(1, 3))
        sites = list(rng.choice(osys.n_sites, n, replace=False))
        if n == 2 and rng.random() < 0.5:
            sites[1] = (sites[0] + 1) % osys.n_sites  # neighbouring changes
        new = [int((occ[s] + 1 + rng.integers(0, nocc - 1)) % nocc) for s in sites]
        o2 = occ.copy()
        for s, v in zip(sites, new):
            o2[s] = v
        d = osys.delta(occ, sites, new)
        assert abs(d - (osys.energy(o2) - E0)) < 1e-12 * max(1, abs(E0))
        dp = O.lib().orc_delta_prepared
        dp.restype = ctypes.c_double
        s_a = np.array(sites, np.int32)
        n_a = np.array(new, np.int32)
        oc = occ.copy()
        v = dp(ctypes.byref(osys.s), oc.ctypes.data_as(O.i8p), ctypes.c_int(n), s_a.ctypes.data_as(O.i32p),
               n_a.ctypes.data_as(O.i32p))
        assert abs(v - d) < 1e-12


def test_translation_and_symmetry_invariance(fcc):
    dims = (4, 4, 4)
    osys = O.OracleSystem(fcc, dims)
    sc = Supercell(fcc.ce.prim, dims)
    rng = np.random.default_rng(4)
    occ = rng.integers(0, 2, sc.n_sites).astype(np.int8)
    c0 = osys.corr(occ)
    # translation by (1, 2, 3)
    occ_t = np.zeros_like(occ)
    for l in range(sc.n_sites):
        b, i, j, k = sc.site(l)
        occ_t[sc.linear_index(b, i + 1, j + 2, k + 3)] = occ[l]
    np.testing.assert_allclose(osys.corr(occ_t), c0, atol=1e-10)
    # a point-group operation that maps the cubic supercell to itself
    ops = factor_group(fcc.ce.prim)
    for op in ops[:12]:
        occ_g = np.zeros_like(occ)
        for l in range(sc.n_sites):
            b2, i2, j2, k2 = op.apply(sc.site(l))
            occ_g[sc.linear_index(b2, i2, j2, k2)] = occ[l]
        np.testing.assert_allclose(osys.corr(occ_g), c0, atol=1e-10)


def test_flower_tables_match_oracle_delta():
    """The product's flattened flower tables V (host-built) reproduce the
    oracle's dE for random single-site changes (binary and ternary)."""
    for flat in (FlatCE(fcc_binary_ce()), ternary_flat()):
        dims = (6, 5, 4)
        sc = Supercell(flat.ce.prim, dims)
        osys = O.OracleSystem(flat, dims)
        rng = np.random.default_rng(9)
        nocc = int(flat.n_occ[0])
        occ = rng.integers(0, nocc, sc.n_sites).astype(np.int8)
        for _ in range(30):
            l = int(rng.integers(sc.n_sites))
            b, i, j, k = sc.site(l)
            o = int(occ[l])
            nw = int((o + 1 + rng.integers(0, nocc - 1)) % nocc)
            G = np.zeros(flat.n_func_max)
            for e in range(flat.flower_begin[b], flat.flower_begin[b + 1]):
                idx, mul = 0, 1
                for q in range(flat.entry_site_begin[e], flat.entry_site_begin[e + 1]):
                    bq, di, dj, dk = flat.entry_sites[q]
                    idx += int(occ[sc.linear_index(bq, i + di, j + dj, k + dk)]) * mul
                    mul *= int(flat.n_occ[bq])
                nf = int(flat.table_nf[flat.entry_table_id[e]])
                G[:nf] += flat.flower_tables[flat.entry_table[e] + idx * nf: flat.entry_table[e] + (idx + 1) * nf]
            d = float(np.dot(flat.site_func[b, :, nw] - flat.site_func[b, :, o], G))
            assert abs(d - osys.delta(occ, [l], [nw])) < 1e-12


def test_independent_sets_and_pair_phases(fcc):
    for dims in ((6, 6, 6), (4, 4, 4)):
        sc = Supercell(fcc.ce.prim, dims)
        check_supercell(fcc, sc)
        offs = interaction_offsets(fcc)
        nbr = [set() for _ in range(sc.n_sites)]
        for l in range(sc.n_sites):
            b, i, j, k = sc.site(l)
            for (b2, di, dj, dk) in offs[b]:
                t = sc.linear_index(b2, i + di, j + dj, k + dk)
                nbr[l].add(t)
                nbr[t].add(l)
        classes = color_classes(fcc, sc)
        assert sum(len(c) for c in classes) == sc.n_sites
        for c in classes:
            s = set(int(x) for x in c)
            for l in s:
                assert not (nbr[l] & s)
        for ph in ENG.pair_phases(fcc, sc, 6, seed=1):
            sites = [int(x) for x in ph.ravel()]
            assert len(set(sites)) == len(sites)
            for a, b in ph:
                assert int(b) not in nbr[int(a)]
            owner = {}
            for p, (a, b) in enumerate(ph):
                owner[int(a)] = p
                owner[int(b)] = p
            for p, (a, b) in enumerate(ph):
                for s in (int(a), int(b)):
                    for t in nbr[s]:
                        assert owner.get(t, p) == p


# ---------------------------------------------------------------------------
# exact enumeration and the oracle's Metropolis
# ---------------------------------------------------------------------------
def small_flat():
    prim = fcc_prim()
    bs = BasisSet(prim, {2: 1.05, 3: 0.75}, "chebychev")
    return FlatCE(ClusterExpansion(bs, np.random.default_rng(0).normal(0, 0.05, bs.n_functions)))


def test_enumeration_infinite_temperature():
    flat = small_flat()
    osys = O.OracleSystem(flat, (2, 2, 2))
    r = osys.enumerate("sgc", 1e12, chem=[0, 0], want_probs=True)
    assert abs(r["n1"] - 4.0) < 1e-6
    assert abs(r["n12"] - (4.0 + 2.0 * 4.0 * 3.0 / 1.0 * 0 + 4.0 * 4.0 + 0)) < 20  # loose sanity
    np.testing.assert_allclose(r["probs"], np.full(256, 1 / 256), rtol=1e-6)
    rc = osys.enumerate("canonical", 1e12, fixed_counts=[4, 4])
    assert abs(rc["n1"] - 4.0) < 1e-12


def test_oracle_metropolis_detailed_balance_chi2():
    """Histogram of states visited by the oracle's reference-loop
    Metropolis versus exact Boltzmann weights (chi-square, 8-site cell)."""
    flat = small_flat()
    dims = (2, 2, 2)
    osys = O.OracleSystem(flat, dims)
    T, mu = 1500.0, 0.02
    ex = osys.enumerate("sgc", T, chem=[0, mu], want_probs=True)
    # per-pass samples carry E only; reconstruct the state histogram with
    # single-pass runs chained through the returned occupation
    rng_occ = np.zeros(8, np.int8)
    counts = np.zeros(256)
    occ = rng_occ
    for it in range(3000):
        r = osys.run(occ, "sgc", T, 0, 1, 1000 + it, chem=[0, mu], full_recompute_samplers=False)
        occ = r["occ"]
        st = int(sum(int(v) << i for i, v in enumerate(occ)))
        counts[st] += 1
    n = counts.sum()
    expct = ex["probs"] * n
    mask = expct > 5
    chi2 = float(np.sum((counts[mask] - expct[mask]) ** 2 / expct[mask]))
    dof = int(mask.sum()) - 1
    # 5-sigma upper bound of chi2 with dof degrees of freedom
    assert chi2 < dof + 5 * math.sqrt(2 * dof), (chi2, dof)


def _canonical_histogram(osys, occ0, T, n_samples, thin, seed0):
    """States visited by the oracle's canonical reference loop, one sample
    every `thin` passes (each pass a separate run continuing the last)."""
    counts = {}
    occ = occ0
    for it in range(n_samples):
        r = osys.run(occ, "canonical", T, 0, thin, seed0 + it, full_recompute_samplers=False)
        occ = r["occ"]
        st = int(sum(int(v) << i for i, v in enumerate(occ)))
        counts[st] = counts.get(st, 0) + 1
    return counts


def test_oracle_canonical_detailed_balance_chi2():
    """The oracle's canonical loop (swap proposals among the candidate lists,
    sequential-flip swap energy, Metropolis acceptance;
    CanonicalCalculator.cc:78-88) visits the states of FCC 2x2x2 at 3 B of 8
    (56 states) in proportion to their exact Boltzmann weights at 400 K
    (chi-square, one sample per 6 passes), and the same histogram rejects
    the weights of 1.2 T (power check)."""
    flat = small_flat()
    dims = (2, 2, 2)
    osys = O.OracleSystem(flat, dims)
    T, nb = 400.0, 3
    ex = osys.enumerate("canonical", T, fixed_counts=[8 - nb, nb], want_probs=True)
    occ0 = np.array([1, 1, 1, 0, 0, 0, 0, 0], np.int8)
    counts = _canonical_histogram(osys, occ0, T, 20000, 6, 5000)
    assert all(bin(st).count("1") == nb for st in counts)
    n = sum(counts.values())

    def chi2_of(probs):
        cnt = np.array([counts.get(st, 0) for st in range(256)], dtype=float)
        expct = probs * n
        mask = expct > 5
        return float(np.sum((cnt[mask] - expct[mask]) ** 2 / expct[mask])), int(mask.sum()) - 1

    chi2, dof = chi2_of(ex["probs"])
    print(f"canonical chi2 {chi2:.1f} dof {dof}")
    assert dof >= 50
    assert chi2 < dof + 5 * math.sqrt(2 * dof), (chi2, dof)
    hot = osys.enumerate("canonical", 1.2 * T, fixed_counts=[8 - nb, nb], want_probs=True)
    chi2_hot, _ = chi2_of(hot["probs"])
    print(f"against 1.2 T: chi2 {chi2_hot:.1f}")
    assert chi2_hot > dof + 10 * math.sqrt(2 * dof), chi2_hot


def test_l12_correlations_closed_form():
    """L1_2 (x_B = 1/4: B on the cube corners, A on the face centres of the
    conventional cell), Chebyshev s = -1 (A), +1 (B), correlations per unit
    cell counted by hand:
    * point: (1 - 3) / 4 = -1/2;
    * 1st-neighbour pairs: each B has 12 A neighbours (12 B-A of the 24 bonds
      per conventional cell), the other 12 bonds are A-A: 0;
    * 2nd-neighbour pairs (a <100>): corner to corner, face centre to a face
      centre of the same kind: all alike, +1;
    * 3rd-neighbour pairs (a <1 1/2 1/2>): each B has 24 A; 24 B-A of 48
      bonds, the rest A-A: 0;
    * 4th-neighbour pairs (a <110>): all alike, +1;
    * nearest-neighbour triangles (8 per site, 32 per cell): each B is in
      24, all A-A-B (B has no B neighbour), the other 8 are A-A-A:
      (24 - 8) / 32 = +1/2."""
    from casmcode_clexmonte_amd.workloads import fcc_binary_basis, l12_occupation
    bs = fcc_binary_basis()
    flat = FlatCE(ClusterExpansion(bs, np.zeros(bs.n_functions)))
    sc = Supercell(bs.prim, (8, 8, 8))
    occ = l12_occupation(sc)
    assert occ.sum() * 4 == sc.n_sites
    corr = O.OracleSystem(flat, sc.dims).corr(occ) / sc.n_cells
    pairs = sorted((o for o in bs.orbits if o.size == 2), key=lambda o: o.max_length)
    tri = sorted((o for o in bs.orbits if o.size == 3), key=lambda o: o.max_length)
    point = [o for o in bs.orbits if o.size == 1][0]
    expect = {point.functions[0]: -0.5, pairs[0].functions[0]: 0.0, pairs[1].functions[0]: 1.0,
              pairs[2].functions[0]: 0.0, pairs[3].functions[0]: 1.0, tri[0].functions[0]: 0.5}
    for f, v in expect.items():
        assert abs(corr[f] - v) <= 1e-12, (f, corr[f], v)


def _b2_occupation(sc):
    """B2 on BCC (prim vectors (-1,1,1)/2, (1,-1,1)/2, (1,1,-1)/2): the cube
    corners and the body centres, i.e. the parity of i + j + k."""
    c = np.arange(sc.n_cells)
    L1, L2 = sc.dims[0], sc.dims[1]
    i, j, k = c % L1, (c // L1) % L2, c // (L1 * L2)
    return ((i + j + k) % 2).astype(np.int8)


@pytest.mark.parametrize("case", ["fcc_ordering_16", "fcc_ordering_32", "bcc_quaternary_64", "hcp_ternary_16x16x10"])
def test_periodic_colourings_independent_and_mixing(case):
    """The periodic colourings: every class is an independent set (no
    interaction offset joins two of its sites), the classes cover the
    variable sites, and on FCC 16^3 / 32^3 (powers of two: no odd residue
    colouring) every class holds a quarter B in perfect L1_2 and on BCC 64^3 every
    class holds half of each B2 sublattice, so canonical swaps inside a class
    can create antisites in these ordered states."""
    from casmcode_clexmonte_amd.system import periodic_colouring
    from casmcode_clexmonte_amd.workloads import (bcc_quaternary_ce, fcc_ordering_ce, hcp_ternary_ce,
                                                  l12_occupation)
    make, dims = {"fcc_ordering_16": (fcc_ordering_ce, (16, 16, 16)), "fcc_ordering_32": (fcc_ordering_ce, (32, 32, 32)),
                  "bcc_quaternary_64": (bcc_quaternary_ce, (64, 64, 64)),
                  "hcp_ternary_16x16x10": (hcp_ternary_ce, (16, 16, 10))}[case]
    ce = make()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, dims)
    classes = color_classes(flat, sc)
    lab = -np.ones(sc.n_sites, dtype=np.int64)
    for q, c in enumerate(classes):
        assert np.all(lab[c] < 0) and np.all(np.diff(c) > 0)
        lab[c] = q
    var = np.concatenate([np.arange(b * sc.n_cells, (b + 1) * sc.n_cells) for b in range(ce.prim.n_sublat)
                          if ce.prim.variable(b)])
    assert np.all(lab[var] >= 0)
    L1, L2, L3 = sc.dims
    cells = np.arange(sc.n_cells)
    i, j, k = cells % L1, (cells // L1) % L2, cells // (L1 * L2)
    for b, lst in interaction_offsets(flat).items():
        for (b2, di, dj, dk) in lst:
            t = b2 * sc.n_cells + ((i + di) % L1) + L1 * (((j + dj) % L2) + L2 * ((k + dk) % L3))
            assert not np.any(lab[b * sc.n_cells + cells] == lab[t]), (b, b2, di, dj, dk)
    kind, _ = periodic_colouring(flat, sc.dims)
    if case.startswith("fcc"):
        occ = l12_occupation(sc)
        assert all(abs(occ[c].mean() - 0.25) < 1e-12 for c in classes)
    elif case.startswith("bcc"):
        occ = _b2_occupation(sc)
        assert all(abs(occ[c].mean() - 0.5) < 1e-12 for c in classes)


def test_oracle_canonical_conserves_composition(fcc):
    osys = O.OracleSystem(fcc, (6, 6, 6))
    occ = random_occupation(osys.n_sites, 0.25, 3)
    r = osys.run(occ, "canonical", 800.0, 2, 5, 11)
    assert int(r["occ"].sum()) == int(occ.sum())
    assert np.all(r["counts"][:, 1] == int(occ.sum()))
    assert abs(r["E_final"] - osys.energy(r["occ"])) < 1e-9


# ---------------------------------------------------------------------------
# RNG known answers
# ---------------------------------------------------------------------------
def test_philox_kat():
    # Random123 known-answer vectors for Philox4x32-10
    assert O.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_mt19937_64_kat():
    # the C++ standard requires the 10000th output of a default-constructed
    # std::mt19937_64 (seed 5489) to be 9981545732273789042
    assert O.mt19937_64_nth(5489, 10000) == 9981545732273789042


# ---------------------------------------------------------------------------
# analysis
# ---------------------------------------------------------------------------
def test_analysis_functions():
    rng = np.random.default_rng(0)
    x = rng.normal(0, 2.0, 200000)
    y = 0.5 * x + rng.normal(0, 1.0, x.size)
    assert abs(A.variance(x) - 4.0) < 0.05
    assert abs(A.covariance(x, y) - 2.0) < 0.05
    T, n = 500.0, 100
    assert abs(A.heat_capacity(x, T, n) - A.variance(x) * n / (A.KB * T * T)) < 1e-12
    # AR(1) with rho: tau_int = (1 + rho) / (2 (1 - rho))
    rho = 0.8
    z = np.zeros(400000)
    e = rng.normal(size=z.size)
    for i in range(1, z.size):
        z[i] = rho * z[i - 1] + e[i]
    t = A.tau_int(z)
    assert abs(t - (1 + rho) / (2 * (1 - rho))) < 0.3
    assert math.isnan(A.variance([]))


# ---------------------------------------------------------------------------
# C ABI library
# ---------------------------------------------------------------------------
def test_library_exports_declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "clexmc.h")).read()
    declared = sorted(set(re.findall(r"\b(clexmc_[a-z_0-9]+)\s*\(", hdr)))
    assert set(declared) == set(ENG.exported_symbols())
    lib = ENG.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert lib.clexmc_version().decode().startswith("clexmc")


def test_engine_without_gpu_fails_loudly():
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from casmcode_clexmonte_amd.engine import Engine, ClexmcError\n"
            "from casmcode_clexmonte_amd.system import FlatCE\n"
            "from casmcode_clexmonte_amd.workloads import fcc_binary_ce\n"
            "try:\n    Engine(FlatCE(fcc_binary_ce()))\nexcept ClexmcError as e:\n    print('raised', e)\n"
            "else:\n    print('created')\n") % ROOT
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = ""
    env["ROCR_VISIBLE_DEVICES"] = ""
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert "raised" in out.stdout, out.stdout + out.stderr


def test_state_energies_follow_enumeration_order():
    """orc_enumerate_energies: state index st holds the energy of the
    occupation with site l = digit l of st (radix 2 here), NaN off the fixed
    counts; its Boltzmann average equals orc_enumerate's."""
    ce = fcc_binary_ce()
    flat = FlatCE(ce)
    osys = O.OracleSystem(flat, (2, 2, 2))
    E = osys.state_energies()
    rng = np.random.default_rng(3)
    for st in rng.integers(0, 256, 12):
        occ = ((int(st) >> np.arange(8)) & 1).astype(np.int8)
        assert E[st] == osys.energy(occ)
    Ec = osys.state_energies(fixed_counts=[5, 3])
    nb = np.array([bin(i).count("1") for i in range(256)])
    assert np.all(np.isnan(Ec) == (nb != 3))
    T = 800.0
    w = np.exp(-(Ec[nb == 3] - Ec[nb == 3].min()) / (8.6173303e-05 * T))
    ex = osys.enumerate("canonical", T, fixed_counts=[5, 3])
    assert abs((w * Ec[nb == 3]).sum() / w.sum() - ex["E"]) < 1e-12
