"""The N-fold way (rejection-free) semi-grand canonical kernel.

* Exact enumeration (FCC binary 2x2x4, 16 sites, T = 1200 K, mu = 0.03 eV):
  residence-weighted averages of E, E^2, N_B, N_B^2, E N_B over 128 chains within
  3 sigma of the Boltzmann averages (error: standard error over chains).
* State bookkeeping: after the events, each chain's running energy equals
  the oracle's energy of its final occupation (exact counts, 1e-9), and the
  occupant counts a following Metropolis pass samples match the occupation.
* Against the plain Metropolis kernel on FCC 12^3 at a low-acceptance point
  (T = 400 K, mu = -0.05 eV): mean energy and composition within 3 sigma.
* The run-time specialised N-fold kernel (one supercell shape; binary FCC
  12^3 and ternary HCP 4x4x3) against the generic one, same states and
  seeds: identical block sums and final states, bit for bit.
* General (table) expansions: ternary HCP 2x2x1 (2 sublattices, 3^8
  states) and quaternary BCC 2x2x2 (4^8 states), T = 1000 K: weighted
  averages of E, E^2, N_1, N_1^2, E N_1 over 128 chains within 3 sigma of exact
  enumeration; final energies equal the oracle's within the FP32
  accumulation bound (4e-6 eV x sqrt(events)); then the per-occupant
  moments (clexmc_nfold_run_moments: <N_o>, <N_o N_q>, <E N_o> for every
  occupant) within 3 sigma of the exact Boltzmann moments over all states;
  occupant counts follow the occupations.
"""
import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.engine import NFOLD_MOMENTS, NFOLD_PAIRS, Engine
from casmcode_clexmonte_amd.system import (BasisSet, ClusterExpansion, FlatCE, Supercell, bcc_prim, fcc_prim,
                                           hcp_prim)
from casmcode_clexmonte_amd.workloads import fcc_binary_ce, random_occupation

pytestmark = pytest.mark.gpu

NSIG = 3.0


def _print_case(name):
    print(f"[case] {name}", flush=True)


def small_flat(seed=0):
    prim = fcc_prim()
    bs = BasisSet(prim, {2: 1.05, 3: 0.75}, "chebychev")
    eci = np.random.default_rng(seed).normal(0, 0.05, bs.n_functions)
    return FlatCE(ClusterExpansion(bs, eci))


def _weighted(blocks, burn_blocks):
    b = blocks[:, burn_blocks:, :].sum(axis=1)
    W = b[:, 0]
    return {"E": b[:, 1] / W, "E2": b[:, 3] / W, "n1": b[:, 2] / W, "n12": b[:, 4] / W, "En1": b[:, 5] / W}


def test_nfold_exact_enumeration():
    _print_case("nfold sgc exact enumeration 2x2x4")
    flat = small_flat()
    sc = Supercell(flat.ce.prim, (2, 2, 4))
    T, mu = 1200.0, 0.03
    chem = [0.0, mu]
    exact = O.OracleSystem(flat, sc.dims).enumerate("sgc", T, chem=chem)
    nch = 128
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    occs = [random_occupation(sc.n_sites, 0.5, 100 + c, exact=False) for c in range(nch)]
    eng.load_chains([sid] * nch, occs, [T] * nch, 1000 + np.arange(nch), chem=[chem] * nch)
    blocks = eng.nfold_run(20000, 500)
    assert blocks.shape == (nch, 40, 6) and np.all(blocks[:, :, 0] > 0)
    per = _weighted(blocks, 2)
    for k in ("E", "E2", "n1", "n12", "En1"):
        m, se = per[k].mean(), per[k].std(ddof=1) / np.sqrt(nch)
        print(f"{k}: nfold {m:.6f} +- {se:.6f} exact {exact[k]:.6f}")
        assert abs(m - exact[k]) <= NSIG * se + 1e-12, (k, m, exact[k], se)
    osys = O.OracleSystem(flat, sc.dims)
    for c in range(0, nch, 16):
        st = eng.read_state(c)
        Eref = osys.energy(st["occ"])
        assert abs(st["E"] - Eref) <= 1e-9 * max(1.0, abs(Eref)), (st["E"], Eref)
    # occupant counts left by the N-fold run: one Metropolis pass samples them
    eng.clear_samples()
    eng.run_passes("semigrand_canonical", 1, 1)
    _, cnt = eng.read_samples()
    for c in range(nch):
        n1 = int(eng.read_state(c)["occ"].sum())
        assert cnt[c, -1, 1] == n1 and cnt[c, -1, 0] == sc.n_sites - n1
    eng.close()


def test_nfold_vs_metropolis_low_acceptance():
    _print_case("nfold vs metropolis fcc 12^3 T=400 K")
    ce = fcc_binary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (12, 12, 12))
    T, mu = 400.0, -0.05
    chem = [0.0, mu]
    nch = 32
    nc = sc.n_cells
    occs = [random_occupation(sc.n_sites, 0.5, 700 + c, exact=False) for c in range(nch)]
    # plain Metropolis
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    eng.load_chains([sid] * nch, occs, [T] * nch, 4000 + np.arange(nch), chem=[chem] * nch)
    eng.run_passes("semigrand_canonical", 3000, 1)
    E, cnt = eng.read_samples()
    att = np.array([eng.read_state(c)["n_attempt"] for c in range(nch)])
    acc = np.array([eng.read_state(c)["n_accept"] for c in range(nch)])
    print(f"metropolis acceptance {acc.sum() / att.sum():.4f}")
    mE = E[:, 1000:].mean(1) / nc
    mX = cnt[:, 1000:, 1].mean(1) / nc
    eng.close()
    # N-fold way from the same starting occupations
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    eng.load_chains([sid] * nch, occs, [T] * nch, 8000 + np.arange(nch), chem=[chem] * nch)
    blocks = eng.nfold_run(60000, 1000)
    steps = blocks[:, :, 0].sum(1) / sc.n_sites
    print(f"nfold: Metropolis-equivalent passes per chain {steps.mean():.0f}")
    # discard the blocks covering the first 1000 Metropolis-equivalent passes
    cum = np.cumsum(blocks[:, :, 0], axis=1) / sc.n_sites
    nE, nX = [], []
    for c in range(nch):
        keep = cum[c] > 1000
        b = blocks[c, keep].sum(0)
        nE.append(b[1] / b[0] / nc)
        nX.append(b[2] / b[0] / nc)
    nE, nX = np.array(nE), np.array(nX)
    for name, g, n in (("E/cell", mE, nE), ("x_B", mX, nX)):
        tol = NSIG * np.sqrt(g.var(ddof=1) / nch + n.var(ddof=1) / nch)
        print(f"{name}: metropolis {g.mean():.6f}  nfold {n.mean():.6f}  tol {tol:.6f}")
        assert abs(g.mean() - n.mean()) <= tol, (name, g.mean(), n.mean(), tol)
    eng.close()


def _hcp_ternary_small(seed=0):
    prim = hcp_prim(1.0, occupants=("A", "B", "C"))
    bs = BasisSet(prim, {2: 1.01, 3: 1.01}, "chebychev")
    return FlatCE(ClusterExpansion(bs, np.random.default_rng(seed).normal(0, 0.04, bs.n_functions)))


def _bcc_quaternary_small(seed=1):
    prim = bcc_prim(1.0, ("A", "B", "C", "D"))
    bs = BasisSet(prim, {2: 0.87, 3: 1.01}, "chebychev")
    return FlatCE(ClusterExpansion(bs, np.random.default_rng(seed).normal(0, 0.04, bs.n_functions)))


@pytest.mark.parametrize("system", ["hcp_ternary", "bcc_quaternary"])
def test_nfold_general_exact_enumeration(system):
    _print_case(f"nfold general path exact enumeration {system}")
    make, dims, nocc, chem_b = {"hcp_ternary": (_hcp_ternary_small, (2, 2, 1), 3, [0.0, 0.02, -0.03]),
                                "bcc_quaternary": (_bcc_quaternary_small, (2, 2, 2), 4, [0.0, 0.03, -0.02, 0.01])}[system]
    flat = make()
    sc = Supercell(flat.ce.prim, dims)
    T = 1000.0
    chem = list(chem_b) * flat.n_sublat
    osys = O.OracleSystem(flat, sc.dims)
    exact = osys.enumerate("sgc", T, chem=chem)
    nch = 128
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    rng = np.random.default_rng(9)
    occs = [rng.integers(0, nocc, sc.n_sites).astype(np.int8) for _ in range(nch)]
    eng.load_chains([sid] * nch, occs, [T] * nch, 2000 + np.arange(nch), chem=[chem] * nch)
    n_events = 20000
    blocks = eng.nfold_run(n_events, 500)
    assert blocks.shape == (nch, 40, 6) and np.all(blocks[:, :, 0] > 0)
    per = _weighted(blocks, 2)
    for k in ("E", "E2", "n1", "n12", "En1"):
        m, se = per[k].mean(), per[k].std(ddof=1) / np.sqrt(nch)
        print(f"{k}: nfold {m:.6f} +- {se:.6f} exact {exact[k]:.6f}")
        assert abs(m - exact[k]) <= NSIG * se + 1e-12, (k, m, exact[k], se)
    for c in range(0, nch, 16):
        st = eng.read_state(c)
        Eref = osys.energy(st["occ"])
        assert abs(st["E"] - Eref) <= 4e-6 * np.sqrt(n_events) + 1e-9 * abs(Eref), (st["E"], Eref)
    # the moments of every occupant count (clexmc_nfold_run_moments) against
    # the exact Boltzmann moments over all states
    mom = eng.nfold_run_moments(n_events, 500)
    assert mom.shape == (nch, 40, NFOLD_MOMENTS) and np.all(mom[:, :, 0] > 0)
    ex = osys.enumerate("sgc", T, chem=chem, want_probs=True)
    pr, Es = ex["probs"], osys.state_energies()
    st_i = np.arange(len(pr))
    occ_s = (st_i[:, None] // (nocc ** np.arange(sc.n_sites))[None, :]) % nocc
    Ns = np.stack([(occ_s == o).sum(1) for o in range(4)], axis=1).astype(float)
    b = mom.sum(axis=1)
    W = b[:, 0]
    checks = [("E", b[:, 1] / W, pr @ Es), ("E2", b[:, 2] / W, pr @ (Es * Es))]
    for o in range(nocc):
        checks.append((f"N{o}", b[:, 3 + o] / W, pr @ Ns[:, o]))
        checks.append((f"E N{o}", b[:, 17 + o] / W, pr @ (Es * Ns[:, o])))
    for t, (o, q) in enumerate(NFOLD_PAIRS):
        if q < nocc:
            checks.append((f"N{o} N{q}", b[:, 7 + t] / W, pr @ (Ns[:, o] * Ns[:, q])))
    for name, per, ref in checks:
        m, se = per.mean(), per.std(ddof=1) / np.sqrt(nch)
        print(f"{name}: nfold {m:.6f} +- {se:.6f} exact {ref:.6f}")
        assert abs(m - ref) <= NSIG * se + 1e-12, (name, m, ref, se)
    assert np.all(b[:, 3 + nocc:7] == 0.0)
    eng.clear_samples()
    eng.run_passes("semigrand_canonical", 1, 1)
    _, cnt = eng.read_samples()
    for c in range(nch):
        nb = np.bincount(eng.read_state(c)["occ"], minlength=4)
        assert np.array_equal(cnt[c, -1, :4], nb[:4])
    eng.close()


@pytest.mark.parametrize("system", ["fcc_binary", "hcp_ternary"])
def test_nfold_specialized_matches_generic(system):
    _print_case(f"nfold specialised vs generic {system}")
    nch = 16
    T = np.linspace(300.0, 900.0, nch)
    if system == "fcc_binary":
        flat = FlatCE(fcc_binary_ce())
        sc = Supercell(flat.ce.prim, (12, 12, 12))
        occs = [random_occupation(sc.n_sites, 0.5, 900 + c, exact=False) for c in range(nch)]
        chem = [[0.0, m] for m in np.linspace(-0.1, 0.1, nch)]
    else:
        flat = _hcp_ternary_small()
        sc = Supercell(flat.ce.prim, (4, 4, 3))
        rng = np.random.default_rng(17)
        occs = [rng.integers(0, 3, sc.n_sites).astype(np.int8) for _ in range(nch)]
        chem = [[0.0, 0.02, -0.03] * flat.n_sublat] * nch
    out = {}
    for spec in (1, 0):
        eng = Engine(flat)
        eng.set_specialization(spec)
        sid = eng.add_supercell(sc, n_pair_phases=0)
        eng.load_chains([sid] * nch, occs, T, 4242 + np.arange(nch), chem=chem)
        blocks = eng.nfold_run(6000, 500)
        assert eng.last_run_info()["specialized"] == bool(spec)
        out[spec] = (blocks, [eng.read_state(c) for c in range(nch)])
        eng.close()
    assert np.array_equal(out[1][0], out[0][0])
    for a, b in zip(out[1][1], out[0][1]):
        assert np.array_equal(a["occ"], b["occ"]) and a["E"] == b["E"]


@pytest.mark.parametrize("system", ["fcc_binary", "hcp_ternary"])
def test_nfold_block_size_bit_identical(system):
    """The N-fold kernel's block size (clexmc_set_nfold_block_size) changes
    only how many threads share the image load, the initial rates and the
    neighbour rates of an event: blocks, occupations and energies are
    bit-identical for 64, 128 and 256 threads and the automatic choice."""
    _print_case(f"nfold block sizes {system}")
    nch = 16
    T = np.linspace(300.0, 900.0, nch)
    if system == "fcc_binary":
        flat = FlatCE(fcc_binary_ce())
        sc = Supercell(flat.ce.prim, (12, 12, 12))
        occs = [random_occupation(sc.n_sites, 0.5, 700 + c, exact=False) for c in range(nch)]
        chem = [[0.0, m] for m in np.linspace(-0.1, 0.1, nch)]
    else:
        flat = _hcp_ternary_small()
        sc = Supercell(flat.ce.prim, (4, 4, 3))
        rng = np.random.default_rng(23)
        occs = [rng.integers(0, 3, sc.n_sites).astype(np.int8) for _ in range(nch)]
        chem = [[0.0, 0.02, -0.03] * flat.n_sublat] * nch
    out = {}
    for thr in (0, 64, 128, 256):
        eng = Engine(flat)
        eng.set_nfold_block_size(thr)
        sid = eng.add_supercell(sc, n_pair_phases=0)
        eng.load_chains([sid] * nch, occs, T, 5151 + np.arange(nch), chem=chem)
        blocks = eng.nfold_run(6000, 500)
        out[thr] = (blocks, [eng.read_state(c) for c in range(nch)])
        eng.close()
    for thr in (64, 128, 256):
        assert np.array_equal(out[thr][0], out[0][0]), thr
        for a, b in zip(out[thr][1], out[0][1]):
            assert np.array_equal(a["occ"], b["occ"]) and a["E"] == b["E"], thr
    eng = Engine(flat)
    with pytest.raises(RuntimeError):
        eng.set_nfold_block_size(96)
    eng.close()


@pytest.mark.parametrize("system", ["fcc_binary", "hcp_ternary"])
def test_nfold_time_sampling_exact_enumeration(system):
    """Sampling by N-fold time (clexmc_nfold_run_time): the states occupied
    at times 0, dt, 2 dt, ... are Boltzmann-distributed with equal weights,
    so plain sample means of E, E^2 and every occupant count (and its
    square) over 128 chains lie within 3 sigma of exact enumeration (FCC
    binary 2x2x4, T 1200 K; ternary HCP 2x2x1, T 1000 K).  The chain stops
    at its last sample: that sample's occupant counts are the final state's
    and its energy the final energy (1e-9: the binary path sums the exact
    counts in another order at the end)."""
    _print_case(f"nfold time sampling exact enumeration {system}")
    if system == "fcc_binary":
        flat = small_flat()
        sc = Supercell(flat.ce.prim, (2, 2, 4))
        T, chem, nocc = 1200.0, [0.0, 0.03], 2
    else:
        flat = _hcp_ternary_small()
        sc = Supercell(flat.ce.prim, (2, 2, 1))
        T, nocc = 1000.0, 3
        chem = [0.0, 0.02, -0.03] * flat.n_sublat
    osys = O.OracleSystem(flat, sc.dims)
    ex = osys.enumerate("sgc", T, chem=chem, want_probs=True)
    pr, Es = ex["probs"], osys.state_energies()
    st_i = np.arange(len(pr))
    occ_s = (st_i[:, None] // (nocc ** np.arange(sc.n_sites))[None, :]) % nocc
    Ns = np.stack([(occ_s == o).sum(1) for o in range(nocc)], axis=1).astype(float)
    nch = 128
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    rng = np.random.default_rng(21)
    occs = [rng.integers(0, nocc, sc.n_sites).astype(np.int8) for _ in range(nch)]
    eng.load_chains([sid] * nch, occs, [T] * nch, 4000 + np.arange(nch), chem=[chem] * nch)
    eng.nfold_run(2000, 500)  # equilibration
    eng.clear_samples()
    n_samples, dt = 3000, 0.25
    eng.nfold_run_time(n_samples, dt, int(sc.n_sites * (n_samples * dt + 1)) + 1)
    E, cnt = eng.read_samples()
    assert E.shape == (nch, n_samples)
    checks = [("E", E, pr @ Es), ("E2", E * E, pr @ (Es * Es))]
    for o in range(nocc):
        n_o = cnt[:, :, o].astype(float)
        checks.append((f"N{o}", n_o, pr @ Ns[:, o]))
        checks.append((f"N{o}^2", n_o * n_o, pr @ (Ns[:, o] ** 2)))
    for name, x, ref in checks:
        per = x.mean(axis=1)
        m, se = per.mean(), per.std(ddof=1) / np.sqrt(nch)
        print(f"{name}: time-sampled {m:.6f} +- {se:.6f} exact {ref:.6f}")
        assert abs(m - ref) <= NSIG * se + 1e-12, (name, m, ref, se)
    for c in range(0, nch, 8):
        st = eng.read_state(c)
        assert abs(E[c, -1] - st["E"]) <= 1e-9 * max(1.0, abs(st["E"])), (E[c, -1], st["E"])
        nb = np.bincount(st["occ"], minlength=4)
        assert np.array_equal(cnt[c, -1, :4], nb[:4])
    eng.close()

