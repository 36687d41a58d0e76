"""GPU parity tests: the HIP path through the C ABI against the FP64 oracle.

Tolerances (stated per result):
* per-proposal dE (FP32 on the GPU): |dE_gpu - dE_oracle| <= 1e-5 * s + 1e-6 eV,
  s = the largest |dE| of the batch (relative to the batch scale, so exact
  cancellations do not make a relative test meaningless);
* full correlations (FP64 on the GPU): 1e-12 relative; binary +-1 sums exact;
* exact integer correlation counts of the binary Metropolis path: bit-exact;
* occupation counts: bit-exact.
"""
import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.engine import Engine
from casmcode_clexmonte_amd.system import BasisSet, ClusterExpansion, FlatCE, Supercell, fcc_prim
from casmcode_clexmonte_amd.workloads import fcc_binary_ce, random_occupation

pytestmark = pytest.mark.gpu

TOL_REL = 1e-5
TOL_ABS = 1e-6


def _print_case(name):
    print(f"[case] {name}", flush=True)


@pytest.fixture(scope="module")
def fcc_bin():
    ce = fcc_binary_ce()
    return FlatCE(ce)


def ternary_fcc_flat(seed=3):
    prim = fcc_prim(1.0, ("A", "B", "C"))
    bs = BasisSet(prim, {2: 1.05, 3: 0.75}, "chebychev")
    rng = np.random.default_rng(seed)
    return FlatCE(ClusterExpansion(bs, rng.normal(0, 0.05, bs.n_functions)))


def test_philox_device_kat(fcc_bin):
    _print_case("philox kat")
    eng = Engine(fcc_bin)
    ctr = np.array([[0, 0, 0, 0], [0xffffffff] * 4, [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344]],
                   dtype=np.uint32)
    key = np.array([[0, 0], [0xffffffff, 0xffffffff], [0xa4093822, 0x299f31d0]], dtype=np.uint32)
    out = eng.philox(ctr, key)
    # Random123 known-answer vectors for Philox4x32-10
    exp = np.array([[0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8],
                    [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd],
                    [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]], dtype=np.uint32)
    assert np.array_equal(out, exp)
    eng.close()


@pytest.mark.parametrize("which", ["binary", "ternary"])
def test_delta_energy_batch_vs_oracle(fcc_bin, which):
    _print_case(f"delta batch {which}")
    flat = fcc_bin if which == "binary" else ternary_fcc_flat()
    nocc = int(flat.n_occ[0])
    sc = Supercell(flat.ce.prim, (8, 8, 8))
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    rng = np.random.default_rng(11)
    occs = [rng.integers(0, nocc, sc.n_sites).astype(np.int8) for _ in range(3)]
    eng.load_chains([sid] * 3, occs, [600.0] * 3, [1, 2, 3])
    osys = O.OracleSystem(flat, sc.dims)
    # single-site changes
    n = 300
    chain = rng.integers(0, 3, n)
    sites = rng.integers(0, sc.n_sites, n)
    new = np.array([(occs[c][s] + 1 + rng.integers(0, nocc - 1)) % nocc for c, s in zip(chain, sites)])
    dE = eng.delta_energy_batch(chain, sites, new)
    ref = np.array([osys.delta(occs[c], [s], [v]) for c, s, v in zip(chain, sites, new)])
    scale = np.abs(ref).max()
    assert np.all(np.abs(dE - ref) <= TOL_REL * scale + TOL_ABS), np.abs(dE - ref).max()
    # swaps (two sequential changes; includes neighbouring pairs)
    s1 = rng.integers(0, sc.n_sites, n)
    s2 = np.where(rng.random(n) < 0.5, rng.integers(0, sc.n_sites, n), (s1 + 1) % sc.n_sites)
    s2 = np.where(s2 == s1, (s1 + 7) % sc.n_sites, s2)
    sites2 = np.stack([s1, s2], axis=1)
    new2 = np.stack([[occs[c][b], occs[c][a]] for c, a, b in zip(chain, s1, s2)])
    dE2 = eng.delta_energy_batch(chain, sites2, new2)
    ref2 = np.array([osys.delta(occs[c], list(s), list(v)) for c, s, v in zip(chain, sites2, new2)])
    scale2 = np.abs(ref2).max()
    assert np.all(np.abs(dE2 - ref2) <= TOL_REL * scale2 + TOL_ABS), np.abs(dE2 - ref2).max()
    eng.close()


@pytest.mark.parametrize("which", ["binary", "ternary"])
def test_correlations_vs_oracle(fcc_bin, which):
    _print_case(f"correlations {which}")
    flat = fcc_bin if which == "binary" else ternary_fcc_flat()
    nocc = int(flat.n_occ[0])
    sc = Supercell(flat.ce.prim, (6, 7, 8))
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    rng = np.random.default_rng(5)
    occ = rng.integers(0, nocc, sc.n_sites).astype(np.int8)
    eng.load_chains([sid], [occ], [600.0], [9])
    corr = eng.correlations(0)
    ref = O.OracleSystem(flat, sc.dims).corr(occ)
    if which == "binary":
        assert np.array_equal(corr, ref)
    else:
        assert np.allclose(corr, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    eng.close()


@pytest.mark.parametrize("packing", ["int8", "packed2"])
@pytest.mark.parametrize("kind", ["semigrand_canonical", "canonical"])
def test_binary_run_exact_counts(fcc_bin, kind, packing):
    """After a Metropolis run, the exact int64 correlation counts tracked on
    the device equal the oracle's full recompute of the final occupation, and
    the occupant counts are exact (conserved for canonical); int8 and 2-bit
    packed LDS images."""
    _print_case(f"binary run exact counts {kind} {packing}")
    flat = fcc_bin
    sc = Supercell(flat.ce.prim, (12, 12, 12))
    eng = Engine(flat)
    eng.set_image_packing(packing)
    sid = eng.add_supercell(sc, n_pair_phases=8, seed=1)
    occs = [random_occupation(sc.n_sites, 0.5, s) for s in range(4)]
    chem = [[0.0, mu] for mu in (-0.05, 0.0, 0.05, 0.1)]
    eng.load_chains([sid] * 4, occs, [400.0, 800.0, 1200.0, 2000.0], [11, 12, 13, 14], chem=chem)
    eng.run_passes(kind, 30, 1, canonical_rounds=2)
    osys = O.OracleSystem(flat, sc.dims)
    E, cnt = eng.read_samples()
    assert E.shape == (4, 30)
    for c in range(4):
        st = eng.read_state(c)
        C = eng.read_binary_counts(c)
        corr = osys.corr(st["occ"])
        assert np.array_equal(C, np.rint(corr * flat.orbit_mult).astype(np.int64))
        Eref = osys.energy(st["occ"])
        assert abs(st["E"] - Eref) <= 1e-9 * max(1.0, abs(Eref))
        assert abs(E[c, -1] - Eref) <= 1e-9 * max(1.0, abs(Eref))
        nb = int(st["occ"].sum())
        assert cnt[c, -1, 1] == nb and cnt[c, -1, 0] == sc.n_sites - nb
        if kind == "canonical":
            assert nb == int(occs[c].sum())
        assert st["n_accept"] > 0 and st["n_attempt"] >= st["n_accept"]
    eng.close()


@pytest.mark.parametrize("packing", ["int8", "packed2"])
@pytest.mark.parametrize("kind", ["semigrand_canonical", "canonical"])
def test_general_path_run_energy_tracking(kind, packing):
    """Ternary (table path): FP64 running energy of FP32 dEs versus the
    oracle's energy of the final occupation; composition exact; int8 and
    2-bit packed LDS images."""
    _print_case(f"general run {kind} {packing}")
    flat = ternary_fcc_flat()
    sc = Supercell(flat.ce.prim, (9, 9, 9))
    eng = Engine(flat)
    eng.set_image_packing(packing)
    sid = eng.add_supercell(sc, n_pair_phases=4, seed=2)
    rng = np.random.default_rng(8)
    occs = [rng.integers(0, 3, sc.n_sites).astype(np.int8) for _ in range(2)]
    eng.load_chains([sid] * 2, occs, [700.0, 1500.0], [21, 22], chem=[[0, 0.01, -0.02], [0, 0.0, 0.0]])
    eng.run_passes(kind, 20, 1, canonical_rounds=2)
    osys = O.OracleSystem(flat, sc.dims)
    for c in range(2):
        st = eng.read_state(c)
        Eref = osys.energy(st["occ"])
        n_acc = max(1, st["n_accept"])
        # FP32 dE rounding, accumulated in FP64 over n_acc accepted moves
        assert abs(st["E"] - Eref) <= 1e-6 * np.sqrt(n_acc) + 1e-9 * abs(Eref), (st["E"], Eref)
        if kind == "canonical":
            assert np.array_equal(np.bincount(st["occ"], minlength=3), np.bincount(occs[c], minlength=3))
    eng.close()
