"""HIP-graph replays and the autotuner.

* A captured block of passes replayed k times leaves every chain bit-for-bit
  where k eager clexmc_run_passes calls of the same block leave it
  (occupations, running energies, counters, samples), checked after every
  replay with eager reads and synchronisation in between.
* The autotuner restores the chains' state after its trials, picks a valid
  configuration, answers a repeated query from its cache, and a run with the
  tuned configuration equals a run with that configuration set by hand.
"""
import numpy as np
import pytest

from casmcode_clexmonte_amd.engine import Engine
from casmcode_clexmonte_amd.system import FlatCE, Supercell
from casmcode_clexmonte_amd.workloads import (composition_counts, fcc_binary_ce, hcp_ternary_ce, random_occupation,
                                              random_occupation_counts)

pytestmark = pytest.mark.gpu


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _setup(which):
    if which == "binary_sgc":
        ce = fcc_binary_ce()
        flat = FlatCE(ce)
        sc = Supercell(ce.prim, (12, 12, 12))
        occs = [random_occupation(sc.n_sites, 0.5, 10 + c, exact=False) for c in range(8)]
        chem = [[0.0, mu] for mu in np.linspace(-0.1, 0.1, 8)]
        return flat, sc, occs, np.linspace(300, 1500, 8), chem, "semigrand_canonical", 0
    ce = hcp_ternary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (8, 8, 6))
    occs = [random_occupation_counts(composition_counts(sc.n_sites, (0.4, 0.3, 0.3)), 20 + c) for c in range(4)]
    return flat, sc, occs, np.linspace(600, 1800, 4), None, "canonical", 4


def _engine(flat, sc, occs, temps, chem, n_phases):
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=n_phases, seed=2)
    eng.load_chains([sid] * len(occs), occs, temps, 500 + np.arange(len(occs)), chem=chem)
    return eng


def _states(eng, n):
    return [eng.read_state(c) for c in range(n)]


def _same(sa, sb):
    for a, b in zip(sa, sb):
        assert np.array_equal(a["occ"], b["occ"])
        assert a["E"] == b["E"]
        assert a["n_attempt"] == b["n_attempt"] and a["n_accept"] == b["n_accept"]


@pytest.mark.parametrize("which", ["binary_sgc", "ternary_canonical"])
def test_graph_replays_equal_eager_blocks(which):
    _print_case(f"graph replays vs eager {which}")
    flat, sc, occs, temps, chem, kind, n_phases = _setup(which)
    eager = _engine(flat, sc, occs, temps, chem, n_phases)
    graph = _engine(flat, sc, occs, temps, chem, n_phases)
    block, reps = 10, 4
    graph.graph_capture(kind, block, reps, sample_period=1, canonical_rounds=2)
    for r in range(reps):
        eager.run_passes(kind, block, 1, canonical_rounds=2)
        graph.graph_replay(1)
        _same(_states(eager, len(occs)), _states(graph, len(occs)))
        Ee, ce = eager.read_samples()
        Eg, cg = graph.read_samples()
        assert Ee.shape == Eg.shape == (len(occs), block * (r + 1))
        assert np.array_equal(Ee, Eg) and np.array_equal(ce, cg)
    with pytest.raises(Exception):
        graph.graph_replay(1)        # beyond the captured sample capacity
    graph.graph_release()
    eager.close()
    graph.close()


@pytest.mark.parametrize("which", ["binary_sgc", "ternary_canonical"])
def test_autotune_restores_state_and_matches_manual_setting(which):
    _print_case(f"autotune {which}")
    flat, sc, occs, temps, chem, kind, n_phases = _setup(which)
    tuned = _engine(flat, sc, occs, temps, chem, n_phases)
    before = _states(tuned, len(occs))
    cfg = tuned.autotune(kind, trial_passes=5)
    print(cfg)
    assert cfg["threads"] in (64, 128, 256, 512) and cfg["packing"] in ("int8", "packed2")
    assert cfg["ms_per_pass"] > 0
    _same(before, _states(tuned, len(occs)))
    assert tuned.autotune(kind, trial_passes=5)["ms_per_pass"] == -1.0    # cached
    manual = _engine(flat, sc, occs, temps, chem, n_phases)
    manual.set_block_size(cfg["threads"])
    manual.set_image_packing(cfg["packing"])
    manual.set_lanes_per_site(cfg["lanes_per_site"])
    tuned.run_passes(kind, 15, 1, canonical_rounds=2)
    manual.run_passes(kind, 15, 1, canonical_rounds=2)
    _same(_states(tuned, len(occs)), _states(manual, len(occs)))
    tuned.close()
    manual.close()


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("kind", ["semigrand_canonical", "canonical"])
@pytest.mark.parametrize("packing", ["int8", "packed2"])
def test_specialized_kernel_bit_identical_to_generic(kind, packing, lanes):
    """The run-time specialised binary kernel (constant site offsets) leaves
    every chain bit-for-bit where the generic kernel (interpreted site
    programs) does: same occupations, exact counts, energies, samples."""
    _print_case(f"specialised vs generic {kind} {packing} lanes/site {lanes}")
    ce = fcc_binary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (12, 12, 12))
    occs = [random_occupation(sc.n_sites, 0.5, 30 + c, exact=(kind == "canonical")) for c in range(6)]
    chem = [[0.0, mu] for mu in np.linspace(-0.1, 0.1, 6)]
    temps = np.linspace(300, 1500, 6)
    runs = []
    for spec in (False, True):
        eng = _engine(flat, sc, occs, temps, chem, 8)
        eng.set_specialization(spec)
        eng.set_image_packing(packing)
        eng.set_lanes_per_site(lanes)
        eng.run_passes(kind, 25, 1, canonical_rounds=2)
        info = eng.last_run_info()
        assert info["specialized"] == spec and info["lanes_per_site"] == lanes, \
            (info, eng.specialization_log()[-3000:])
        E, cnt = eng.read_samples()
        runs.append((_states(eng, len(occs)), [eng.read_binary_counts(c) for c in range(len(occs))], E, cnt))
        eng.close()
    (s0, c0, E0, n0), (s1, c1, E1, n1) = runs
    _same(s0, s1)
    for a, b in zip(c0, c1):
        assert np.array_equal(a, b)
    assert np.array_equal(E0, E1) and np.array_equal(n0, n1)


@pytest.mark.parametrize("kind", ["semigrand_canonical", "canonical"])
@pytest.mark.parametrize("packing", ["int8", "packed2"])
def test_specialized_general_kernel_bit_identical_to_generic(kind, packing):
    """Same for the table path: ternary HCP (two sublattices)."""
    _print_case(f"specialised vs generic general {kind} {packing}")
    ce = hcp_ternary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (8, 8, 6))
    rng = np.random.default_rng(4)
    if kind == "canonical":
        occs = [random_occupation_counts(composition_counts(sc.n_sites, (0.4, 0.3, 0.3)), 60 + c) for c in range(4)]
    else:
        occs = [rng.integers(0, 3, sc.n_sites).astype(np.int8) for _ in range(4)]
    chem = [[0.0, 0.02 * c, -0.01 * c] * 2 for c in range(4)]
    temps = np.linspace(600, 1800, 4)
    runs = []
    for spec in (False, True):
        eng = _engine(flat, sc, occs, temps, chem, 4)
        eng.set_specialization(spec)
        eng.set_image_packing(packing)
        eng.run_passes(kind, 8, 1, canonical_rounds=2)
        info = eng.last_run_info()
        assert info["specialized"] == spec, (info, eng.specialization_log()[-3000:])
        E, cnt = eng.read_samples()
        runs.append((_states(eng, len(occs)), E, cnt))
        eng.close()
    (s0, E0, n0), (s1, E1, n1) = runs
    _same(s0, s1)
    assert np.array_equal(E0, E1) and np.array_equal(n0, n1)


def test_grouped_shape_launches_bit_identical_to_one_launch():
    """Chains on uneven supercells: the fused specialised launch (one kernel
    holding every shape's field) and the per-shape specialised launches on
    concurrent streams both equal the single generic launch over all chains
    bit for bit (every shape's blocks have 64 threads here)."""
    _print_case("grouped specialised launches vs one generic launch")
    ce = hcp_ternary_ce()
    flat = FlatCE(ce)
    shapes = [(6, 6, 4), (8, 8, 5), (10, 10, 6)]
    occs, dims = [], []
    for c in range(6):
        d = shapes[c % 3]
        dims.append(d)
        occs.append(random_occupation_counts(composition_counts(2 * d[0] * d[1] * d[2], (0.4, 0.3, 0.3)), 80 + c))
    temps = np.linspace(700, 1700, 6)
    runs = []
    for spec, fuse in ((False, 1), (True, 1), (True, 0)):
        eng = Engine(flat)
        eng.set_specialization(spec)
        eng.set_fused_launch(fuse)
        sids = [eng.add_supercell(Supercell(ce.prim, d), n_pair_phases=4, seed=i) for i, d in enumerate(shapes)]
        eng.load_chains([sids[c % 3] for c in range(6)], occs, temps, 300 + np.arange(6))
        eng.run_passes("canonical", 10, 1, canonical_rounds=2)
        assert eng.last_run_info()["specialized"] == spec, eng.specialization_log()[-2000:]
        assert eng.last_run_fused() == (spec and fuse == 1)
        E, cnt = eng.read_samples()
        runs.append((_states(eng, 6), E, cnt))
        eng.close()
    (s0, E0, n0) = runs[0]
    for (s1, E1, n1) in runs[1:]:
        _same(s0, s1)
        assert np.array_equal(E0, E1) and np.array_equal(n0, n1)


@pytest.mark.parametrize("which", ["binary_sgc", "ternary_canonical"])
@pytest.mark.parametrize("specialize", [1, 0])
def test_lds_row_pitch_does_not_change_results(which, specialize):
    """The padded image's row pitch (automatic bank-conflict choice, a fixed
    odd padding, none) changes only the LDS layout: every chain ends in the
    same state with the same samples."""
    _print_case(f"row pitch identity {which} specialize {specialize}")
    flat, sc, occs, temps, chem, kind, pp = _setup(which)
    out = []
    for pitch in (0, -1, 3):
        eng = Engine(flat)
        eng.set_specialization(specialize)
        eng.set_lds_row_pitch(pitch)
        sid = eng.add_supercell(sc, n_pair_phases=pp, seed=3)
        eng.load_chains([sid] * len(occs), occs, temps, 40 + np.arange(len(occs)), chem=chem)
        eng.run_passes(kind, 12, 1, canonical_rounds=2)
        E, cnt = eng.read_samples()
        states = [eng.read_state(c) for c in range(len(occs))]
        out.append((E, cnt, [s["occ"] for s in states], [s["E"] for s in states]))
        eng.close()
    for E, cnt, occ, e in out[1:]:
        assert np.array_equal(E, out[0][0]) and np.array_equal(cnt, out[0][1])
        assert all(np.array_equal(a, b) for a, b in zip(occ, out[0][2]))
        assert e == out[0][3]


@pytest.mark.parametrize("packing", ["int8", "packed2"])
def test_wide_row_pitch_small_supercell(packing):
    """A row pitch far wider than the rows (64 unused bytes after each
    10-byte row of a 6x6x6 image): the unused bytes are filled from a clamped
    valid cell (the image load's global reads stay inside the chain), and the
    Metropolis and N-fold runs end exactly as with the natural pitch."""
    _print_case(f"wide row pitch 6^3 {packing}")
    ce = fcc_binary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (6, 6, 6))
    n = 8
    occs = [random_occupation(sc.n_sites, 0.5, 60 + c, exact=False) for c in range(n)]
    chem = [[0.0, mu] for mu in np.linspace(-0.1, 0.1, n)]
    temps = np.linspace(300, 1500, n)
    out = []
    for pitch in (0, 64):
        eng = Engine(flat)
        eng.set_lds_row_pitch(pitch)
        eng.set_image_packing(packing)
        sid = eng.add_supercell(sc, n_pair_phases=0, seed=3)
        eng.load_chains([sid] * n, occs, temps, 70 + np.arange(n), chem=chem)
        eng.run_passes("semigrand_canonical", 12, 1)
        E, cnt = eng.read_samples()
        blocks = eng.nfold_run(2000, 500)
        states = [eng.read_state(c) for c in range(n)]
        out.append((E, cnt, blocks, [s["occ"] for s in states], [s["E"] for s in states]))
        eng.close()
    E, cnt, blocks, occ, e = out[1]
    assert np.array_equal(E, out[0][0]) and np.array_equal(cnt, out[0][1])
    assert np.array_equal(blocks, out[0][2])
    assert all(np.array_equal(a, b) for a, b in zip(occ, out[0][3]))
    assert e == out[0][4]


@pytest.mark.parametrize("dims,n_chains", [((24, 24, 24), 256), ((12, 12, 12), 8), ((6, 24, 24), 4)])
def test_register_cube_bit_identical_to_class_loop(dims, n_chains):
    """The register-cube sweep of the specialised semi-grand canonical
    binary kernel (each lane keeps the cells around its 3x3x3 block in
    registers and reads back only the class just updated) against the class
    loop: the same states, counts, energies and samples bit for bit after
    30 passes, on the headline workload (FCC 24^3, 256 chains on the
    T x mu grid) and smaller supercells (12^3; 6x24x24, two blocks along the
    first axis, so a block's neighbours on both sides are the same block)."""
    _print_case(f"register cube vs class loop fcc {dims} x {n_chains}")
    from casmcode_clexmonte_amd.workloads import fcc_sgc_grid
    ce = fcc_binary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, dims)
    batch = fcc_sgc_grid(sc) if n_chains == 256 else None
    runs = []
    for cube in (1, 0):
        eng = Engine(flat, device=0)
        eng.set_register_cube(cube)
        sid = eng.add_supercell(sc, n_pair_phases=0)
        if batch is not None:
            eng.load_chains([sid] * n_chains, batch.occ, batch.temperatures, batch.seeds, chem=batch.chem)
        else:
            occs = [random_occupation(sc.n_sites, 0.5, 70 + c, exact=False) for c in range(n_chains)]
            eng.load_chains([sid] * n_chains, occs, np.linspace(300.0, 1500.0, n_chains), 90 + np.arange(n_chains),
                            chem=[[0.0, m] for m in np.linspace(-0.2, 0.2, n_chains)])
        eng.run_passes("semigrand_canonical", 30, 1)
        assert eng.last_run_cube() == bool(cube), eng.specialization_log()[-2000:]
        E, cnt = eng.read_samples()
        states = [eng.read_state(c) for c in range(n_chains)]
        counts = [eng.read_binary_counts(c) for c in range(n_chains)]
        runs.append((states, counts, E, cnt))
        eng.close()
    (s1, c1, E1, n1), (s0, c0, E0, n0) = runs
    for a, b in zip(s1, s0):
        assert np.array_equal(a["occ"], b["occ"]) and a["E"] == b["E"] and a["n_accept"] == b["n_accept"]
    for a, b in zip(c1, c0):
        assert np.array_equal(a, b)
    assert np.array_equal(E1, E0) and np.array_equal(n1, n0)
