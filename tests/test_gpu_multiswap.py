"""Semi-grand canonical multiswaps (SemiGrandCanonicalCalculator.cc:104-115;
the system's semigrand_canonical_multiswaps, System_json_io.cc:955-1006).

A charge-neutral toy: ternary FCC (occupants A, B, C read as charges 0, +1,
-1), 2x2x2 = 8 sites, random ECIs (general table path), and the multiswap
types {A->B, A->C}, its reverse {B->A, C->A}, and {B->C, C->B} (its own
reverse), so composition moves only along n_B = n_C from the all-A start.
The stationary law is then the semi-grand canonical one restricted to
n_B = n_C: P(x) ~ exp(-beta (E(x) - chem . n(x))) over those 1,107 states,
computed exactly from the oracle's energies of all 3^8 states.

* 128 chains, 3,000 passes (the fixed-composition sweep plus 8 multiswap
  attempts per pass), both with the int8 and the packed image: <E>, <n_B>,
  <n_B^2> and P(n_B = k), k = 0..4, within 3 sigma (standard error over
  chains) of the exact values; n_B = n_C in every sample; the final running
  energies equal the oracle's within the FP32 accumulation bound.
* The proposal ratio matters here: from n_B = n_C = k with a A sites,
  {A->B, A->C} has q_rev / q_fwd = a (a - 1) / (k + 1)^2 (56 at k = 0), so
  accepting with exp(-beta dOmega) alone would put P(n_B = k) far from the
  exact values.
* The multiswap list is validated on the host: a list without the reverse of
  a type is refused.
"""
import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.engine import ClexmcError, Engine
from casmcode_clexmonte_amd.system import BasisSet, ClusterExpansion, FlatCE, Supercell, fcc_prim

pytestmark = pytest.mark.gpu

KB = 8.6173303e-05
NSIG = 3.0
TYPES = [[(0, 0, 1, 1), (0, 0, 2, 1)], [(0, 1, 0, 1), (0, 2, 0, 1)], [(0, 1, 2, 1), (0, 2, 1, 1)]]


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _ternary_flat(seed=4):
    prim = fcc_prim(1.0, ("A", "B", "C"))
    bs = BasisSet(prim, {2: 1.05, 3: 0.75}, "chebychev")
    return FlatCE(ClusterExpansion(bs, np.random.default_rng(seed).normal(0, 0.04, bs.n_functions)))


def _exact(flat, dims, T, chem):
    osys = O.OracleSystem(flat, dims)
    E = osys.state_energies()
    n = int(np.prod(dims))
    st = np.arange(3 ** n)
    occ = (st[:, None] // (3 ** np.arange(n))[None, :]) % 3
    nB = (occ == 1).sum(1)
    nC = (occ == 2).sum(1)
    keep = nB == nC
    omega = E - chem[1] * nB - chem[2] * nC
    w = np.where(keep, np.exp(-(omega - omega[keep].min()) / (KB * T)), 0.0)
    w /= w.sum()
    return {"E": float(w @ E), "nB": float(w @ nB), "nB2": float(w @ nB ** 2),
            "P": [float(w[nB == k].sum()) for k in range(5)]}


@pytest.mark.parametrize("packing", ["int8", "packed2"])
def test_multiswap_exact_enumeration(packing):
    _print_case(f"multiswap exact enumeration {packing}")
    flat = _ternary_flat()
    dims = (2, 2, 2)
    sc = Supercell(flat.ce.prim, dims)
    T = 1200.0
    chem = [0.0, 0.05, -0.02]
    ex = _exact(flat, dims, T, chem)
    nch = 128
    eng = Engine(flat)
    eng.set_image_packing(packing)
    sid = eng.add_supercell(sc, n_pair_phases=8, seed=3)
    eng.set_multiswaps(TYPES, per_pass=8)
    occ0 = np.zeros(sc.n_sites, dtype=np.int8)
    eng.load_chains([sid] * nch, [occ0] * nch, [T] * nch, 8100 + np.arange(nch), chem=[chem] * nch)
    eng.run_passes("semigrand_canonical", 3000, 1, canonical_rounds=0)
    E, cnt = eng.read_samples()
    assert np.all(cnt[:, :, 1] == cnt[:, :, 2])
    burn = 200
    e = E[:, burn:]
    nB = cnt[:, burn:, 1].astype(float)
    per = {"E": e.mean(1), "nB": nB.mean(1), "nB2": (nB * nB).mean(1)}
    for k in range(5):
        per[f"P{k}"] = (nB == k).mean(1)
        ex[f"P{k}"] = ex["P"][k]
    for key in ("E", "nB", "nB2", "P0", "P1", "P2", "P3", "P4"):
        m, se = per[key].mean(), per[key].std(ddof=1) / np.sqrt(nch)
        print(f"{key}: gpu {m:.6f} +- {se:.6f} exact {ex[key]:.6f}")
        assert abs(m - ex[key]) <= NSIG * se + 1e-12, (key, m, ex[key], se)
    osys = O.OracleSystem(flat, dims)
    for c in range(0, nch, 16):
        st = eng.read_state(c)
        Eref = osys.energy(st["occ"])
        assert abs(st["E"] - Eref) <= 4e-6 * np.sqrt(st["n_accept"] + 1) + 1e-9, (st["E"], Eref)
        assert st["n_accept"] > 0
    eng.close()


def test_multiswap_list_needs_reverses():
    _print_case("multiswap list validation")
    flat = _ternary_flat()
    eng = Engine(flat)
    with pytest.raises(ClexmcError, match="undoes"):
        eng.set_multiswaps(TYPES[:1])
    with pytest.raises(ClexmcError, match="occupants"):
        eng.set_multiswaps([[(0, 0, 0, 1)]])
    eng.set_multiswaps(TYPES)
    eng.set_multiswaps([])
    eng.close()
