"""Rare events: Metropolis acceptance and N-fold rates far below 2^-24.

All cases use expansions whose ECIs are zero, so each site is independent and
the semi-grand canonical answers are closed forms: with site-occupant
potentials chem_o, a site holds occupant o with probability
exp(beta chem_o) / sum_q exp(beta chem_q).

* Metropolis acceptance (FCC binary 24^3, 256 chains, 1,000 passes, all-A
  start, chem = [0, -D]): an A site flips to B with probability
  p = exp(-beta D) per attempt, so the number of A -> B acceptances is
  Poisson with mean (attempts on A sites) x p.  At beta D = 18
  (p = 1.5e-8, below the 2^-24 = 6e-8 resolution of a 24-bit uniform) the
  count must lie within 5 sigma of that mean; at beta D = 24 (p = 3.8e-11)
  the expected count is 0.13, so at most 4 are allowed (a 24-bit uniform
  would give about 210 in both cases).
* N-fold way at rates ~1e-9 (beta D = 20): FCC binary 12^3 (27 rate groups)
  and ternary HCP 4x4x3 (general path, two excited occupants), all-A
  start, 64 chains: the residence-weighted mean of N_1 per site equals the
  closed form within 3 sigma over chains.  Here a chain sits in the all-A
  state (total rate ~N p) between excursions whose own rate is 1, so group
  sums updated by differences in FP32 would cancel to 0 after the first
  excursion.
* A chain whose every rate underflows (beta D = 110) stops and the run
  reports an error; the handle stays usable.
"""
import numpy as np
import pytest

from casmcode_clexmonte_amd.engine import ClexmcError, Engine
from casmcode_clexmonte_amd.system import BasisSet, ClusterExpansion, FlatCE, Supercell, fcc_prim, hcp_prim
from casmcode_clexmonte_amd.workloads import fcc_binary_basis

pytestmark = pytest.mark.gpu

KB = 8.6173303e-05


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _zero_fcc_binary():
    bs = fcc_binary_basis()
    return FlatCE(ClusterExpansion(bs, np.zeros(bs.n_functions)))


def _zero_hcp_ternary():
    prim = hcp_prim(1.0, occupants=("A", "B", "C"))
    bs = BasisSet(prim, {2: 1.01}, "chebychev")
    return FlatCE(ClusterExpansion(bs, np.zeros(bs.n_functions)))


@pytest.mark.parametrize("beta_d", [18.0, 24.0])
def test_metropolis_acceptance_below_24_bits(beta_d):
    _print_case(f"metropolis rare acceptance beta*D={beta_d}")
    flat = _zero_fcc_binary()
    sc = Supercell(flat.ce.prim, (24, 24, 24))
    D = 0.3
    T = D / (KB * beta_d)
    nch, n_passes = 256, 1000
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    occ0 = np.zeros(sc.n_sites, dtype=np.int8)
    eng.load_chains([sid] * nch, [occ0] * nch, [T] * nch, 5150 + np.arange(nch), chem=[[0.0, -D]] * nch)
    eng.run_passes("semigrand_canonical", n_passes, n_passes)
    n_acc = n_b = n_att = 0
    for c in range(nch):
        st = eng.read_state(c)
        n_acc += st["n_accept"]
        n_att += st["n_attempt"]
        n_b += int(st["occ"].sum())
    eng.close()
    # every accepted B -> A change undoes an earlier A -> B change
    n_ab = (n_acc + n_b) // 2
    p = np.exp(-beta_d)
    lam = (n_att - n_ab) * p  # attempts made on A sites (all but the few on B sites)
    print(f"attempts {n_att}  A->B accepted {n_ab}  expected {lam:.2f}  (24-bit floor: {n_att * 2.0**-24:.0f})")
    assert n_att == nch * sc.n_sites * n_passes
    if lam > 20:
        assert abs(n_ab - lam) <= 5.0 * np.sqrt(lam), (n_ab, lam)
    else:
        assert n_ab <= 4, (n_ab, lam)


@pytest.mark.parametrize("system", ["fcc_binary_12", "hcp_ternary_4x4x3"])
def test_nfold_rates_near_1e9(system):
    _print_case(f"nfold rare rates {system}")
    if system == "fcc_binary_12":
        flat, dims = _zero_fcc_binary(), (12, 12, 12)
    else:
        flat, dims = _zero_hcp_ternary(), (4, 4, 3)
    sc = Supercell(flat.ce.prim, dims)
    nocc = int(flat.n_occ_max)
    D = 0.2
    T = D / (KB * 20.0)
    beta = 1.0 / (KB * T)
    chem_b = [0.0, -D] if nocc == 2 else [0.0, -D, -D - 0.2 * KB * T * 10]
    chem = list(chem_b) * flat.n_sublat
    w = np.exp(beta * np.array(chem_b))
    x1 = w[1] / w.sum()  # closed form: sites holding occupant 1, per site
    nch = 64
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    occ0 = np.zeros(sc.n_sites, dtype=np.int8)
    eng.load_chains([sid] * nch, [occ0] * nch, [T] * nch, 6100 + np.arange(nch), chem=[chem] * nch)
    blocks = eng.nfold_run(20000, 1000)
    eng.close()
    assert np.all(np.isfinite(blocks)) and np.all(blocks[:, :, 0] > 0)
    b = blocks.sum(axis=1)
    est = b[:, 2] / b[:, 0] / sc.n_sites
    m, se = est.mean(), est.std(ddof=1) / np.sqrt(nch)
    print(f"x_1 per site: nfold {m:.6e} +- {se:.2e}  closed form {x1:.6e}")
    assert abs(m - x1) <= 3.0 * se + 1e-3 * x1, (m, x1, se)


def test_nfold_underflowing_rates_report_an_error():
    _print_case("nfold stalled chain error")
    flat = _zero_fcc_binary()
    sc = Supercell(flat.ce.prim, (6, 6, 6))
    D = 0.3
    T = D / (KB * 110.0)
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    occ0 = np.zeros(sc.n_sites, dtype=np.int8)
    eng.load_chains([sid] * 2, [occ0] * 2, [T, 1500.0], [1, 2], chem=[[0.0, -D], [0.0, -D]])
    with pytest.raises(ClexmcError, match="stopped"):
        eng.nfold_run(100, 10)
    # the flag is cleared: a run at a temperature with live rates succeeds
    eng.load_chains([sid] * 2, [occ0] * 2, [1500.0, 1500.0], [1, 2], chem=[[0.0, -D], [0.0, -D]])
    blocks = eng.nfold_run(100, 10)
    assert np.all(blocks[:, :, 0] > 0)
    eng.close()
