"""Detailed balance by chi-square: the visited (energy level, B count)
histogram of the GPU chains against the exact Boltzmann distribution over
the same bins.

System: FCC binary, pairs to the 2nd shell and nearest-neighbour triplets
with N(0, 0.05 eV) ECIs; semi-grand canonical on 2x2x4 = 16 sites (2^16
states, T = 1200 K, mu = 0.03 eV), canonical on 2x2x6 = 24 sites with 12 B
(C(24,12) states, T = 900 K; independent sets of 3 sites, so the grouped
proposals' weights enter the canonical ratio).  The exact distribution is
built here from the oracle's energy of every state (weights
exp(-beta (E - mu n_B)), or exp(-beta E) over the fixed-count states).

Cases: the plain kernel (semi-grand canonical; canonical with 2 rounds and
with automatic rounds) and the grouped proposals (learned with a random
network, uniform, informed; semi-grand canonical and canonical).  Samples
are taken every 10 passes after 200 (the energy's autocorrelation time is
about one pass at these temperatures); bins expected to hold fewer than 5 samples are
merged.  A case passes if the chi-square upper-tail probability is above
1e-4.  Power: the plain semi-grand canonical histogram tested against the
distribution at 1.2 T must give a probability below 1e-6 (exact independent
samples of the same size give about 1e-26 there, and 6e-5 at 1.1 T).
"""
import numpy as np
import pytest
from scipy.stats import chi2

import oracle as O
from casmcode_clexmonte_amd.engine import Engine
from casmcode_clexmonte_amd.system import KB, BasisSet, ClusterExpansion, FlatCE, Supercell, fcc_prim
from casmcode_clexmonte_amd.workloads import random_mlp_params, random_occupation

pytestmark = pytest.mark.gpu

SGC_DIMS, CAN_DIMS = (2, 2, 4), (2, 2, 6)
T_SGC, T_CAN = 1200.0, 900.0
MU = 0.03
NB_CANONICAL = 12
N_CHAINS = 128
N_PASSES = 4200
BURN, THIN = 200, 10
P_MIN = 1e-4


def _print_case(name):
    print(f"[case] {name}", flush=True)


def small_flat(seed=0):
    prim = fcc_prim()
    bs = BasisSet(prim, {2: 1.05, 3: 0.75}, "chebychev")
    eci = np.random.default_rng(seed).normal(0, 0.05, bs.n_functions)
    return FlatCE(ClusterExpansion(bs, eci))


_STATES = {}


def _states(kind):
    """Oracle energies and B counts of the states of the ensemble: all 2^16
    states of the 2x2x4 cell (semi-grand canonical) or the C(24, 12) states of
    the 2x2x6 cell with 12 B (canonical)."""
    if kind not in _STATES:
        flat = small_flat()
        if kind == "sgc":
            E = O.OracleSystem(flat, SGC_DIMS).state_energies()
            idx = np.arange(len(E))
            nb = ((idx[:, None] >> np.arange(16)[None, :]) & 1).sum(axis=1)
        else:
            n = 24
            E = O.OracleSystem(flat, CAN_DIMS).state_energies(fixed_counts=[n - NB_CANONICAL, NB_CANONICAL])
            E = E[~np.isnan(E)]
            nb = np.full(len(E), NB_CANONICAL)
        levels = np.unique(np.round(E, 9))
        lev = np.clip(np.searchsorted(levels, np.round(E, 9)), 0, len(levels) - 1)
        _STATES[kind] = (levels, lev, nb, E)
    return _STATES[kind]


def _exact_bins(kind, temperature):
    levels, lev, nb, E = _states(kind)
    beta = 1.0 / (KB * temperature)
    w = E - MU * nb if kind == "sgc" else E
    p = np.exp(-beta * (w - w.min()))
    p /= p.sum()
    probs = np.bincount(lev * 25 + nb, weights=p, minlength=len(levels) * 25)
    return levels, probs


def _chi2_p(E, nB, kind, temperature):
    levels, probs = _exact_bins(kind, temperature)
    e = E[:, BURN::THIN].ravel()
    n = nB[:, BURN::THIN].ravel().astype(int)
    j = np.clip(np.searchsorted(levels, e), 1, len(levels) - 1)
    j = np.where(np.abs(levels[j - 1] - e) < np.abs(levels[j] - e), j - 1, j)
    assert np.max(np.abs(levels[j] - e)) < 1e-6, "sampled energy is not a level of the exact spectrum"
    obs = np.bincount(j * 25 + n, minlength=len(probs)).astype(float)
    exp_ = probs * len(e)
    assert obs[exp_ == 0].sum() == 0, "a state of zero weight was visited"
    big = exp_ >= 5
    o = np.append(obs[big], obs[~big].sum())
    x = np.append(exp_[big], exp_[~big].sum())
    if x[-1] < 5:   # fold the merged rest into the smallest kept bin
        k = int(np.argmin(x[:-1]))
        o[k] += o[-1]
        x[k] += x[-1]
        o, x = o[:-1], x[:-1]
    stat = float(((o - x) ** 2 / x).sum())
    dof = len(o) - 1
    p = float(chi2.sf(stat, dof))
    print(f"  chi2 {stat:.1f} dof {dof} p {p:.3g} ({len(e)} samples)")
    return p


def _run(kind, proposal=None, rounds=2):
    flat = small_flat()
    sc = Supercell(flat.ce.prim, SGC_DIMS if kind == "sgc" else CAN_DIMS)
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=64 if kind == "canonical" else 0, seed=5)
    if proposal is not None:
        n_env = eng.mlp_info()[0]
        eng.mlp_set_params(random_mlp_params(17, n_env, scale=2.0), eps=0.05)
        eng.set_proposal(proposal, rounds_per_group=4)
    if kind == "canonical":
        occs = [random_occupation(sc.n_sites, NB_CANONICAL / sc.n_sites, 300 + c) for c in range(N_CHAINS)]
        eng.load_chains([sid] * N_CHAINS, occs, [T_CAN] * N_CHAINS, 7000 + np.arange(N_CHAINS))
        eng.run_passes("canonical", N_PASSES, 1, canonical_rounds=rounds)
    else:
        occs = [random_occupation(sc.n_sites, 0.5, 100 + c, exact=False) for c in range(N_CHAINS)]
        eng.load_chains([sid] * N_CHAINS, occs, [T_SGC] * N_CHAINS, 5000 + np.arange(N_CHAINS),
                        chem=[[0.0, MU]] * N_CHAINS)
        eng.run_passes("semigrand_canonical", N_PASSES, 1)
    E, cnt = eng.read_samples()
    eng.close()
    return E, cnt[:, :, 1]


def test_plain_sgc_chi2_and_power():
    _print_case("plain sgc chi-square 2x2x4")
    E, nB = _run("sgc")
    assert _chi2_p(E, nB, "sgc", T_SGC) > P_MIN
    print("  power check against 1.2 T:")
    assert _chi2_p(E, nB, "sgc", 1.2 * T_SGC) < 1e-6


@pytest.mark.parametrize("rounds", [2, 0])
def test_plain_canonical_chi2(rounds):
    _print_case(f"plain canonical chi-square 2x2x6 rounds {rounds}")
    E, nB = _run("canonical", rounds=rounds)
    assert np.all(nB == NB_CANONICAL)
    assert _chi2_p(E, nB, "canonical", T_CAN) > P_MIN


@pytest.mark.parametrize("kind", ["sgc", "canonical"])
@pytest.mark.parametrize("proposal", ["learned", "grouped_uniform", "informed"])
def test_grouped_proposal_chi2(proposal, kind):
    _print_case(f"{proposal} {kind} chi-square")
    E, nB = _run(kind, proposal=proposal)
    assert _chi2_p(E, nB, kind, T_SGC if kind == "sgc" else T_CAN) > P_MIN
