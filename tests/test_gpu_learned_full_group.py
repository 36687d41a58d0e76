"""The grouped weighted proposals at full group size (64 sites per group),
on the ordering workload's supercell (FCC 32^3, 128 chains, 32 proposals per
group and class), against closed forms.

The cluster expansion keeps the workload's basis (so the network reads the
same 42-site environment) with every ECI zero: the energy is constant, the
proposal weights are the only thing that depends on the configuration, and
a proposal ratio that is not exact shows up as a bias of the sampled law.

* Semi-grand canonical, chem = [0, mu] with beta mu = -ln 3: sites are
  independent with <x_B> = 1 / (1 + exp(-beta mu)) = 1/4 and nearest-
  neighbour <s_i s_j> = (2 x_B - 1)^2 = 1/4 (s = -1 for A, +1 for B).
* Canonical at x_B = 1/4 (N_B = 8192 of N = 32768): the stationary law is
  uniform over arrangements, so for any two distinct sites
  <s_i s_j> = ((N_A - N_B)^2 - N) / (N (N - 1)).

Weights: the network with random weights (scale 2), the network trained on
the ordering CE by the device trainer (as in the ESS comparison), and for
the semi-grand canonical case the informed weights (eps + sigmoid(-beta e_s),
which here depend on the site's occupant through mu).  Errors: standard
error over the 128 independent chains; 3 sigma.
"""
import numpy as np
import pytest

from casmcode_clexmonte_amd.engine import Engine
from casmcode_clexmonte_amd.system import ClusterExpansion, FlatCE, Supercell
from casmcode_clexmonte_amd.workloads import (fcc_binary_basis, fcc_ordering_ce, l12_occupation,
                                              random_mlp_params, random_occupation)

pytestmark = pytest.mark.gpu

KB = 8.6173303e-05
NSIG = 3.0
L = 32
NCH = 128
ROUNDS = 32


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _nn_table(sc):
    """Linear indices of the 6 nearest neighbours per site in one direction
    of each of the 6 nearest-neighbour bond directions (each bond once)."""
    prim = sc.prim
    lat = np.array(prim.lattice, dtype=float)
    nn = []
    for d in np.ndindex(3, 3, 3):
        v = np.array(d) - 1
        if abs(np.linalg.norm(v @ lat) - 1.0 / np.sqrt(2.0)) < 1e-6 and tuple(-v) not in {tuple(x) for x in nn}:
            nn.append(v)
    assert len(nn) == 6
    L1, L2, L3 = sc.dims
    l = np.arange(sc.n_sites)
    i, j, k = l % L1, (l // L1) % L2, l // (L1 * L2)
    return np.stack([((i + v[0]) % L1) + L1 * (((j + v[1]) % L2) + L2 * ((k + v[2]) % L3)) for v in nn], axis=1)


@pytest.fixture(scope="module")
def trained_params():
    """The network trained by the device trainer on plain-sampler
    configurations of the ordering CE (canonical, x_B = 1/4, below T_c)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "fcc_ordering_tc_scan.json")) as f:
        Tc = float(json.load(f)["Tc"])
    ce = fcc_ordering_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (16, 16, 16))
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=16, seed=2)
    nch = 32
    temps = Tc * np.linspace(0.8, 1.2, nch)
    occs = [l12_occupation(sc) if T < Tc else random_occupation(sc.n_sites, 0.25, 60 + c) for c, T in enumerate(temps)]
    eng.load_chains([sid] * nch, occs, temps, 600 + np.arange(nch))
    eng.run_passes("canonical", 400, 400, canonical_rounds=2)
    loss = eng.mlp_train(n_per_chain=512, n_steps=200, lr=3e-3, seed=5)
    assert loss[-1] < loss[0]
    p = eng.mlp_get_params()
    eng.close()
    return p


def _zero_ce_engine():
    bs = fcc_binary_basis()
    flat = FlatCE(ClusterExpansion(bs, np.zeros(bs.n_functions)))
    sc = Supercell(bs.prim, (L, L, L))
    eng = Engine(flat)
    sid = eng.add_supercell(sc, n_pair_phases=32, seed=11)
    return eng, sc, sid


def _weights(eng, which, trained):
    if which == "trained":
        eng.mlp_set_params(trained, eps=0.05)
    else:
        eng.mlp_set_params(random_mlp_params(29, eng.mlp_info()[0], scale=2.0), eps=0.05)


def _nn_corr(eng, sc, nbr):
    out = np.zeros(NCH)
    for c in range(NCH):
        s = 2.0 * eng.read_state(c)["occ"].astype(np.float64) - 1.0
        out[c] = float(np.mean(s[:, None] * s[nbr]))
    return out


def _check(name, per_chain, exact):
    m, se = per_chain.mean(), per_chain.std(ddof=1) / np.sqrt(len(per_chain))
    print(f"{name}: {m:.6f} +- {se:.6f}  closed form {exact:.6f}  z {(m - exact) / se:+.2f}")
    assert abs(m - exact) <= NSIG * se, (name, m, exact, se)


@pytest.mark.parametrize("proposal,weights", [("learned", "random"), ("learned", "trained"), ("informed", None)])
def test_full_group_sgc_closed_form(proposal, weights, trained_params):
    _print_case(f"full group sgc {proposal} {weights}")
    eng, sc, sid = _zero_ce_engine()
    nbr = _nn_table(sc)
    T = 300.0
    mu = -KB * T * np.log(3.0)
    x_exact = 1.0 / (1.0 + np.exp(-mu / (KB * T)))
    _weights(eng, weights, trained_params)
    eng.set_proposal(proposal, rounds_per_group=ROUNDS)
    occs = [random_occupation(sc.n_sites, 0.5, 100 + c, exact=False) for c in range(NCH)]
    eng.load_chains([sid] * NCH, occs, [T] * NCH, 5000 + np.arange(NCH), chem=[[0.0, mu]] * NCH)
    eng.run_passes("semigrand_canonical", 100, 100)   # burn-in
    eng.clear_samples()
    corr = []
    for seg in range(6):
        eng.run_passes("semigrand_canonical", 50, 1)
        corr.append(_nn_corr(eng, sc, nbr))
    _, cnt = eng.read_samples()
    x = cnt[:, :, 1].mean(1) / sc.n_sites
    _check("x_B", x, x_exact)
    _check("nn <s s>", np.mean(corr, axis=0), (2 * x_exact - 1) ** 2)
    eng.close()


@pytest.mark.parametrize("weights", ["random", "trained"])
def test_full_group_canonical_closed_form(weights, trained_params):
    _print_case(f"full group canonical learned {weights}")
    eng, sc, sid = _zero_ce_engine()
    nbr = _nn_table(sc)
    N = sc.n_sites
    nb = N // 4
    exact = ((N - 2 * nb) ** 2 - N) / (N * (N - 1.0))
    _weights(eng, weights, trained_params)
    eng.set_proposal("learned", rounds_per_group=ROUNDS)
    occs = [random_occupation(N, nb / N, 300 + c) for c in range(NCH)]
    eng.load_chains([sid] * NCH, occs, [300.0] * NCH, 7000 + np.arange(NCH))
    eng.run_passes("canonical", 100, 100, canonical_rounds=2)
    corr = []
    for seg in range(6):
        eng.run_passes("canonical", 50, 50, canonical_rounds=2)
        corr.append(_nn_corr(eng, sc, nbr))
    for c in range(NCH):
        assert int(eng.read_state(c)["occ"].sum()) == nb
    _check("nn <s s>", np.mean(corr, axis=0), exact)
    eng.close()
