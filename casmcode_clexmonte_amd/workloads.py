"""Synthetic cluster expansions and chain batches for the measured workloads.

* FCC binary (A, B), 1 site per primitive cell, Chebyshev basis, clusters:
  empty, point, pairs to the 4th neighbour shell, triplets with every edge
  within the 2nd shell (8 correlation functions).  ECIs are drawn
  ``N(0, 0.05 eV) / (1 + L)`` with ``L`` the cluster's maximum length in
  nearest-neighbour units, from a recorded seed.
* ``fcc_canonical_12``:  12x12x12 supercell (1,728 sites), canonical swaps at
  x_B = 0.5, T = 600 K, one chain, 1,000 + 5,000 passes.
* ``fcc_sgc_24_grid``:   24x24x24 supercell (13,824 sites), semi-grand
  canonical flips, 256 chains on a 16 T (200..1500 K) x 16 mu (-0.3..0.3 eV)
  grid, 2,000 passes.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from .system import (BasisSet, ClusterExpansion, FlatCE, Prim, Supercell, bcc_prim, fcc_prim, hcp_prim, KB)

FCC_ECI_SEED = 20240611
FCC_NN = 1.0 / math.sqrt(2.0)


def fcc_binary_basis() -> BasisSet:
    prim = fcc_prim(1.0, ("A", "B"))
    # 4th-neighbour pair shell at sqrt(2) a, 5th at sqrt(5/2) a; 2nd shell at a
    return BasisSet(prim, {2: 1.5, 3: 1.05}, "chebychev")


def seeded_eci(bs: BasisSet, seed: int, sigma: float = 0.05, nn: float = FCC_NN) -> np.ndarray:
    rng = np.random.default_rng(seed)
    eci = np.zeros(bs.n_functions)
    for orb in bs.orbits:
        L = orb.max_length / nn
        for fi in orb.functions:
            eci[fi] = rng.normal(0.0, sigma) / (1.0 + L)
    return eci


def fcc_binary_ce(seed: int = FCC_ECI_SEED) -> ClusterExpansion:
    bs = fcc_binary_basis()
    return ClusterExpansion(bs, seeded_eci(bs, seed))


def random_occupation(n_sites: int, x_b: float, seed: int, exact: bool = True) -> np.ndarray:
    """Binary occupation with a fraction x_b of B (occupant index 1);
    exact=True shuffles exactly round(x_b*N) B atoms (Fisher-Yates)."""
    rng = np.random.default_rng(seed)
    if exact:
        occ = np.zeros(n_sites, dtype=np.int8)
        occ[: int(round(x_b * n_sites))] = 1
        rng.shuffle(occ)
        return occ
    return (rng.random(n_sites) < x_b).astype(np.int8)


@dataclass
class ChainBatch:
    temperatures: np.ndarray
    chem: np.ndarray        # [n_chains][n_sublat * n_occ_max]
    occ: List[np.ndarray]
    seeds: np.ndarray
    mu: np.ndarray


def fcc_sgc_grid(sc: Supercell, n_T: int = 16, n_mu: int = 16, T_range=(200.0, 1500.0),
                 mu_range=(-0.3, 0.3), seed: int = 7) -> ChainBatch:
    Ts = np.linspace(T_range[0], T_range[1], n_T)
    mus = np.linspace(mu_range[0], mu_range[1], n_mu)
    temps, mu_list, chem, occs, seeds = [], [], [], [], []
    c = 0
    for T in Ts:
        for mu in mus:
            temps.append(T)
            mu_list.append(mu)
            # Omega = E - mu * N_B  ->  weight of occupant B is mu
            chem.append([0.0, mu])
            occs.append(random_occupation(sc.n_sites, 0.5, seed * 100003 + c, exact=False))
            seeds.append((seed << 32) + c)
            c += 1
    return ChainBatch(np.array(temps), np.array(chem), occs, np.array(seeds, dtype=np.uint64),
                      np.array(mu_list))


# ---------------------------------------------------------------------------
# Learned-proposal workload: ordering-favouring FCC binary CE
# ---------------------------------------------------------------------------
ORDERING_ECI_SEED = 20240612


def fcc_ordering_ce(seed: int = ORDERING_ECI_SEED, nn_pair: float = 0.04, nnn_pair: float = -0.004,
                    sigma: float = 0.001) -> ClusterExpansion:
    """Binary FCC (same clusters as fcc_binary_ce) with ordering ECIs: a
    positive nearest-neighbour pair ECI (+0.04 eV, favouring unlike
    neighbours), a weak negative second-neighbour pair ECI (-0.004 eV; it
    lifts the degeneracy of the nearest-neighbour model at x_B = 1/4 in
    favour of L1_2, whose second-neighbour pairs are all like pairs) and
    small seeded ECIs N(0, 0.001 eV) on the other pairs and triplets."""
    bs = fcc_binary_basis()
    rng = np.random.default_rng(seed)
    eci = np.zeros(bs.n_functions)
    pairs = sorted((o for o in bs.orbits if o.size == 2), key=lambda o: o.max_length)
    for orb in bs.orbits:
        for fi in orb.functions:
            if orb is pairs[0]:
                eci[fi] = nn_pair
            elif orb is pairs[1]:
                eci[fi] = nnn_pair
            elif orb.size >= 2:
                eci[fi] = rng.normal(0.0, sigma)
    return ClusterExpansion(bs, eci)


def random_mlp_params(seed: int, n_env: int, scale: float = 0.3) -> np.ndarray:
    """Random network parameters [W1 (64x48) | b1 | W2 | b2] (FP32)."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((64, 48))
    W1[:, :n_env] = rng.normal(0.0, scale / np.sqrt(max(n_env, 1)), (64, n_env))
    b1 = rng.normal(0.0, 0.1, 64)
    W2 = rng.normal(0.0, scale, 64)
    b2 = np.array([0.0])
    return np.concatenate([W1.ravel(), b1, W2, b2]).astype(np.float32)


def l12_occupation(sc: Supercell) -> np.ndarray:
    """Perfect L1_2 (A3B) arrangement on an FCC supercell: B (occupant 1) on
    the sites whose Cartesian coordinates are integer multiples of the
    conventional lattice parameter, A on the face centres.  Needs even
    supercell dimensions."""
    prim = sc.prim
    a = 2.0 * abs(prim.lattice[0][1])
    occ = np.zeros(sc.n_sites, dtype=np.int8)
    L1, L2, _ = sc.dims
    for l in range(sc.n_sites):
        c = l % sc.n_cells
        i, j, k = c % L1, (c // L1) % L2, c // (L1 * L2)
        r = prim.cart(0, (i, j, k)) / a
        if np.allclose(r, np.round(r), atol=1e-9):
            occ[l] = 1
    return occ


def random_occupation_counts(counts, seed: int) -> np.ndarray:
    """Shuffled occupation with exactly counts[o] sites of occupant o."""
    rng = np.random.default_rng(seed)
    occ = np.concatenate([np.full(int(n), o, dtype=np.int8) for o, n in enumerate(counts)])
    rng.shuffle(occ)
    return occ


def composition_counts(n_sites: int, x) -> List[int]:
    """Integer occupant counts closest to fractions x (summing to n_sites)."""
    c = [int(round(xi * n_sites)) for xi in x[:-1]]
    return c + [n_sites - sum(c)]


# ---------------------------------------------------------------------------
# Ternary HCP workload: uneven supercells in one launch
# ---------------------------------------------------------------------------
HCP_ECI_SEED = 20240613
HCP_SUPERCELLS = [(16, 16, 10), (16, 16, 12), (20, 20, 12), (24, 24, 16),
                  (32, 32, 20), (36, 36, 24), (40, 40, 24), (48, 48, 32)]
HCP_COMPOSITION = (0.4, 0.3, 0.3)


def hcp_ternary_basis() -> BasisSet:
    """Ternary (A, B, C) HCP, ideal c/a, 2 sites per primitive cell,
    Chebyshev basis (2 site functions); pairs to the 6th neighbour shell
    (2a), triplets with edges within the 3rd shell (c = 1.633a),
    quadruplets with edges within the 2nd shell (sqrt(2)a): 99 functions."""
    prim = hcp_prim(1.0, occupants=("A", "B", "C"))
    return BasisSet(prim, {2: 2.01, 3: 1.64, 4: 1.42}, "chebychev")


def hcp_ternary_ce(seed: int = HCP_ECI_SEED) -> ClusterExpansion:
    bs = hcp_ternary_basis()
    return ClusterExpansion(bs, seeded_eci(bs, seed, sigma=0.05, nn=1.0))


@dataclass
class MixedBatch:
    dims: List[Tuple[int, int, int]]   # supercell of each chain
    temperatures: np.ndarray
    occ: List[np.ndarray]
    seeds: np.ndarray


def hcp_canonical_batch(prim: Prim, n_chains: int = 64, T_range=(600.0, 1800.0), seed: int = 13) -> MixedBatch:
    """Chains cycle through HCP_SUPERCELLS (16x16x10 ... 48x48x32, 5,120 to
    147,456 sites), canonical at x = (0.4, 0.3, 0.3), temperatures
    linspace(600, 1800) K."""
    dims, occs = [], []
    for c in range(n_chains):
        d = HCP_SUPERCELLS[c % len(HCP_SUPERCELLS)]
        n = prim.n_sublat * d[0] * d[1] * d[2]
        dims.append(d)
        occs.append(random_occupation_counts(composition_counts(n, HCP_COMPOSITION), seed * 1000003 + c))
    temps = np.linspace(T_range[0], T_range[1], n_chains)
    return MixedBatch(dims, temps, occs, (np.uint64(seed) << np.uint64(32)) + np.arange(n_chains, dtype=np.uint64))


# ---------------------------------------------------------------------------
# Quaternary BCC workload: 64^3 = 262,144 sites per chain
# ---------------------------------------------------------------------------
BCC_ECI_SEED = 20240614
BCC_NN = math.sqrt(3.0) / 2.0


def bcc_quaternary_basis() -> BasisSet:
    """Quaternary (A, B, C, D) BCC, 1 site per primitive cell, Chebyshev
    basis (3 site functions); pairs to the 8th neighbour shell (sqrt(5)a,
    112 neighbours), triplets with edges within the 3rd shell (sqrt(2)a),
    quadruplets with edges within the 2nd shell (a): 137 functions."""
    prim = bcc_prim(1.0, ("A", "B", "C", "D"))
    return BasisSet(prim, {2: 2.24, 3: 1.42, 4: 1.01}, "chebychev")


def bcc_quaternary_ce(seed: int = BCC_ECI_SEED) -> ClusterExpansion:
    bs = bcc_quaternary_basis()
    return ClusterExpansion(bs, seeded_eci(bs, seed, sigma=0.05, nn=BCC_NN))


def bcc_equiatomic_batch(sc: Supercell, n_chains: int = 32, T_range=(300.0, 2000.0), seed: int = 17) -> ChainBatch:
    """Canonical equiatomic chains, temperatures linspace(300, 2000) K."""
    counts = composition_counts(sc.n_sites, (0.25, 0.25, 0.25, 0.25))
    occs = [random_occupation_counts(counts, seed * 1000003 + c) for c in range(n_chains)]
    temps = np.linspace(T_range[0], T_range[1], n_chains)
    return ChainBatch(temps, np.zeros((n_chains, 4)), occs,
                      (np.uint64(seed) << np.uint64(32)) + np.arange(n_chains, dtype=np.uint64), np.zeros(n_chains))
