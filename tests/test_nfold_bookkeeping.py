"""CPU emulation of the N-fold kernel's rate bookkeeping (numpy, FP32 rates).

The kernel keeps one FP32 rate per site and, per group of 64 sites, the
FP64 sum of the group's rates, formed again from the 64 rates whenever one of
them changes (never updated by differences).  The emulation runs the events
of independent binary sites at beta dE = 20 (rates ~2e-9; a flipped site's
way back has rate 1), which is where difference updates cancel: it checks
that the formed total stays equal to the exact sum of the rates, and that
the same events with FP32 difference updates of the group sums (the
bookkeeping this replaced) leave the total about 10 % off after the first
excursion (a group's 64 x 2e-9 does not survive + 1 - 1 in FP32).
"""
import numpy as np


def _events(n_sites=1728, n_events=400, beta_d=20.0, seed=3, difference_updates=False):
    rng = np.random.default_rng(seed)
    occ = np.zeros(n_sites, dtype=np.int8)
    p = np.float32(np.exp(-beta_d))
    rate = np.full(((n_sites + 63) // 64) * 64, 0.0, dtype=np.float32)
    rate[:n_sites] = p
    groups = rate.reshape(-1, 64)
    G = groups.astype(np.float64).sum(axis=1)
    G32 = groups.sum(axis=1, dtype=np.float32)
    worst = 0.0
    for _ in range(n_events):
        true = rate.astype(np.float64).sum()
        R = float(G32.astype(np.float64).sum()) if difference_updates else float(G.sum())
        worst = max(worst, abs(R - true) / true)
        if not R > 0:
            return np.inf
        s = rng.choice(n_sites, p=rate[:n_sites].astype(np.float64) / true)
        occ[s] ^= 1
        new = np.float32(1.0) if occ[s] else p
        if difference_updates:
            G32[s // 64] = np.float32(G32[s // 64] + np.float32(new - rate[s]))
        rate[s] = new
        g = s // 64
        G[g] = rate[g * 64:(g + 1) * 64].astype(np.float64).sum()
    return worst


def test_group_sums_formed_from_rates_track_the_total():
    assert _events() <= 1e-12


def test_fp32_difference_updates_lose_the_total():
    # the replaced bookkeeping: one excursion costs about 10 % of the total
    assert _events(difference_updates=True) > 0.05
