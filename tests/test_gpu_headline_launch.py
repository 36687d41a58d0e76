"""The launch the benchmark times: FCC binary 24^3 (13,824 sites), semi-grand
canonical, 256 chains on the 16 T (200..1500 K) x 16 mu (-0.3..0.3 eV)
grid, built exactly as bench.py builds it (fcc_binary_ce, fcc_sgc_grid,
n_pair_phases = 0, every launch choice automatic: run-time specialised
kernel, int8 image with the modelled row pitch, one lane per class site).

* After the bench's 20-pass warm-up and 50 sampled passes: every chain's
  exact int64 cluster counts equal the oracle's full FP64 recompute of its
  final occupation (rounded to integers), its running energy equals the
  oracle's energy (1e-9 relative), and its last sample's energy and occupant
  counts equal the final state's.
* Continuing to 2,000 passes, the four grid corners (200 K and 1500 K at
  mu = -0.3 and +0.3 eV) against the oracle's restatement of the reference
  loop (SemiGrandCanonicalCalculator.cc:423-466, analysis_functions.cc:
  17-118), 3 chains of 150 + 1,000 passes per corner: <E>/cell and <x_B> within
  3 sigma (autocorrelation-corrected errors), heat capacity and
  susceptibility within 3 sigma (errors from 10 blocks per chain).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
from casmcode_clexmonte_amd.analysis import mean_and_error
from casmcode_clexmonte_amd.engine import Engine
from casmcode_clexmonte_amd.system import FlatCE, Supercell
from casmcode_clexmonte_amd.workloads import fcc_binary_ce, fcc_sgc_grid, random_occupation

pytestmark = pytest.mark.gpu

KB = 8.6173303e-05
NSIG = 3.0
CORNERS = (0, 15, 240, 255)


def _print_case(name):
    print(f"[case] {name}", flush=True)


def _block_stats(pe, x, T, nc, n_blocks=10):
    """Heat capacity and susceptibility per block of a chain's series."""
    C, chi = [], []
    for pb, xb in zip(np.array_split(pe, n_blocks), np.array_split(x, n_blocks)):
        C.append(np.var(pb) * nc / (KB * T * T))
        chi.append(np.var(xb) * nc / (KB * T))
    return C, chi


def _summary(E_sc, n1, T, mu, nc):
    """Means with errors, and blocked C and chi, of one or more series."""
    e = [es / nc for es in E_sc]
    x = [n / nc for n in n1]
    pe = [a - mu * b for a, b in zip(e, x)]
    out = {}
    for name, ser in (("E", e), ("x", x)):
        ms = [mean_and_error(s) for s in ser]
        m = float(np.mean([a for a, _ in ms]))
        err = float(np.sqrt(np.sum([b * b for _, b in ms])) / len(ms))
        out[name] = (m, err)
    C, chi = [], []
    for p, xx in zip(pe, x):
        c1, c2 = _block_stats(p, xx, T, nc)
        C += c1
        chi += c2
    for name, v in (("C", C), ("chi", chi)):
        v = np.array(v)
        out[name] = (float(v.mean()), float(v.std(ddof=1) / np.sqrt(len(v))))
    return out


def test_headline_launch_parity():
    _print_case("headline launch: fcc 24^3 sgc 256 chains")
    ce = fcc_binary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (24, 24, 24))
    nc = sc.n_cells
    batch = fcc_sgc_grid(sc)
    n = len(batch.temperatures)
    assert n == 256
    eng = Engine(flat, device=0)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    eng.load_chains([sid] * n, batch.occ, batch.temperatures, batch.seeds, chem=batch.chem)
    eng.run_passes("semigrand_canonical", 20, 1)
    eng.clear_samples()
    eng.run_passes("semigrand_canonical", 50, 1)
    info = eng.last_run_info()
    print("launch", info, flush=True)
    assert info["specialized"] and not info["packed"] and info["threads"] == 512
    E, cnt = eng.read_samples()
    assert E.shape == (n, 50)
    osys = O.OracleSystem(flat, sc.dims)
    for c in range(n):
        st = eng.read_state(c)
        C = eng.read_binary_counts(c)
        corr = osys.corr(st["occ"])
        assert np.array_equal(C, np.rint(corr * flat.orbit_mult).astype(np.int64)), c
        Eref = osys.energy(st["occ"])
        assert abs(st["E"] - Eref) <= 1e-9 * max(1.0, abs(Eref)), (c, st["E"], Eref)
        assert E[c, -1] == st["E"]
        n1 = int(st["occ"].sum())
        assert cnt[c, -1, 1] == n1 and cnt[c, -1, 0] == sc.n_sites - n1
    # continue to 2,000 passes; the corners against the oracle loop
    eng.clear_samples()
    eng.run_passes("semigrand_canonical", 1930, 1)
    E, cnt = eng.read_samples()
    eng.close()
    burn = 400
    runs = []
    for c in CORNERS:
        for r in range(3):
            runs.append((c, r))

    def orc(cr):
        c, r = cr
        occ = random_occupation(sc.n_sites, 0.5, 9100 + 7 * c + r, exact=False)
        return O.OracleSystem(flat, sc.dims).run(occ, "sgc", float(batch.temperatures[c]), 150, 1000, 31 + 131 * c + r,
                                                 chem=batch.chem[c], full_recompute_samplers=False)

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        res = list(ex.map(orc, runs))
    for c in CORNERS:
        T, mu = float(batch.temperatures[c]), float(batch.mu[c])
        g = _summary([E[c, burn:]], [cnt[c, burn:, 1].astype(float)], T, mu, nc)
        rr = [r for (cc, _), r in zip(runs, res) if cc == c]
        o = _summary([r["E"] for r in rr], [r["counts"][:, 1].astype(float) for r in rr], T, mu, nc)
        for k in ("E", "x", "C", "chi"):
            (gm, ge), (om, oe) = g[k], o[k]
            tol = NSIG * np.sqrt(ge * ge + oe * oe)
            print(f"T {T:6.1f} mu {mu:+.2f} {k:3s}: gpu {gm:.6g} +- {ge:.2g}  oracle {om:.6g} +- {oe:.2g}",
                  flush=True)
            assert abs(gm - om) <= tol + 1e-12, (c, k, gm, om, tol)
