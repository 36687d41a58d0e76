"""Benchmark of the Metropolis hot path on one MI355X.

Default workload (the headline configuration): binary FCC cluster expansion
(8 correlations: empty, point, pairs to the 4th shell, triplets within the 2nd
shell), 24x24x24 supercell (13,824 sites), semi-grand canonical flips, 256
chains on a 16 T (200..1500 K) x 16 mu (-0.3..0.3 eV) grid.

A step is one Monte Carlo pass (13,824 site attempts) of every chain.  W warmup
passes run untimed, then K passes run in one persistent launch bracketed by
device synchronisation.  Inputs (occupations, tables) are resident in HBM
before the timed region.

Prints one JSON line (see the keys below).  --gpus N with N > 1 exits: this
build measures one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LDS_PEAK_GBS = 256 * 128 * 2.4e9 / 1e9  # 256 CU x 128 B/clk x 2.4 GHz = 78.6 TB/s


def tau_int(x: np.ndarray, c: float = 5.0) -> float:
    """Integrated autocorrelation time with Sokal's self-consistent window."""
    x = np.asarray(x, dtype=float)
    n = len(x)
    if n < 4 or np.var(x) == 0:
        return 0.5
    y = x - x.mean()
    f = np.fft.rfft(y, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    acf /= acf[0]
    tau = 0.5
    for M in range(1, n):
        tau = 0.5 + acf[1:M + 1].sum()
        if M >= c * tau:
            break
    return max(tau, 0.5)


def cpu_baseline(flat, sc, batch, n_chains: int, n_passes: int, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # CPU restatement of the reference loop (timed baseline)
    osys = O.OracleSystem(flat, sc.dims)
    occ = np.stack([batch.occ[i] for i in range(n_chains)])
    r = osys.run_batch(occ, "sgc", batch.temperatures[:n_chains], batch.chem[:n_chains], n_passes,
                       batch.seeds[:n_chains], threads, full_recompute_samplers=True)
    return r["n_attempt"] / r["elapsed_s"], r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-chains", type=int, default=32)
    ap.add_argument("--cpu-passes", type=int, default=40)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="copy N MiB device-to-device (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        # a plain copy of known size: FETCH_SIZE / WRITE_SIZE of this run
        # calibrate the counters read on the Metropolis kernel
        import torch
        n = args.calibrate_copy << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": n}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1"}))
        sys.exit(2)

    from casmcode_clexmonte_amd.engine import Engine
    from casmcode_clexmonte_amd.system import FlatCE, Supercell
    from casmcode_clexmonte_amd.workloads import fcc_binary_ce, fcc_sgc_grid

    ce = fcc_binary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (24, 24, 24))
    batch = fcc_sgc_grid(sc)
    n_chains = len(batch.temperatures)
    eng = Engine(flat, device=0)
    if args.block:
        eng.set_block_size(args.block)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    eng.load_chains([sid] * n_chains, batch.occ, batch.temperatures, batch.seeds, chem=batch.chem)
    if args.warmup > 0:
        eng.run_passes("semigrand_canonical", args.warmup, 1)
    eng.clear_samples()
    att0 = sum(eng.read_state(c)["n_attempt"] for c in range(n_chains))
    # clexmc_run_passes returns after the launch has completed on the device
    # (event synchronisation on the handle's stream), so the wall clock below
    # brackets exactly the K timed passes.
    t0 = time.perf_counter()
    kernel_ms = eng.run_passes("semigrand_canonical", args.steps, 1)
    wall = time.perf_counter() - t0
    att1 = sum(eng.read_state(c)["n_attempt"] for c in range(n_chains))
    attempts = att1 - att0
    expected = n_chains * sc.n_sites * args.steps
    assert attempts == expected, (attempts, expected)
    value = attempts / wall
    # effective sample size of E per chain from the per-pass samples
    E, cnt = eng.read_samples()
    taus = np.array([tau_int(E[c]) for c in range(n_chains)])
    ess_total = float(np.sum(E.shape[1] / (2.0 * taus)))
    bytes_per_attempt = 1 + 54  # the site and its 54 interaction neighbours, int8 in LDS
    achieved = attempts * bytes_per_attempt / (kernel_ms * 1e-3) / 1e9
    out = {
        "metric": "MC site-attempts/s (FCC binary 24^3 SGC, 256 chains, 16 T x 16 mu)",
        "value": value,
        "unit": "site-attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded ECIs and random initial occupations)",
        "config": {"workload": "fcc_binary_24x24x24_sgc_256chains_T200-1500_mu-0.3-0.3",
                   "sites_per_chain": sc.n_sites, "chains": n_chains, "passes": args.steps,
                   "cluster_expansion": "FCC Chebyshev binary, pairs<=4NN, triplets<=2NN, 8 correlations"},
        "roofline": {"bound": "lds", "achieved": achieved, "peak": LDS_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / LDS_PEAK_GBS, "traffic": None,
                     "bytes_per_attempt": bytes_per_attempt, "kernel_ms": kernel_ms},
        "ess_per_s": ess_total / wall,
        "tau_int_E_median_passes": float(np.median(taus)),
        "acceptance": float(sum(eng.read_state(c)["n_accept"] for c in range(n_chains)) / max(1, att1)),
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu_rate, r = cpu_baseline(flat, sc, batch, args.cpu_chains, args.cpu_passes, threads)
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "site-attempts/s", "cores": threads, "kind": "port",
                               "sample": f"{args.cpu_chains} chains x {args.cpu_passes} passes of the same "
                                         f"workload, oracle restatement of occupation_metropolis_v2 "
                                         f"(FP64, mt19937_64, per-pass full-recompute samplers), "
                                         f"{r['elapsed_s']:.1f} s wall"}
        out["speedup_vs_cpu"] = value / cpu_rate
    print(json.dumps(out))
    eng.close()


if __name__ == "__main__":
    main()
