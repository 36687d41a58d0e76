"""Benchmark of the Metropolis hot path on one MI355X.

Default workload (the headline configuration): binary FCC cluster expansion
(8 correlations: empty, point, pairs to the 4th shell, triplets within the 2nd
shell), 24x24x24 supercell (13,824 sites), semi-grand canonical flips, 256
chains on a 16 T (200..1500 K) x 16 mu (-0.3..0.3 eV) grid.

A step is one Monte Carlo pass (13,824 site attempts) of every chain.  W warmup
passes run untimed, then K passes run in one persistent launch bracketed by
device synchronisation.  Inputs (occupations, tables) are resident in HBM
before the timed region.

Second workload (effective samples per second, learned versus plain
proposal): binary FCC ordering CE (nearest-neighbour pair +0.04 eV), x_B =
0.25, 32x32x32 supercell (32,768 sites), 128 canonical chains: 16
temperatures bracketing the L1_2 ordering temperature T_c (0.7..1.3 T_c;
T_c from the CPU oracle's heat-capacity scan in
tests/golden/fcc_ordering_tc_scan.json) x 8 independent replicas.  The plain
sampler equilibrates the chains and generates the training data; the
network is trained on the device; then each proposal runs the same number of
sampled passes and ESS/s = sum over chains of n_samples / (2 tau_int(E)) per
second of sampling.  Every proposal starts from the same equilibrated
configurations; per temperature, <E> is compared between proposals with
errors from the spread over replicas.

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

# LDS read peak of MI355X (CDNA4: 160 KB LDS per CU with twice the CDNA3 read
# bandwidth, 256 B/clk): 256 CU x 256 B/clk x 2.4 GHz = 157.3 TB/s.  bench.py
# also measures the device's own LDS throughput (16-byte reads and byte gathers).
LDS_PEAK_GBS = 256 * 256 * 2.4e9 / 1e9
# VALU issue peak for unpacked 32-bit ops: 256 CUs x 64 lanes x 2.4 GHz
VALU_PEAK_LANE_OPS = 256 * 64 * 2.4e9
# VALU lane-ops per attempt of the specialised kernel on this workload (the
# register-cube sweep; SQ_INSTS_VALU x 64 / attempts,
# profiles/metropolis_specialized_fcc24_sgc256.md)
VALU_OPS_PER_ATTEMPT = 283
# HBM bytes per launch of that kernel, measured with rocprofv3 on the 20-pass
# launch of this workload: FETCH_SIZE 2,617 KiB x 2.000 + WRITE_SIZE 3,652 KiB
# x 1.000, the factors from the same counters on a plain 256 MiB copy
# (profiles/metropolis_specialized_fcc24_sgc256.md).  About 9 MB whatever the
# number of passes: each chain's occupation read once into LDS and written
# back once, plus state, tables and samples.
HBM_BYTES_PER_LAUNCH = (2617 * 2 + 3652) * 1024


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


def host_cpu():
    """The host CPUs the CPU baselines run on: model, CPUs visible to this
    process, threads used (the process's CPU share: OMP_NUM_THREADS when set,
    which the GPU hosts set to the CPUs granted per GPU, else every visible
    CPU) and the oracle's compiler."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    visible = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(visible, int(env)) if env.isdigit() and int(env) > 0 else visible
    try:
        import subprocess
        cc = subprocess.run(["gcc", "--version"], capture_output=True, text=True).stdout.splitlines()[0]
    except (OSError, IndexError):
        cc = "gcc"
    return {"model": model, "cpus_visible": visible, "threads": threads,
            "compiler": f"{cc}; oracle/Makefile: -O3 -march=x86-64-v3 (built where the library is built, run here)"}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # CPU restatement of the reference loop (timed baseline)
    return O


def cpu_baseline(flat, sc, batch, n_chains: int, n_passes: int, threads: int):
    O = _oracle()
    osys = O.OracleSystem(flat, sc.dims)
    occ = np.stack([batch.occ[i] for i in range(n_chains)])
    r = osys.run_batch(occ, "sgc", batch.temperatures[:n_chains], batch.chem[:n_chains], n_passes,
                       batch.seeds[:n_chains], threads, full_recompute_samplers=True)
    return r["n_attempt"] / r["elapsed_s"], r


def cpu_baseline_chains(flat, jobs, threads: int):
    """Independent chains of the oracle's reference loop spread over host
    threads (one chain per task, no shared state; the foreign calls release
    the GIL).  jobs: (dims, occ, kind, temperature, chem, n_passes, seed).
    Returns site-attempts/s over the wall time of all chains and the wall time."""
    from concurrent.futures import ThreadPoolExecutor
    O = _oracle()
    systems = {}
    for j in jobs:
        if j[0] not in systems:
            systems[j[0]] = O.OracleSystem(flat, j[0])

    def one(j):
        dims, occ, kind, T, chem, n_passes, seed = j
        r = systems[dims].run(occ, kind, float(T), 0, n_passes, int(seed), chem=chem, full_recompute_samplers=True)
        return r["n_attempt"]

    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        att = sum(ex.map(one, jobs))
    wall = time.perf_counter() - t0
    return att / wall, wall


def learned_comparison(args):
    """ESS/s of the plain, learned and informed proposals on the ordering
    workload (see the module docstring), and their agreement on <E>.

    The 128 chains are 16 temperatures x 8 independent replicas (own initial
    state, own plain-sampler equilibration, own seeds).  Per temperature and
    proposal, the mean energy is the mean over replicas of each replica's
    time average and its error the replicas' standard error, so the error
    needs no autocorrelation estimate.  Two proposals agree at a temperature
    when Welch's t of their means stays below the two-sided critical value
    for a family-wise error of 1 % over the 16 temperatures (Bonferroni).
    ESS per chain is n / (2 tau_int(E)) with Sokal's window; the run length
    is reported in units of the longest tau_int."""
    from scipy import stats

    from casmcode_clexmonte_amd.engine import Engine
    from casmcode_clexmonte_amd.system import FlatCE, Supercell
    from casmcode_clexmonte_amd.workloads import fcc_ordering_ce, l12_occupation, random_occupation

    with open(os.path.join(ROOT, "tests", "golden", "fcc_ordering_tc_scan.json")) as f:
        Tc = float(json.load(f)["Tc"])
    ce = fcc_ordering_ce()
    flat = FlatCE(ce)
    L = args.c4_size
    sc = Supercell(ce.prim, (L, L, L))
    nT, nrep = args.c4_temps, args.c4_replicas
    nch = nT * nrep
    tgrid = Tc * np.linspace(0.7, 1.3, nT)
    temps = np.repeat(tgrid, nrep)
    occ_ord = l12_occupation(sc)
    occs = [occ_ord if T < Tc else random_occupation(sc.n_sites, 0.25, 4000 + c) for c, T in enumerate(temps)]
    eng = Engine(flat, device=0)
    sid = eng.add_supercell(sc, n_pair_phases=32, seed=11)
    eng.load_chains([sid] * nch, occs, temps, 77000 + np.arange(nch))
    # automatic matchings and pair phases (the calculators' default)
    eng.run_passes("canonical", args.c4_equil, args.c4_equil, canonical_rounds=0)
    t0 = time.perf_counter()
    loss = eng.mlp_train(n_per_chain=args.c4_train_samples, n_steps=args.c4_train_steps, lr=3e-3, seed=5)
    train_s = time.perf_counter() - t0
    # every proposal starts from the same equilibrated configurations
    snap = [eng.read_state(c)["occ"].copy() for c in range(nch)]
    res = {}
    per_T = {}
    cpu_line = None
    if not args.no_cpu:
        cpu = host_cpu()
        P = args.cpu_extra_passes
        jobs = [(sc.dims, snap[c], "canonical", temps[c], None, P, 5150 + c) for c in range(nch)]
        v, w = cpu_baseline_chains(flat, jobs, cpu["threads"])
        cpu_line = {"value": v, "unit": "site-attempts/s", "cores": cpu["threads"], "kind": "port",
                    "sample": f"plain proposal, all {nch} chains x {P} passes from the equilibrated states, {w:.1f} s"}
    for i, prop in enumerate(("plain", "learned", "informed")):
        eng.load_chains([sid] * nch, snap, temps, 91000 + 1000 * i + np.arange(nch))
        eng.set_proposal(prop, rounds_per_group=args.c4_rounds)
        eng.clear_samples()
        a0 = [eng.read_state(c) for c in range(nch)]
        t0 = time.perf_counter()
        eng.run_passes("canonical", args.c4_passes, 1, canonical_rounds=0)
        wall = time.perf_counter() - t0
        a1 = [eng.read_state(c) for c in range(nch)]
        E, _ = eng.read_samples()
        E = E / sc.n_cells
        burn = args.c4_passes // 10
        taus = np.array([tau_int(E[c, burn:]) for c in range(nch)])
        n = E.shape[1] - burn
        ess = float(np.sum(n / (2.0 * taus)))
        att = sum(b["n_attempt"] - a["n_attempt"] for a, b in zip(a0, a1))
        acc = sum(b["n_accept"] - a["n_accept"] for a, b in zip(a0, a1))
        m = E[:, burn:].mean(1).reshape(nT, nrep)
        per_T[prop] = (m.mean(1), m.std(1, ddof=1) / np.sqrt(nrep))
        # ESS over the post-burn-in samples per second of the passes that made them
        res[prop] = {"ess_per_s": ess / (wall * n / E.shape[1]),
                     "passes_per_s": args.c4_passes / wall, "attempts_per_s": att / wall,
                     "acceptance": acc / max(1, att), "tau_int_E_median_passes": float(np.median(taus)),
                     "tau_int_E_max_passes": float(np.max(taus)),
                     "sampled_passes_over_max_tau": float(n / np.max(taus))}
        # ESS/s by temperature band (fractions of T_c)
        bands = {}
        for lo, hi in ((0.0, 0.9), (0.9, 1.0), (1.0, 1.1), (1.1, 9.9)):
            sel = (temps >= lo * Tc) & (temps < hi * Tc)
            if sel.any():
                bands[f"{lo:.1f}-{hi:.1f}Tc"] = float(np.sum(n / (2.0 * taus[sel])) / (wall * n / E.shape[1]))
        res[prop]["ess_per_s_by_T_band"] = bands
    dof = 2 * (nrep - 1)
    t_crit = float(stats.t.ppf(1.0 - 0.01 / (2 * nT), dof))
    for prop in ("learned", "informed"):
        (m0, s0), (m1, s1) = per_T["plain"], per_T[prop]
        t = np.abs(m1 - m0) / np.sqrt(s0 ** 2 + s1 ** 2 + 1e-30)
        res[prop]["E_agreement"] = {
            "temperatures": nT, "replicas_per_temperature": nrep, "error": "standard error over replicas",
            "welch_t_max": float(t.max()), "t_critical_1pct_familywise": t_crit,
            "temperatures_agreeing": int(np.sum(t <= t_crit)),
            "frac_within_3_sigma": float(np.mean(t <= 3.0))}
        res[prop]["ess_per_s_vs_plain"] = res[prop]["ess_per_s"] / res["plain"]["ess_per_s"]
    eng.close()
    return {"workload": f"fcc_ordering_{L}x{L}x{L}_canonical_xB0.25_{nch}chains_{nT}T_x_{nrep}replicas_"
                        f"T{tgrid[0]:.0f}-{tgrid[-1]:.0f}K",
            "Tc_K": Tc, "sampled_passes": args.c4_passes, "equilibration_passes": args.c4_equil,
            "rounds_per_group": args.c4_rounds,
            "training": {"samples": args.c4_train_samples * nch, "adam_steps": args.c4_train_steps,
                         "seconds": train_s, "loss_first": float(loss[0]), "loss_last": float(loss[-1])},
            "cpu_baseline_plain": cpu_line,
            "mean_E_per_cell_by_T": {p: [float(x) for x in per_T[p][0]] for p in per_T},
            "err_E_per_cell_by_T": {p: [float(x) for x in per_T[p][1]] for p in per_T},
            **res}


def _run_timed(eng, kind, n_passes, n_chains):
    a0 = sum(eng.read_state(c)["n_attempt"] for c in range(n_chains))
    t0 = time.perf_counter()
    kernel_ms = eng.run_passes(kind, n_passes, 1, canonical_rounds=2)
    wall = time.perf_counter() - t0
    a1 = sum(eng.read_state(c)["n_attempt"] for c in range(n_chains))
    return (a1 - a0) / wall, wall, kernel_ms


def workload_c1(args):
    """The reference's CPU-runnable case: FCC 12^3, canonical, x_B = 0.5,
    T = 600 K, one chain, 1,000 + 5,000 passes, sampled every pass; the CPU
    oracle (one core) timed on the first 1,000 passes of the same chain."""
    from casmcode_clexmonte_amd.engine import Engine
    from casmcode_clexmonte_amd.system import FlatCE, Supercell
    from casmcode_clexmonte_amd.workloads import fcc_binary_ce, random_occupation
    ce = fcc_binary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (12, 12, 12))
    occ = random_occupation(sc.n_sites, 0.5, 2024)
    eng = Engine(flat, device=0)
    sid = eng.add_supercell(sc, n_pair_phases=16, seed=3)
    eng.load_chains([sid], [occ], [600.0], [11])
    t0 = time.perf_counter()
    eng.run_passes("canonical", 1000, 1, canonical_rounds=2)
    rate, wall, kms = _run_timed(eng, "canonical", 5000, 1)
    total = time.perf_counter() - t0
    E, _ = eng.read_samples()
    e = E[0, 1000:] / sc.n_cells
    tau = tau_int(e)
    eng.close()
    out = {"workload": "fcc_binary_12x12x12_canonical_xB0.5_T600_1chain", "site_attempts_per_s": rate,
           "passes_per_s": 5000 / wall, "run_s_1000_plus_5000_passes": total,
           "E_per_cell_mean": float(e.mean()), "E_per_cell_err": float(e.std() * np.sqrt(2 * tau / len(e))),
           "tau_int_E_passes": tau}
    if not args.no_cpu:
        O = _oracle()
        osys = O.OracleSystem(flat, sc.dims)
        t0 = time.perf_counter()
        r = osys.run(occ, "canonical", 600.0, 0, 1000, 11, full_recompute_samplers=True)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": r["n_attempt"] / el, "unit": "site-attempts/s", "cores": 1, "kind": "port",
                               "sample": f"1,000 passes of the same chain, {el:.1f} s"}
        out["speedup_vs_cpu_1core"] = rate / out["cpu_baseline"]["value"]
    return out


def workload_c3(args):
    """Ternary HCP, 64 canonical chains on uneven supercells 16x16x10 ...
    48x48x32 (5,120 ... 147,456 sites): one persistent launch per supercell
    shape (kernels specialised to each shape), all on concurrent streams."""
    from casmcode_clexmonte_amd.engine import Engine
    from casmcode_clexmonte_amd.system import FlatCE, Supercell
    from casmcode_clexmonte_amd.workloads import hcp_canonical_batch, hcp_ternary_ce
    ce = hcp_ternary_ce()
    flat = FlatCE(ce)
    b = hcp_canonical_batch(ce.prim)
    eng = Engine(flat, device=0)
    ids = {}
    for d in b.dims:
        if d not in ids:
            ids[d] = eng.add_supercell(Supercell(ce.prim, d), n_pair_phases=8, seed=len(ids))
    n = len(b.dims)
    eng.load_chains([ids[d] for d in b.dims], b.occ, b.temperatures, b.seeds)
    # one launch per supercell shape, each specialised and sized for its shape
    eng.run_passes("canonical", 2, 1, canonical_rounds=2)
    info = eng.last_run_info()
    rate, wall, kms = _run_timed(eng, "canonical", args.extra_passes, n)
    info["slabs_per_chain_max"] = eng.last_run_slabs()
    drift = eng.reanchor_energies()
    sites = sum(len(o) for o in b.occ)
    eng.close()
    out = {"workload": "hcp_ternary_canonical_64chains_16x16x10-48x48x32", "site_attempts_per_s": rate,
           "sites_total": sites, "passes": args.extra_passes, "wall_s": wall, "kernel_ms": kms,
           "launch": info, "max_energy_drift_eV": drift}
    if not args.no_cpu:
        cpu = host_cpu()
        P = args.cpu_extra_passes
        jobs = [(tuple(d), b.occ[i], "canonical", b.temperatures[i], None, P, b.seeds[i]) for i, d in enumerate(b.dims)]
        v, w = cpu_baseline_chains(flat, jobs, cpu["threads"])
        out["cpu_baseline"] = {"value": v, "unit": "site-attempts/s", "cores": cpu["threads"], "kind": "port",
                               "sample": f"all 64 chains x {P} passes, one chain per host thread task, {w:.1f} s"}
        out["speedup_vs_cpu"] = rate / v
    return out


def workload_c5(args):
    """Quaternary BCC 64^3 (262,144 sites), 32 canonical equiatomic chains,
    T = 300..2000 K, autotuned (2-bit packed LDS image)."""
    from casmcode_clexmonte_amd.engine import Engine
    from casmcode_clexmonte_amd.system import FlatCE, Supercell
    from casmcode_clexmonte_amd.workloads import bcc_equiatomic_batch, bcc_quaternary_ce
    ce = bcc_quaternary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (64, 64, 64))
    b = bcc_equiatomic_batch(sc)
    eng = Engine(flat, device=0)
    sid = eng.add_supercell(sc, n_pair_phases=8, seed=5)
    n = len(b.occ)
    eng.load_chains([sid] * n, b.occ, b.temperatures, b.seeds)
    tune = eng.autotune("canonical", trial_passes=3)
    rate, wall, kms = _run_timed(eng, "canonical", args.extra_passes, n)
    tune["slabs_per_chain"] = eng.last_run_slabs()
    drift = eng.reanchor_energies()
    eng.close()
    out = {"workload": "bcc_quaternary_64x64x64_canonical_equiatomic_32chains_T300-2000",
           "site_attempts_per_s": rate, "passes": args.extra_passes, "wall_s": wall, "kernel_ms": kms,
           "autotune": tune, "max_energy_drift_eV": drift}
    if not args.no_cpu:
        cpu = host_cpu()
        P = max(1, args.cpu_extra_passes // 4)
        jobs = [(sc.dims, b.occ[i], "canonical", b.temperatures[i], None, P, b.seeds[i]) for i in range(n)]
        v, w = cpu_baseline_chains(flat, jobs, cpu["threads"])
        out["cpu_baseline"] = {"value": v, "unit": "site-attempts/s", "cores": cpu["threads"], "kind": "port",
                               "sample": f"all 32 chains x {P} passes, one chain per host thread task, {w:.1f} s"}
        out["speedup_vs_cpu"] = rate / v
    return out


def workload_nfold(args):
    """N-fold way on the FCC binary 24^3 CE, 256 semi-grand canonical chains at
    low temperature (T = 200..400 K x mu = -0.3..0.3 eV, where Metropolis
    acceptance is small).  Every N-fold event is an accepted move; the plain
    kernel's accepted moves per second at the same points (its site-attempt
    rate times its acceptance) are measured beside it."""
    from casmcode_clexmonte_amd.engine import Engine
    from casmcode_clexmonte_amd.system import FlatCE, Supercell
    from casmcode_clexmonte_amd.workloads import fcc_binary_ce, fcc_sgc_grid
    ce = fcc_binary_ce()
    flat = FlatCE(ce)
    sc = Supercell(ce.prim, (24, 24, 24))
    batch = fcc_sgc_grid(sc, T_range=(200.0, 400.0))
    n = len(batch.temperatures)
    eng = Engine(flat, device=0)
    sid = eng.add_supercell(sc, n_pair_phases=0)
    eng.load_chains([sid] * n, batch.occ, batch.temperatures, batch.seeds, chem=batch.chem)
    eng.run_passes("semigrand_canonical", 200, 200)          # equilibrate with the plain kernel
    a0 = [eng.read_state(c) for c in range(n)]
    t0 = time.perf_counter()
    eng.run_passes("semigrand_canonical", 50, 50)
    plain_wall = time.perf_counter() - t0
    a1 = [eng.read_state(c) for c in range(n)]
    n_acc = sum(b["n_accept"] - a["n_accept"] for a, b in zip(a0, a1))
    n_att = sum(b["n_attempt"] - a["n_attempt"] for a, b in zip(a0, a1))
    n_events = args.nfold_events
    if args.nfold_generic:
        eng.set_specialization(False)
    eng.nfold_run(100, 100)                                    # warm up
    t0 = time.perf_counter()
    eng.nfold_run(n_events, n_events // 10)
    wall = time.perf_counter() - t0
    specialized = eng.last_run_info()["specialized"]
    eng.close()
    ev_rate = n * n_events / wall
    plain_acc_rate = n_acc / plain_wall
    return {"workload": "fcc_binary_24x24x24_sgc_256chains_T200-400_nfold", "events": n * n_events,
            "events_per_s": ev_rate, "wall_s": wall, "specialized": specialized,
            "plain_site_attempts_per_s": n_att / plain_wall,
            "plain_acceptance": n_acc / n_att, "plain_accepted_moves_per_s": plain_acc_rate,
            "nfold_over_plain_accepted_moves": ev_rate / plain_acc_rate}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-chains", type=int, default=0, help="chains of the CPU baseline (0: all 256)")
    ap.add_argument("--cpu-passes", type=int, default=40)
    ap.add_argument("--cpu-extra-passes", type=int, default=4,
                    help="passes per chain of the CPU baselines of the HCP, ordering and BCC workloads")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal ESS comparison")
    ap.add_argument("--no-extra", action="store_true", help="skip the single-chain, HCP and BCC workloads")
    ap.add_argument("--extra-passes", type=int, default=200)
    ap.add_argument("--nfold-events", type=int, default=20000)
    ap.add_argument("--nfold-generic", action="store_true",
                    help="run the N-fold workload on the generic (interpreted) kernel")
    ap.add_argument("--c4-size", type=int, default=32)
    ap.add_argument("--c4-temps", type=int, default=16)
    ap.add_argument("--c4-replicas", type=int, default=8)
    ap.add_argument("--c4-equil", type=int, default=2000)
    ap.add_argument("--c4-passes", type=int, default=12000)
    ap.add_argument("--c4-rounds", type=int, default=16)
    ap.add_argument("--c4-train-samples", type=int, default=512)
    ap.add_argument("--c4-train-steps", type=int, default=200)
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="GPU_MAX_HW_QUEUES for this process, at most 8 (set before the first HIP call; "
                         "0, the default, keeps the environment's value, the runtime's 4 unless set): the "
                         "uneven-shape batch runs one stream per hardware queue, its shape groups balanced "
                         "over them")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="copy N MiB device-to-device (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.hw_queues > 0:
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(args.hw_queues, 8))
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
    # attempt counters before the warm-up, so that the timed launch follows
    # the warm-up launch directly (the device does not idle in between)
    att0 = sum(eng.read_state(c)["n_attempt"] for c in range(n_chains))
    if args.warmup > 0:
        eng.run_passes("semigrand_canonical", args.warmup, 1)
    # clexmc_run_passes returns after the launch has completed on the device
    # (event synchronisation on the handle's stream), so the wall clock below
    # brackets exactly the K timed passes.
    t0 = time.perf_counter()
    kernel_ms = eng.run_passes("semigrand_canonical", args.steps, 1)
    wall = time.perf_counter() - t0
    att1 = sum(eng.read_state(c)["n_attempt"] for c in range(n_chains))
    attempts = att1 - att0 - n_chains * sc.n_sites * max(0, args.warmup)
    expected = n_chains * sc.n_sites * args.steps
    assert attempts == expected, (attempts, expected)
    value = attempts / wall
    # effective sample size of E per chain from the per-pass samples (only
    # for runs long enough to estimate tau_int: at least 500 passes)
    ess_total, tau_med = None, None
    if args.steps >= 500:
        E, cnt = eng.read_samples()
        E = E[:, E.shape[1] - args.steps:]  # the timed passes (the warm-up's samples come first)
        taus = np.array([tau_int(E[c]) for c in range(n_chains)])
        ess_total = float(np.sum(E.shape[1] / (2.0 * taus)))
        tau_med = float(np.median(taus))
    bytes_per_attempt = 1 + 54  # the site and its 54 interaction neighbours, int8 in LDS
    achieved = attempts * bytes_per_attempt / (kernel_ms * 1e-3) / 1e9
    # measured LDS read throughput of this device: byte gathers and 16-byte reads
    lds_u8 = eng.lds_probe("u8")
    lds_b128 = eng.lds_probe("b128")
    launch = eng.last_run_info()
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
                     "frac": achieved / LDS_PEAK_GBS,
                     "traffic": HBM_BYTES_PER_LAUNCH if launch.get("specialized") else None,
                     "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, calibrated)",
                     "bytes_per_attempt": bytes_per_attempt, "kernel_ms": kernel_ms,
                     "measured_lds_byte_gather_GBs": lds_u8, "frac_of_measured_byte_gather": achieved / lds_u8,
                     "measured_lds_16B_read_GBs": lds_b128,
                     # the roof that comes closest: VALU issue (profiled lane-ops per attempt x live rate)
                     "valu": {"lane_ops_per_attempt": VALU_OPS_PER_ATTEMPT,
                              "achieved_lane_ops_per_s": attempts * VALU_OPS_PER_ATTEMPT / (kernel_ms * 1e-3),
                              "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
                              "frac": attempts * VALU_OPS_PER_ATTEMPT / (kernel_ms * 1e-3) / VALU_PEAK_LANE_OPS}},
        "launch": dict(launch, gpu_max_hw_queues=os.environ.get("GPU_MAX_HW_QUEUES", "runtime default")),
        "ess_per_s": ess_total / wall if ess_total is not None else None,
        "tau_int_E_median_passes": tau_med,
        "acceptance": float(sum(eng.read_state(c)["n_accept"] for c in range(n_chains)) / max(1, att1)),
    }
    eng.close()
    if not args.no_learned:
        out["learned_vs_plain"] = learned_comparison(args)
    if not args.no_extra:
        out["workloads"] = {"single_chain": workload_c1(args), "hcp_ternary_uneven": workload_c3(args),
                            "bcc_quaternary_64": workload_c5(args), "nfold_low_T": workload_nfold(args)}
    if not args.no_cpu:
        cpu = host_cpu()
        threads = cpu["threads"]
        n_cpu = min(args.cpu_chains, n_chains) if args.cpu_chains > 0 else n_chains
        cpu_rate, r = cpu_baseline(flat, sc, batch, n_cpu, args.cpu_passes, threads)
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "site-attempts/s", "cores": threads, "kind": "port",
                               "sample": f"{n_cpu} chains x {args.cpu_passes} passes of the same "
                                         f"workload, oracle restatement of occupation_metropolis_v2 "
                                         f"(FP64, mt19937_64, per-pass full-recompute samplers), chains "
                                         f"spread over {threads} threads, {r['elapsed_s']:.1f} s wall"}
        out["cpu_host"] = cpu
        out["speedup_vs_cpu"] = value / cpu_rate
    print(json.dumps(out))


if __name__ == "__main__":
    main()
