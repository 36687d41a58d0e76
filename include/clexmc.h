/* clexmc: C ABI of the MI355X (gfx950) Metropolis hot path for cluster-expansion
 * Monte Carlo.
 *
 * The entry points replace the per-step arithmetic the reference calls through
 * C++ objects.  Each declaration names the reference interface it stands in
 * for (paths relative to the reference repository root):
 *
 *  - clexmc_create / clexmc_add_supercell: the System + SupercellSystemData
 *    construction that builds clusters, neighbour lists and ECIs
 *    (src/casm/clexmonte/system/System.cc:41-72, get_clex :633-641,
 *    StateData::StateData src/casm/clexmonte/monte_calculator/StateData.cc:9-79);
 *  - clexmc_load_chains: MonteCalculator::set_state_and_potential + the
 *    OccLocation initialisation of one or many states
 *    (include/casm/clexmonte/monte_calculator/MonteCalculator.hh:127-135,161-186);
 *  - clexmc_delta_energy_batch: BaseMontePotential::occ_delta_per_supercell
 *    (include/casm/clexmonte/monte_calculator/BaseMonteCalculator.hh:38-43;
 *    CanonicalCalculator.cc:137-140, SemiGrandCanonicalCalculator.cc:186-213);
 *  - clexmc_correlations / clexmc_energy: Correlations::per_supercell and
 *    ClusterExpansion::per_supercell used by the "corr.<key>" and
 *    "clex.<key>" samplers (src/casm/clexmonte/monte_calculator/
 *    sampling_functions.cc:121-151);
 *  - clexmc_run_passes: the Metropolis main loop occupation_metropolis_v2
 *    (include/casm/clexmonte/methods/occupation_metropolis.hh:72-123) driven
 *    by CanonicalCalculator::run (CanonicalCalculator.cc:413-462) and
 *    SemiGrandCanonicalCalculator::run (SemiGrandCanonicalCalculator.cc:423-466),
 *    and, for many chains at once, the experimental multi-state run
 *    (BaseMonteCalculator.hh:245-247);
 *  - clexmc_read_samples / clexmc_read_state: the RunManager sample storage
 *    read by the analysis functions (analysis_functions.cc:17-173).
 *
 * Conventions: plain C types only; no exceptions cross this boundary; every
 * call returns a clexmc_status and a message is kept for
 * clexmc_last_error().  A handle owns its device memory and one HIP stream;
 * it is thread-compatible (one handle per host thread), not thread-safe.
 * Linear site index l = b * n_cells + i + L1 * (j + L2 * k).
 */
#ifndef CLEXMC_H
#define CLEXMC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  CLEXMC_OK = 0,
  CLEXMC_ERR_ARG = 1,
  CLEXMC_ERR_HIP = 2,
  CLEXMC_ERR_UNSUPPORTED = 3,
  CLEXMC_ERR_STATE = 4,
  CLEXMC_ERR_NO_DEVICE = 5
} clexmc_status;

typedef struct clexmc_handle clexmc_handle;

/* Flattened cluster expansion.  Produced by the host system builder
 * (casmcode_clexmonte_amd/system.py FlatCE); all arrays are copied. */
typedef struct {
  int n_sublat;
  const int *n_occ;            /* [n_sublat] */
  int n_occ_max;               /* <= 4 */
  int n_func_max;              /* site functions per site, <= 3 */
  const double *site_func;     /* [n_sublat][n_func_max][n_occ_max] */
  /* correlations: orbit element lists and orbit tables */
  int n_orbits;
  const int *orbit_k;
  const int *orbit_mult;
  const int *orbit_nfunc;
  const int *orbit_func_begin;
  const int *corr_orbit_begin;   /* [n_orbits+1] */
  const int *corr_elem_begin;    /* [n_orbits+1] */
  const int *corr_elem_sites;    /* [n_elem_sites][4] b,i,j,k */
  int n_corr_elem_sites;
  const int *orbit_table_offset; /* [n_orbits] */
  const double *orbit_tables;
  int n_orbit_tables;            /* doubles */
  int n_corr;
  const double *eci;             /* [n_corr] */
  /* flower: cluster instances containing a site of sublattice b */
  const int *flower_begin;       /* [n_sublat+1] */
  int n_entries;
  const int *entry_orbit;        /* [n_entries] */
  const int *entry_k;            /* other sites per entry */
  const int *entry_site_begin;   /* [n_entries+1] */
  const int *entry_sites;        /* [n_other_sites][4] b,di,dj,dk */
  int n_entry_sites;
  const int *entry_table;        /* [n_entries] offset (values) into flower_tables */
  const double *flower_tables;   /* V[tau_others][f] */
  int n_flower_tables;
  /* binary (+-1 Chebyshev) integer path; binary == 0 disables it */
  int binary;
  const int *bin_seg_begin;      /* [n_sublat+1] */
  const int *bin_segs;           /* [n_segs][3]: orbit, entry_begin, entry_end */
  int n_bin_segs;
  const double *bin_orbit_weight;/* [n_orbits] J_a / m_a */
  /* asymmetric unit of each sublattice ([n_sublat], may be NULL): canonical
   * swaps exchange occupants only between sites of one unit, as the
   * reference's OccSwap candidates (monte/events/OccCandidate.hh) are keyed
   * by (asym unit, species).  NULL: sublattices with equal occupant counts
   * share a unit. */
  const int *asym;
} clexmc_system_desc;

/* Supercell (diagonal, L1 x L2 x L3 primitive cells) with its independent-set
 * partition: sites of one class share no cluster.  With n_classes == 0 the
 * library computes the partition itself (periodic colouring by residues of
 * the cell coordinates, or a greedy colouring) and n_pair_phases canonical
 * mixing phases; class/pair arrays are then ignored. */
typedef struct {
  int L[3];
  int n_classes;
  const int *class_begin;   /* [n_classes+1] */
  const int *class_sites;   /* linear site indices */
  /* canonical cross-class mixing: phases of disjoint, mutually
   * non-interacting site pairs (optional; n_pair_phases may be 0); the two
   * sites of a pair may lie on different sublattices of one asymmetric unit
   * (occupant indices must denote the same species) */
  int n_pair_phases;
  const int *pair_phase_begin; /* [n_pair_phases+1] in pairs */
  const int *pair_sites;       /* [n_pairs][2] */
} clexmc_supercell_desc;

/* Creates a handle on HIP device `device`.  Replaces System construction
 * (System.cc:41-72).  Fails with CLEXMC_ERR_NO_DEVICE without a GPU. */
clexmc_status clexmc_create(const clexmc_system_desc *sys, int device, clexmc_handle **out);

/* Registers a supercell; returns its id.  (SupercellSystemData, System.cc:62-72) */
clexmc_status clexmc_add_supercell(clexmc_handle *h, const clexmc_supercell_desc *sc, int *id_out);

/* Loads n_chains states (replaces set_state_and_potential + OccLocation
 * initialisation, MonteCalculator.hh:127-135,161-186).
 * occ: concatenated per-chain occupations (int8 occupant index per site).
 * temperature [n_chains] (K); chem [n_chains][n_sublat*n_occ_max] per
 * site-occupant potential weights such that dOmega = dE - (chem[new]-chem[old])
 * (SemiGrandCanonicalCalculator.cc:186-213 with exchange potentials folded
 * in), or NULL; seeds [n_chains] Philox keys.  kb: Boltzmann constant (eV/K). */
clexmc_status clexmc_load_chains(clexmc_handle *h, int n_chains, const int *supercell_id,
                                 const int8_t *occ, const double *temperature, const double *chem,
                                 const uint64_t *seeds, double kb);

/* Per-proposal energy change (per supercell, eV) evaluated on the GPU in FP32
 * with sequential application of the sites of a proposal (a swap is two
 * changes, the second seeing the first), at each chain's current occupation.
 * Replaces BaseMontePotential::occ_delta_per_supercell
 * (BaseMonteCalculator.hh:38-43; ClusterExpansion::occ_delta_value call site
 * CanonicalCalculator.cc:137-140).  Host arrays: chain[n_prop],
 * sites[n_prop][sites_per_prop], new_occ[n_prop][sites_per_prop];
 * sites_per_prop <= 4.  If include_chem != 0 the chain's chem weights are
 * subtracted (semi-grand canonical potential). */
clexmc_status clexmc_delta_energy_batch(clexmc_handle *h, int n_prop, const int32_t *chain,
                                        const int32_t *sites, const int32_t *new_occ,
                                        int sites_per_prop, int include_chem, float *dE_out);

/* clexmc_delta_energy_batch without any allocation: no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops that call
 * occ_delta_per_supercell per proposal, harnesses that forbid hipMalloc).
 * Same arguments and results as clexmc_delta_energy_batch, but the proposals
 * and results are staged in a fixed array in device memory that the library
 * holds on each GPU for the life of the process (2.5 MiB), so at most 65,536
 * proposals per call (more is rejected with CLEXMC_ERR_ARG).  Every call on a
 * GPU uses that GPU's one array: calls that overlap on one GPU (from several
 * handles or threads at once) race and must not be made. */
clexmc_status clexmc_delta_energy_batch_noalloc(clexmc_handle *h, int n_prop, const int32_t *chain,
                                                const int32_t *sites, const int32_t *new_occ, int sites_per_prop,
                                                int include_chem, float *dE_out);

/* Full per-supercell correlations of one chain in FP64 (Correlations::
 * per_supercell, sampling_functions.cc:121-135).  corr_out[n_corr]. */
clexmc_status clexmc_correlations(clexmc_handle *h, int chain, double *corr_out);

/* Runs n_passes Metropolis passes on every loaded chain in one persistent
 * launch (occupation_metropolis.hh:72-123).  kind: 0 canonical swaps,
 * 1 semi-grand canonical changes.  A sample (formation energy per
 * supercell, occupant counts) is recorded every sample_period passes.
 * canonical_rounds: random matchings per independent set and pass
 * (canonical only; <= 4096).  0 (automatic): as many matchings as make
 * about one attempt between different occupants per mutable site and pass
 * -- the reference's steps per pass -- shared among the sets in proportion
 * to their pairs of different occupants (from the occupant counts per set
 * at the start of each pass; at most 1024 per set); a set holding one
 * occupant only is skipped. */
clexmc_status clexmc_run_passes(clexmc_handle *h, int kind, int n_passes, int sample_period,
                                int canonical_rounds);

/* Per-chain totals of the last run and the samples recorded so far.
 * E [n_chains][*n_samples] (formation energy per supercell),
 * counts [n_chains][*n_samples][n_occ_max]; pass NULL to query *n_samples. */
clexmc_status clexmc_read_samples(clexmc_handle *h, int *n_samples, double *E, int32_t *counts);
clexmc_status clexmc_clear_samples(clexmc_handle *h);

/* Current occupation (N sites of the chain's supercell), running formation
 * energy (FP64) and attempt / accept counters of one chain. */
clexmc_status clexmc_read_state(clexmc_handle *h, int chain, int8_t *occ, double *E_running,
                                long long *n_attempt, long long *n_accept);

/* Exact integer cluster-product sums C_a = m_a * corr_a of the binary path
 * (tracked incrementally, int64 [n_orbits]). */
clexmc_status clexmc_read_binary_counts(clexmc_handle *h, int chain, long long *counts_out);

/* HIP graph of one block of passes (the persistent Metropolis launch),
 * captured once and replayed: each replay continues every chain exactly as a
 * clexmc_run_passes call of passes_per_replay passes would.  Samples for up
 * to max_replays replays are preallocated at capture.  The graph is released
 * by clexmc_graph_release, clexmc_load_chains or clexmc_destroy.  (No
 * reference counterpart: the reference's run loop is host code,
 * occupation_metropolis.hh:72-123.) */
clexmc_status clexmc_graph_capture(clexmc_handle *h, int kind, int passes_per_replay, int max_replays,
                                   int sample_period, int canonical_rounds);
clexmc_status clexmc_graph_replay(clexmc_handle *h, int n_replays);
clexmc_status clexmc_graph_release(clexmc_handle *h);

/* Autotuner: times trial_passes-pass launches of the loaded chains over
 * block sizes (64..512), LDS image formats (int8 where it fits, 2-bit
 * packed) and, for binary expansions, 1 or 2 lanes per site; restores the
 * chains' device state after each trial, and sets the fastest as the
 * handle's block size, packing and lanes per site.  Cached per (kind,
 * proposal, CE shape, chain count, largest padded supercell, largest class).
 * Outputs may be NULL; ms_per_pass is -1 when the cache answered. */
clexmc_status clexmc_autotune(clexmc_handle *h, int kind, int trial_passes, int canonical_rounds, int *threads_out,
                              int *packing_out, int *lanes_out, double *ms_per_pass_out);

/* Measured LDS read throughput of the device (GB/s): mode 0 byte gathers
 * (the Metropolis kernels' access type), mode 1 bank-conflict-free 16-byte
 * reads.  A roofline reference, not part of a run. */
clexmc_status clexmc_lds_probe(clexmc_handle *h, int mode, double *gbps);

/* Replaces every chain's FP64 running energy (the sum of FP32 dEs on the
 * table path) by a full FP64 recompute of its correlations and returns the
 * largest |running - recomputed| (eV per supercell).  The binary path's
 * energies come from exact integer counts and show zero drift. */
clexmc_status clexmc_reanchor_energies(clexmc_handle *h, double *max_drift);

/* Kernel timing of the last clexmc_run_passes (HIP events on the handle's
 * stream, milliseconds). */
clexmc_status clexmc_last_run_ms(clexmc_handle *h, double *ms);

/* Tuning knobs (block size of the Metropolis kernel, a multiple of 64 up to
 * 1024; 0 = automatic: one lane per site of the largest class.  Blocks above
 * 512 threads run only on run-time specialised kernels; without one the
 * library caps the block at 512). */
clexmc_status clexmc_set_block_size(clexmc_handle *h, int threads);

/* Block size of the N-fold kernel (clexmc_nfold_run*): 64, 128 or 256
 * threads per chain, or 0 (default) for the smallest multiple of 64 that
 * covers a site's interaction-neighbour list, at most 256.  An event's
 * select and flip run in the first wavefront, its neighbours' rates one per
 * thread; more wavefronts share only the launch's image load and initial
 * rates.  Results do not depend on it.  No counterpart in the reference
 * (its nfold loop, include/casm/clexmonte/nfold/nfold_impl.hh:18-98, runs
 * on one CPU thread). */
clexmc_status clexmc_set_nfold_block_size(clexmc_handle *h, int threads);

/* Slab decomposition of large chains (no reference counterpart: the
 * reference runs one chain on one thread, occupation_metropolis.hh:72-123).
 * A chain split into S slabs along the third cell axis runs on S workgroups
 * at once; workgroup s updates only the sites of its slab's planes except a
 * frozen band, as thick as the clusters reach along that axis, at the slab's
 * top, so updated sites of different slabs share no cluster and every site a
 * workgroup reads is its own or frozen.  The slab boundaries move by a random
 * plane offset (from the chain's key and pass counter) every launch of at
 * most 4 passes; each slab records its additions at the sampled passes and a
 * follow-up kernel forms the samples and chain totals.  slabs: 0 = automatic
 * (large chains in batches that do not fill the device, about
 * total_sites / compute_units sites per workgroup), 1 = off, 2..64 = that
 * many slabs where each keeps an updated plane.  Plain Metropolis runs
 * (clexmc_run_passes) only; graph captures and the autotuner run whole chains.
 * Canonical runs under slabs skip the serial random-pair swaps. */
clexmc_status clexmc_set_slabs(clexmc_handle *h, int slabs);
/* Slabs per chain of the last clexmc_run_passes (the largest over its
 * launches; 1 = whole chains). */
clexmc_status clexmc_last_run_slabs(clexmc_handle *h, int *slabs);

/* Chains on several supercell shapes with run-time specialisation on
 * (plain proposal, one lane per site, automatic block size, at most 16
 * shapes): mode 1 (default) runs all of them in ONE launch of a kernel that
 * holds the specialised field of every shape, each workgroup dispatching on
 * its shape (image format, chain map, slabs), so all shapes run at once;
 * mode 0 runs one specialised launch per shape on concurrent streams, which
 * share the process's hardware queues.  Results depend on the block size
 * only: the fused launch uses the largest block of the shapes.  No
 * counterpart in the reference (its runs hold one configuration); this is
 * the batching of chains on uneven supercells. */
clexmc_status clexmc_set_fused_launch(clexmc_handle *h, int mode);
/* 1 when the last clexmc_run_passes ran as one fused launch, else 0. */
clexmc_status clexmc_last_run_fused(clexmc_handle *h, int *fused);

/* Register-cube sweep of the run-time specialised semi-grand canonical
 * binary kernel (one sublattice, classes = the residues of the cell
 * coordinates mod 3, one lane per 3x3x3 block of cells; automatic where the
 * supercell has that layout): each lane keeps the (3 + 2h)^3 cells around
 * its block in registers (one bit per cell: 11 registers for h = 2),
 * computes every site's field
 * from them and, after each class, reads back only the cells of that class
 * that other lanes may have changed, instead of one LDS read per neighbour
 * and a halo copy per class.  The proposals, random draws and results are
 * bit-identical to the class loop.  mode 1 (default): where available; 0:
 * the class loop.  No counterpart in the reference (a layout of the sweep). */
clexmc_status clexmc_set_register_cube(clexmc_handle *h, int mode);
/* 1 when the last clexmc_run_passes used the register-cube sweep, else 0. */
clexmc_status clexmc_last_run_cube(clexmc_handle *h, int *cube);

/* Canonical runs: serial swaps of uniformly random site pairs per pass
 * (default 1), beside the parallel independent-set matchings and pair
 * phases.  The pair may share clusters (the reference's unrestricted
 * OccSwap proposal, CanonicalCalculator.cc:78-83); these moves keep every
 * supercell ergodic, including ones too small for non-interacting pairs. */
clexmc_status clexmc_set_global_swaps(clexmc_handle *h, int per_pass);

/* Semi-grand canonical multiswaps: the reference proposes only multiswaps
 * when its system's single semi-grand canonical swaps are empty
 * (SemiGrandCanonicalCalculator.cc:104-115; the system's
 * "semigrand_canonical_multiswaps", System_json_io.cc:955-1006), for
 * example charge-neutral combinations of single-site changes.
 * n_types types (at most 256); type t has the parts
 * parts[type_begin[t] .. type_begin[t+1])[4] = (asymmetric unit, occupant
 * from, occupant to, count): 1..4 parts and at most 8 sites per type, no
 * part repeated; the list must hold, for every type, the type that undoes
 * it (the same parts with from and to exchanged).  per_pass: multiswap
 * attempts per pass (0: one per mutable site, the reference's pass).
 * With n_types > 0, semi-grand canonical runs (clexmc_run_passes kind 1,
 * graph captures) change composition only through multiswaps: each pass is
 * the fixed-composition canonical sweep (matchings, pair phases, random-pair
 * swaps) followed by per_pass serial attempts.  An attempt draws a type
 * uniformly and, per part, `count` distinct sites uniformly among the sites
 * of the unit holding `from`; it is accepted with
 * min(1, exp(-beta dOmega) q_reverse / q_forward), q the proposal
 * probability of the move and of the move that undoes it (exact
 * Metropolis-Hastings; dOmega with sequential site changes).  n_types == 0
 * restores single-site changes.  The N-fold method refuses multiswaps. */
clexmc_status clexmc_set_multiswaps(clexmc_handle *h, int n_types, const int *type_begin, const int *parts,
                                    int per_pass);

/* Row pitch of the padded LDS occupation image of supercells added after
 * this call: -1 (default) chooses, per supercell, the pitch L1 + 2h + x
 * (x <= max(4, (L1 + 2h)/4), image growth <= 25 %, images up to 64 KiB)
 * that minimises modelled LDS bank conflicts of the classes' neighbour
 * gathers; 0..64 fixes x.  The
 * extra bytes of each row are never read; results do not depend on the
 * pitch.  No reference counterpart (memory layout of this build). */
clexmc_status clexmc_set_lds_row_pitch(clexmc_handle *h, int extra);

/* N-fold way (rejection-free) semi-grand canonical runs of every loaded
 * chain (supercells up to 262,144 sites): the reference's nfold method
 * (include/casm/clexmonte/nfold/nfold_impl.hh:18-98) with its event rates
 * (src/casm/clexmonte/nfold/nfold_events.cc:28-62: 1 if the semi-grand
 * potential change dE <= 0, else exp(-beta dE)).  The events are the plain
 * semi-grand canonical proposal's (a site and a uniformly chosen other
 * occupant), so a site's rate is (1/(n_occ-1)) times the sum of its
 * changes' rates, and a chosen site takes a new occupant in proportion to
 * the change's rate.  Binary expansions keep exact cluster counts, others
 * an FP64 running energy of FP32 changes; with every chain on one
 * supercell the run-time specialised kernel runs when specialisation is on
 * (clexmc_set_specialization; identical results).  Each chain makes
 * n_events events; every state visited is weighted by its expected residence
 * time w = n_mutable / R (Metropolis steps; R the total rate), so weighted
 * averages are Boltzmann averages.  blocks_out[n_chains][n_events /
 * events_per_block][6] receives, per block of events, (sum w, sum w E,
 * sum w N1, sum w E^2, sum w N1^2, sum w E N1) with E the formation energy
 * per supercell and N1 the number of sites holding occupant 1.  Chain state (occupation,
 * exact counts, energy, counters) advances as in clexmc_run_passes. */
clexmc_status clexmc_nfold_run(clexmc_handle *h, int n_events, int events_per_block, double *blocks_out);
/* The same run with the moments of every occupant count (any number of
 * occupants, the calculator-level nfold method for ternary and larger
 * expansions): blocks_out[n_chains][n_events / events_per_block][21] receives
 * per block [0] sum w, [1] sum w E, [2] sum w E^2, [3 + o] sum w N_o
 * (o = 0..3), [7 + t] sum w N_o N_q for o <= q in row order (t = 0 for
 * (0,0), 1 (0,1), 2 (0,2), 3 (0,3), 4 (1,1), 5 (1,2), 6 (1,3), 7 (2,2),
 * 8 (2,3), 9 (3,3)), [17 + o] sum w E N_o; N_o the sites holding occupant
 * o (0 for occupants the system does not have). */
clexmc_status clexmc_nfold_run_moments(clexmc_handle *h, int n_events, int events_per_block, double *blocks_out);
/* N-fold way sampled by time (the reference's nfold method allows
 * time-based sampling, include/casm/clexmonte/nfold/nfold.hh:22-23): every
 * chain records n_samples samples into its sample storage (read with
 * clexmc_read_samples: the energy per supercell and the occupant counts) at
 * times 0, period, 2 period, ... since the call, each the state the chain
 * occupies at that time, and stops after its last sample.  Time advances by
 * each visited state's expected residence time 1/R in passes (expected
 * Metropolis steps per mutable site; R the total rate), so the samples are
 * Boltzmann-distributed states with equal weights.  max_events bounds the
 * events per chain; a chain that takes fewer samples (or whose rates all
 * vanish) makes the call fail.  Since every rate is at most 1, n_mutable x
 * (n_samples x period + 1) events always suffice. */
clexmc_status clexmc_nfold_run_time(clexmc_handle *h, int n_samples, double period, int max_events);

/* Run-time specialisation (hipRTC, gfx950) of the binary Metropolis kernel
 * to the loaded supercell: constant site offsets instead of interpreted site
 * programs; used when every chain shares one supercell.  1 (default): when
 * possible, 0: never.  Results are bit-identical either way. */
clexmc_status clexmc_set_specialization(clexmc_handle *h, int mode);

/* Compiler log of the last specialisation attempt (NUL-terminated, at most
 * cap bytes). */
clexmc_status clexmc_specialization_log(clexmc_handle *h, char *buf, int cap);

/* What the last clexmc_run_passes launched: specialised kernel (0/1), 2-bit
 * packed image (0/1), block size. */
clexmc_status clexmc_last_run_info(clexmc_handle *h, int *specialized, int *packed, int *threads);

/* Binary expansions: lanes per site in the class sweeps (2: the two lanes of
 * a pair each sum half of every orbit's clusters, doubling the wavefronts per
 * chain, at the cost of neighbour reads the two halves both need).  0
 * automatic (1, or what clexmc_autotune chose), 1, 2.  Results do not depend
 * on it except through the RNG stream assignment. */
clexmc_status clexmc_set_lanes_per_site(clexmc_handle *h, int lanes);
clexmc_status clexmc_last_run_lanes(clexmc_handle *h, int *lanes_per_site);

/* LDS image format of the Metropolis kernel: 0 automatic (int8 per site when
 * the largest padded supercell fits, else 2 bits per site), 1 int8, 2 packed
 * 2-bit (at most 4 occupants per site). */
clexmc_status clexmc_set_image_packing(clexmc_handle *h, int mode);

/* ---- learned proposal (no counterpart in the reference: it plugs a
 * proposal into the Metropolis loop occupation_metropolis_v2,
 * occupation_metropolis.hh:97-105, keeping Metropolis-Hastings acceptance
 * with the exact proposal ratio) ---- */

/* Proposal used by clexmc_run_passes: 0 plain (independent-set sweeps and
 * random matchings), 1 learned (MLP-weighted sequential proposals within
 * groups of 16 independent sites, four groups side by side per wavefront),
 * 2 the same grouped proposals with uniform
 * weights, 3 the same with informed weights eps + sigmoid(-beta e_s) of the
 * exact flip energy (no network).  rounds_per_group: proposals per group and
 * class phase.  Binary cluster expansions only. */
clexmc_status clexmc_set_proposal(clexmc_handle *h, int proposal, int rounds_per_group);

/* Network size: n_env inputs (sites of the three shortest pair orbits),
 * n_params = 64*48 + 64 + 64 + 1 FP32 parameters [W1 (64x48 row-major) | b1 |
 * W2 | b2]; env_sites [n_sublat][48][4] (b, di, dj, dk) if non-NULL. */
clexmc_status clexmc_mlp_info(clexmc_handle *h, int *n_env, int *n_params, int *env_sites);
clexmc_status clexmc_mlp_set_params(clexmc_handle *h, const float *params, float eps);
clexmc_status clexmc_mlp_get_params(clexmc_handle *h, float *params);

/* MFMA (BF16) inference on feature rows X [n][48] (+1/-1 on the first n_env
 * inputs, the rest ignored): out [n][2] =
 * network outputs z for the current and the flipped occupation (a chain at
 * inverse temperature beta uses the logit beta * 0.1 eV * z). */
clexmc_status clexmc_mlp_forward(clexmc_handle *h, int n, const int8_t *X, float *out);

/* FP32 training forward/backward on (X, y, t): mean binary cross-entropy of
 * sigmoid(t * z) against y and its gradient with respect to the n_params
 * parameters (t NULL: t = 1). */
clexmc_status clexmc_mlp_gradients(clexmc_handle *h, int n, const int8_t *X, const float *y, const float *t,
                                   float *grad_out, float *loss_out);

/* Training samples from the loaded chains' current configurations:
 * X [n_chains*n_per_chain][48], y = sigmoid(-beta * flip energy),
 * t = beta * 0.1 eV (t_out may be NULL). */
clexmc_status clexmc_mlp_collect(clexmc_handle *h, int n_per_chain, uint32_t seed, int include_chem, int8_t *X_out,
                                 float *y_out, float *t_out);

/* Collects n_per_chain samples per chain and runs n_steps Adam steps on the
 * device (hand-written forward/backward kernels); loss_out [n_steps] or NULL.
 * A network never set by clexmc_mlp_set_params nor trained before starts
 * from a seeded He initialisation. */
clexmc_status clexmc_mlp_train(clexmc_handle *h, int n_per_chain, int n_steps, float lr, uint32_t seed,
                               int include_chem, float *loss_out);

/* The grouped proposals' wavefront primitives on given data (known-answer
 * checks): per row of 64 weights v >= 0, the inclusive prefix sums and the
 * chosen lane for the uniform u[row] (first lane with prefix > u*total and
 * v > 0; -1 if all weights are 0). */
clexmc_status clexmc_wave_pick_device(clexmc_handle *h, int n_rows, const float *v /*[n_rows][64]*/,
                                      const float *u /*[n_rows]*/, float *scan_out /*[n_rows][64]*/,
                                      int *pick_out /*[n_rows]*/);

/* Philox4x32-10 evaluated on the device (known-answer checks). */
clexmc_status clexmc_philox_device(clexmc_handle *h, int n, const uint32_t *ctr /*[n][4]*/,
                                   const uint32_t *key /*[n][2]*/, uint32_t *out /*[n][4]*/);

const char *clexmc_last_error(clexmc_handle *h);
void clexmc_destroy(clexmc_handle *h);

/* Library identification (compile-time target and version). */
const char *clexmc_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CLEXMC_H */
