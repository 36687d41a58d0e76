/* FP64 CPU oracle for the clexmonte Metropolis hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * (or the timed CPU baseline), never as the thing measured or shipped.
 *
 * It restates, in plain C, what the reference computes on this path:
 *  - correlations / formation energy: the generated Clexulator's
 *    eval_bfunc_* summed over the supercell
 *    (tests/unit/clexmonte/data/FCC_binary_vacancy/basis_sets/bset.default/
 *     FCC_binary_vacancy_Clexulator_default.cc:805-832) contracted with ECIs
 *    (ClusterExpansion::per_supercell, call site
 *     src/casm/clexmonte/monte_calculator/CanonicalCalculator.cc:126-128);
 *  - occ_delta_value: sequential single-site changes of every cluster that
 *    contains the changed site (FCC...Clexulator_default.cc:582-610,923-930;
 *    call site CanonicalCalculator.cc:137-140);
 *  - the semi-grand-canonical potential dE - mu.(R^T dN)
 *    (src/casm/clexmonte/monte_calculator/SemiGrandCanonicalCalculator.cc:186-213);
 *  - occupation_metropolis_v2 (include/casm/clexmonte/methods/
 *    occupation_metropolis.hh:72-123) with canonical swap / SGC flip
 *    proposals (CanonicalCalculator.cc:78-88,
 *    SemiGrandCanonicalCalculator.cc:104-120), std::mt19937_64 engine
 *    (include/casm/clexmonte/definitions.hh:35) and the reference's default
 *    per-pass samplers (CanonicalCalculator.cc:245-246,
 *    sampling_functions.cc:37-87,141-151,274-285), each recomputed in full
 *    as the reference does;
 *  - exact enumeration of the Boltzmann distribution on tiny supercells;
 *  - Philox4x32-10 (Salmon et al., SC'11) and mt19937_64 (Matsumoto &
 *    Nishimura 1998) used to pin the GPU generator and the CPU engine.
 */
#ifndef CLEXMC_ORACLE_H
#define CLEXMC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int n_sublat;
  const int *n_occ;           /* [n_sublat] */
  const int *asym;            /* [n_sublat] asymmetric-unit index (candidate lists) */
  int n_orbits;
  const int *orbit_k;         /* sites per cluster */
  const int *orbit_mult;      /* clusters per unit cell */
  const int *orbit_nfunc;
  const int *orbit_func_begin;/* linear function index of first function */
  const int *corr_orbit_begin;/* [n_orbits+1] element prefix */
  const int *corr_elem_begin; /* [n_orbits+1] site prefix into corr_elem_sites */
  const int *corr_elem_sites; /* [.][4] (b, i, j, k) */
  const int *orbit_table_offset;
  const double *orbit_tables; /* CT[tau][func_local], tau mixed radix site0 fastest */
  int n_corr;
  const double *eci;          /* [n_corr] */
  int L[3];                   /* diagonal supercell */
} orc_system;

/* correlations per supercell (sum over unit cells of the per-cell values) */
void orc_corr(const orc_system *s, const int8_t *occ, double *corr_sc);
double orc_energy(const orc_system *s, const int8_t *occ);
/* change of per-supercell energy for sequential site changes */
double orc_delta(const orc_system *s, int8_t *occ, int n, const int *sites, const int *new_occ);
/* same via the prepared neighbour-list path used by orc_run */
double orc_delta_prepared(const orc_system *s, int8_t *occ, int n, const int *sites, const int *new_occ);
/* same, for the correlations (length n_corr) */
void orc_delta_corr(const orc_system *s, int8_t *occ, int n, const int *sites, const int *new_occ,
                    double *dcorr);

/* Metropolis run (reference restatement).
 * kind 0: canonical swaps (uniform over swap types, then uniform candidate
 *         sites of each species on one sublattice class);
 * kind 1: semi-grand canonical single-site changes (uniform site, uniform
 *         other allowed occupant).
 * chem[b*n_occ_max + occ]: per-site-occupant potential weight, so that
 *   dOmega = dE - sum_changes (chem[new] - chem[old]).
 * Samples every pass after `n_equil` passes: E per supercell (FP64, full
 * recompute as the reference's samplers do) and occupant counts.
 * Returns accepted count; writes samples_E[n_samples], samples_counts
 * [n_samples][n_occ_max] if non-null. */
typedef struct {
  int kind;
  double temperature;
  double kb;
  const double *chem; /* [n_sublat * n_occ_max] or NULL */
  int n_occ_max;
  long long n_equil_passes;
  long long n_passes;        /* sampled passes */
  int full_recompute_samplers; /* 1: reference-like per-pass E recompute x2 */
  uint64_t seed;
} orc_run_params;

typedef struct {
  long long n_attempt;
  long long n_accept;
  double E_final;       /* running energy (FP64) */
  double elapsed_s;
} orc_run_result;

int orc_run(const orc_system *s, int8_t *occ, const orc_run_params *p, double *samples_E,
            int *samples_counts, orc_run_result *res);

/* multi-chain batch over threads: chains independent, occ [n_chains][N] */
int orc_run_batch(const orc_system *s, int n_chains, int8_t *occ, const double *temperature,
                  const double *chem /* [n_chains][n_sublat*n_occ_max] */,
                  const orc_run_params *p, const uint64_t *seeds, int n_threads,
                  long long *n_attempt_total, double *elapsed_s);

/* Exact enumeration over all occupations (kind 1) or all occupations with
 * fixed counts per occupant index (kind 0, counts in fixed_counts[n_occ_max]).
 * Outputs Boltzmann averages of E, E^2, count of occupant index 1..: avg[0]=<E>,
 * avg[1]=<E^2>, avg[2]=<n1>, avg[3]=<n1^2>, avg[4]=<E n1>, avg[5]=Z (shifted).
 * Also the full distribution if probs != NULL (length = number of states). */
int orc_enumerate(const orc_system *s, int kind, double temperature, double kb,
                  const double *chem, int n_occ_max, const int *fixed_counts, double *avg,
                  double *probs, long long n_probs);

/* Energy of every state in orc_enumerate's order (site 0 fastest, radix
 * n_occ of its sublattice); with fixed_counts != NULL, states whose
 * variable-site occupant counts differ get NaN.  n_states must equal the
 * number of states; only states [st0, st1) are written (E_out indexed by
 * state).  Returns 0, or -1 (too many states) / -2 (size or range). */
int orc_enumerate_energies(const orc_system *s, int n_occ_max, const int *fixed_counts, double *E_out,
                           long long n_states, long long st0, long long st1);

/* Philox4x32-10 */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* mt19937_64: returns the n-th output (1-based) for seed */
uint64_t orc_mt19937_64_nth(uint64_t seed, long long n);

#ifdef __cplusplus
}
#endif
#endif
