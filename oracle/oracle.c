/* FP64 CPU oracle for the clexmonte Metropolis hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product path.
 *
 * Each function cites the reference code it restates.  The arithmetic of
 * the cluster expansion itself lives in the non-vendored libcasm-clexulator
 * and in generated Clexulator sources; the restatement here follows the
 * generated pattern of the FCC_binary_vacancy fixture:
 *   per-supercell correlation of function phi of orbit a
 *     = (1/m_a) sum_{unit cells x} sum_{clusters c of a in cell x} Phi(sigma_c)
 *   (FCC_binary_vacancy_Clexulator_default.cc:825-832: "(...)/6." for the
 *    nearest-neighbour pair orbit of multiplicity 6);
 *   occ_delta_value: for each changed site, the change of every cluster
 *   function on the clusters containing that site, with earlier changes of
 *   the same event already applied (the Clexulator's
 *   _calc_restricted_delta_point_corr, :582-610).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* indexing: l = b * n_cells + i + L1 * (j + L2 * k), periodic          */
/* ------------------------------------------------------------------ */
static inline int pmod(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

static inline long site_index(const orc_system *s, int b, int i, int j, int k) {
  long nc = (long)s->L[0] * s->L[1] * s->L[2];
  return b * nc + pmod(i, s->L[0]) + (long)s->L[0] * (pmod(j, s->L[1]) + (long)s->L[1] * pmod(k, s->L[2]));
}

static inline long n_cells(const orc_system *s) { return (long)s->L[0] * s->L[1] * s->L[2]; }

/* ------------------------------------------------------------------ */
/* correlations: sampling_functions.cc:121-135 (corr.<key>) restated    */
/* ------------------------------------------------------------------ */
void orc_corr(const orc_system *s, const int8_t *occ, double *corr_sc) {
  memset(corr_sc, 0, sizeof(double) * s->n_corr);
  const int L1 = s->L[0], L2 = s->L[1], L3 = s->L[2];
  for (int a = 0; a < s->n_orbits; ++a) {
    const int k = s->orbit_k[a];
    const int nf = s->orbit_nfunc[a];
    const int fb = s->orbit_func_begin[a];
    const double *T = s->orbit_tables + s->orbit_table_offset[a];
    const int e0 = s->corr_orbit_begin[a], e1 = s->corr_orbit_begin[a + 1];
    double *acc = (double *)calloc(nf, sizeof(double));
    for (int e = e0; e < e1; ++e) {
      const int *es = s->corr_elem_sites + 4 * (s->corr_elem_begin[a] + (e - e0) * k);
      for (int kk = 0; kk < L3; ++kk)
        for (int jj = 0; jj < L2; ++jj)
          for (int ii = 0; ii < L1; ++ii) {
            long tau = 0, mul = 1;
            for (int q = 0; q < k; ++q) {
              const int b = es[4 * q];
              long l = site_index(s, b, ii + es[4 * q + 1], jj + es[4 * q + 2], kk + es[4 * q + 3]);
              tau += occ[l] * mul;
              mul *= s->n_occ[b];
            }
            const double *row = T + tau * nf;
            for (int f = 0; f < nf; ++f) acc[f] += row[f];
          }
    }
    for (int f = 0; f < nf; ++f) corr_sc[fb + f] = acc[f] / s->orbit_mult[a];
    free(acc);
  }
}

double orc_energy(const orc_system *s, const int8_t *occ) {
  double *c = (double *)malloc(sizeof(double) * s->n_corr);
  orc_corr(s, occ, c);
  double E = 0.0;
  for (int i = 0; i < s->n_corr; ++i) E += s->eci[i] * c[i];
  free(c);
  return E;
}

/* Change of all correlations (per supercell) for changing site l from its
 * current occupant to `nw`, at the current occupation. */
static void delta_corr_site(const orc_system *s, const int8_t *occ, long l, int nw, double *dcorr) {
  const long nc = n_cells(s);
  const int b0 = (int)(l / nc);
  long c = l % nc;
  const int i0 = (int)(c % s->L[0]);
  const int j0 = (int)((c / s->L[0]) % s->L[1]);
  const int k0 = (int)(c / ((long)s->L[0] * s->L[1]));
  const int old = occ[l];
  if (old == nw) return;
  for (int a = 0; a < s->n_orbits; ++a) {
    const int k = s->orbit_k[a];
    if (k == 0) continue;
    const int nf = s->orbit_nfunc[a];
    const int fb = s->orbit_func_begin[a];
    const double inv_m = 1.0 / s->orbit_mult[a];
    const double *T = s->orbit_tables + s->orbit_table_offset[a];
    const int e0 = s->corr_orbit_begin[a], e1 = s->corr_orbit_begin[a + 1];
    for (int e = e0; e < e1; ++e) {
      const int *es = s->corr_elem_sites + 4 * (s->corr_elem_begin[a] + (e - e0) * k);
      for (int j = 0; j < k; ++j) {
        if (es[4 * j] != b0) continue;
        /* instance with site j placed on l */
        const int ci = i0 - es[4 * j + 1], cj = j0 - es[4 * j + 2], ck = k0 - es[4 * j + 3];
        long tau = 0, mul = 1, mulj = 0;
        for (int q = 0; q < k; ++q) {
          const int b = es[4 * q];
          long lq = site_index(s, b, ci + es[4 * q + 1], cj + es[4 * q + 2], ck + es[4 * q + 3]);
          if (q == j) mulj = mul;
          tau += occ[lq] * mul;
          mul *= s->n_occ[b];
        }
        const long tau_new = tau + (long)(nw - old) * mulj;
        const double *r0 = T + tau * nf;
        const double *r1 = T + tau_new * nf;
        for (int f = 0; f < nf; ++f) dcorr[fb + f] += (r1[f] - r0[f]) * inv_m;
      }
    }
  }
}

void orc_delta_corr(const orc_system *s, int8_t *occ, int n, const int *sites, const int *new_occ,
                    double *dcorr) {
  memset(dcorr, 0, sizeof(double) * s->n_corr);
  int8_t *saved = (int8_t *)malloc(n);
  for (int i = 0; i < n; ++i) {
    saved[i] = occ[sites[i]];
    delta_corr_site(s, occ, sites[i], new_occ[i], dcorr);
    occ[sites[i]] = (int8_t)new_occ[i];
  }
  for (int i = n - 1; i >= 0; --i) occ[sites[i]] = saved[i];
  free(saved);
}

double orc_delta(const orc_system *s, int8_t *occ, int n, const int *sites, const int *new_occ) {
  double *d = (double *)malloc(sizeof(double) * s->n_corr);
  orc_delta_corr(s, occ, n, sites, new_occ, d);
  double dE = 0.0;
  for (int i = 0; i < s->n_corr; ++i) dE += s->eci[i] * d[i];
  free(d);
  return dE;
}

/* ------------------------------------------------------------------ */
/* mt19937_64 (Matsumoto & Nishimura; the std::mt19937_64 parameters)  */
/* ------------------------------------------------------------------ */
typedef struct {
  uint64_t mt[312];
  int mti;
} mt64;

static void mt64_seed(mt64 *g, uint64_t seed) {
  g->mt[0] = seed;
  for (int i = 1; i < 312; ++i)
    g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
  g->mti = 312;
}

static uint64_t mt64_next(mt64 *g) {
  static const uint64_t mag01[2] = {0ULL, 0xB5026F5AA96619E9ULL};
  const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
  uint64_t x;
  if (g->mti >= 312) {
    int i;
    for (i = 0; i < 312 - 156; ++i) {
      x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
      g->mt[i] = g->mt[i + 156] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    }
    for (; i < 311; ++i) {
      x = (g->mt[i] & UM) | (g->mt[i + 1] & LM);
      g->mt[i] = g->mt[i + (156 - 312)] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    }
    x = (g->mt[311] & UM) | (g->mt[0] & LM);
    g->mt[311] = g->mt[155] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
    g->mti = 0;
  }
  x = g->mt[g->mti++];
  x ^= (x >> 29) & 0x5555555555555555ULL;
  x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
  x ^= (x << 37) & 0xFFF7EEE000000000ULL;
  x ^= (x >> 43);
  return x;
}

uint64_t orc_mt19937_64_nth(uint64_t seed, long long n) {
  mt64 g;
  mt64_seed(&g, seed);
  uint64_t x = 0;
  for (long long i = 0; i < n; ++i) x = mt64_next(&g);
  return x;
}

/* uniform integer in [0, n) by rejection (unbiased) */
static uint64_t rand_int(mt64 *g, uint64_t n) {
  if (n <= 1) return 0;
  uint64_t lim = UINT64_MAX - (UINT64_MAX % n);
  uint64_t x;
  do {
    x = mt64_next(g);
  } while (x >= lim);
  return x % n;
}

/* uniform real in [0, 1) with 53 random bits */
static double rand_real(mt64 *g) { return (double)(mt64_next(g) >> 11) * (1.0 / 9007199254740992.0); }

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                   */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

/* ------------------------------------------------------------------ */
/* Prepared delta: the reference evaluates occ_delta_value through a     */
/* per-supercell neighbour list (SuperNeighborList, StateData.cc:64-79)  */
/* and generated per-site delta functions restricted to non-zero ECIs    */
/* (FCC...Clexulator_default.cc:582-610).  The run loop uses the same    */
/* structure so the timed CPU path is a fair restatement.                */
/* ------------------------------------------------------------------ */
typedef struct {
  int n_inst;          /* instances containing a site of this sublattice */
  int *inst_orbit;     /* [n_inst] */
  int *inst_k;         /* sites in cluster */
  int *inst_pos;       /* position j of the changed site */
  int *inst_nbr;       /* [n_inst][kmax] neighbour-list index of each site (pos j: -1) */
  int *inst_mul;       /* [n_inst][kmax] radix multiplier of each site */
  int n_nbr;           /* unique neighbour offsets */
  int *nbr_off;        /* [n_nbr][4] */
} prep_sublat;

typedef struct {
  int kmax;
  prep_sublat *sub;    /* [n_sublat] */
  int *nlist;          /* [N][nbr_max] linear index of neighbour q of site l */
  int nbr_max;
  int *nz_funcs;       /* functions with non-zero ECI, per orbit: [nz_begin[a], nz_begin[a+1]) local idx */
  int *nz_begin;
} prepared;

static int find_or_add(int *offs, int *n, int b, int i, int j, int k) {
  for (int q = 0; q < *n; ++q)
    if (offs[4 * q] == b && offs[4 * q + 1] == i && offs[4 * q + 2] == j && offs[4 * q + 3] == k) return q;
  offs[4 * *n] = b;
  offs[4 * *n + 1] = i;
  offs[4 * *n + 2] = j;
  offs[4 * *n + 3] = k;
  return (*n)++;
}

static void prep_init(prepared *P, const orc_system *s) {
  int kmax = 1;
  for (int a = 0; a < s->n_orbits; ++a)
    if (s->orbit_k[a] > kmax) kmax = s->orbit_k[a];
  P->kmax = kmax;
  P->sub = (prep_sublat *)calloc(s->n_sublat, sizeof(prep_sublat));
  int total_inst = 0;
  for (int a = 0; a < s->n_orbits; ++a) total_inst += (s->corr_orbit_begin[a + 1] - s->corr_orbit_begin[a]) * s->orbit_k[a];
  P->nbr_max = 0;
  for (int b = 0; b < s->n_sublat; ++b) {
    prep_sublat *S = &P->sub[b];
    S->inst_orbit = (int *)malloc(sizeof(int) * (total_inst + 1));
    S->inst_k = (int *)malloc(sizeof(int) * (total_inst + 1));
    S->inst_pos = (int *)malloc(sizeof(int) * (total_inst + 1));
    S->inst_nbr = (int *)malloc(sizeof(int) * (total_inst + 1) * kmax);
    S->inst_mul = (int *)malloc(sizeof(int) * (total_inst + 1) * kmax);
    S->nbr_off = (int *)malloc(sizeof(int) * 4 * (total_inst * kmax + 1));
    S->n_inst = 0;
    S->n_nbr = 0;
    for (int a = 0; a < s->n_orbits; ++a) {
      const int k = s->orbit_k[a];
      if (k == 0) continue;
      const int e0 = s->corr_orbit_begin[a], e1 = s->corr_orbit_begin[a + 1];
      for (int e = e0; e < e1; ++e) {
        const int *es = s->corr_elem_sites + 4 * (s->corr_elem_begin[a] + (e - e0) * k);
        for (int j = 0; j < k; ++j) {
          if (es[4 * j] != b) continue;
          int t = S->n_inst++;
          S->inst_orbit[t] = a;
          S->inst_k[t] = k;
          S->inst_pos[t] = j;
          int mul = 1;
          for (int q = 0; q < k; ++q) {
            S->inst_mul[t * kmax + q] = mul;
            mul *= s->n_occ[es[4 * q]];
            if (q == j) {
              S->inst_nbr[t * kmax + q] = -1;
              continue;
            }
            S->inst_nbr[t * kmax + q] =
                find_or_add(S->nbr_off, &S->n_nbr, es[4 * q], es[4 * q + 1] - es[4 * j + 1],
                            es[4 * q + 2] - es[4 * j + 2], es[4 * q + 3] - es[4 * j + 3]);
          }
        }
      }
    }
    if (S->n_nbr > P->nbr_max) P->nbr_max = S->n_nbr;
  }
  const long nc = n_cells(s);
  const long N = nc * s->n_sublat;
  int nm = P->nbr_max > 0 ? P->nbr_max : 1;
  P->nlist = (int *)malloc(sizeof(int) * N * nm);
  for (long l = 0; l < N; ++l) {
    int b = (int)(l / nc);
    long c = l % nc;
    int i0 = (int)(c % s->L[0]), j0 = (int)((c / s->L[0]) % s->L[1]), k0 = (int)(c / ((long)s->L[0] * s->L[1]));
    prep_sublat *S = &P->sub[b];
    for (int q = 0; q < S->n_nbr; ++q)
      P->nlist[l * nm + q] = (int)site_index(s, S->nbr_off[4 * q], i0 + S->nbr_off[4 * q + 1],
                                             j0 + S->nbr_off[4 * q + 2], k0 + S->nbr_off[4 * q + 3]);
  }
  P->nz_begin = (int *)malloc(sizeof(int) * (s->n_orbits + 1));
  P->nz_funcs = (int *)malloc(sizeof(int) * (s->n_corr + 1));
  int nz = 0;
  for (int a = 0; a < s->n_orbits; ++a) {
    P->nz_begin[a] = nz;
    for (int f = 0; f < s->orbit_nfunc[a]; ++f)
      if (s->eci[s->orbit_func_begin[a] + f] != 0.0) P->nz_funcs[nz++] = f;
  }
  P->nz_begin[s->n_orbits] = nz;
}

static void prep_free(prepared *P, const orc_system *s) {
  for (int b = 0; b < s->n_sublat; ++b) {
    prep_sublat *S = &P->sub[b];
    free(S->inst_orbit);
    free(S->inst_k);
    free(S->inst_pos);
    free(S->inst_nbr);
    free(S->inst_mul);
    free(S->nbr_off);
  }
  free(P->sub);
  free(P->nlist);
  free(P->nz_begin);
  free(P->nz_funcs);
}

/* dE for changing site l to nw at the current occupation (restricted to
 * non-zero ECIs); occ_nb is scratch of size nbr_max */
static double prep_delta_site(const prepared *P, const orc_system *s, const int8_t *occ, long l, int nw,
                              int *occ_nb) {
  const long nc = n_cells(s);
  const int b = (int)(l / nc);
  const prep_sublat *S = &P->sub[b];
  const int nm = P->nbr_max > 0 ? P->nbr_max : 1;
  const int old = occ[l];
  if (old == nw) return 0.0;
  /* _point_prepare: read the neighbourhood once */
  const int *nl = P->nlist + l * nm;
  for (int q = 0; q < S->n_nbr; ++q) occ_nb[q] = occ[nl[q]];
  double dE = 0.0;
  const int kmax = P->kmax;
  for (int t = 0; t < S->n_inst; ++t) {
    const int a = S->inst_orbit[t];
    const int z0 = P->nz_begin[a], z1 = P->nz_begin[a + 1];
    if (z0 == z1) continue;
    const int k = S->inst_k[t];
    const int j = S->inst_pos[t];
    long tau = 0;
    for (int q = 0; q < k; ++q) {
      if (q == j) continue;
      tau += (long)occ_nb[S->inst_nbr[t * kmax + q]] * S->inst_mul[t * kmax + q];
    }
    const long t0 = tau + (long)old * S->inst_mul[t * kmax + j];
    const long t1 = tau + (long)nw * S->inst_mul[t * kmax + j];
    const int nf = s->orbit_nfunc[a];
    const double *T = s->orbit_tables + s->orbit_table_offset[a];
    const double *eci = s->eci + s->orbit_func_begin[a];
    double acc = 0.0;
    for (int z = z0; z < z1; ++z) {
      const int f = P->nz_funcs[z];
      acc += eci[f] * (T[t1 * nf + f] - T[t0 * nf + f]);
    }
    dE += acc / s->orbit_mult[a];
  }
  return dE;
}

/* sequential changes (swap = two flips, the second seeing the first) */
static double prep_delta(const prepared *P, const orc_system *s, int8_t *occ, int n, const int *sites,
                         const int *new_occ, int *occ_nb) {
  double dE = 0.0;
  int8_t saved[8];
  for (int i = 0; i < n; ++i) {
    saved[i] = occ[sites[i]];
    dE += prep_delta_site(P, s, occ, sites[i], new_occ[i], occ_nb);
    occ[sites[i]] = (int8_t)new_occ[i];
  }
  for (int i = n - 1; i >= 0; --i) occ[sites[i]] = saved[i];
  return dE;
}

/* exported for the test of the prepared path against orc_delta */
double orc_delta_prepared(const orc_system *s, int8_t *occ, int n, const int *sites, const int *new_occ) {
  prepared P;
  prep_init(&P, s);
  int *nb = (int *)malloc(sizeof(int) * (P.nbr_max + 1));
  double d = prep_delta(&P, s, occ, n, sites, new_occ, nb);
  free(nb);
  prep_free(&P, s);
  return d;
}

/* ------------------------------------------------------------------ */
/* Metropolis run: occupation_metropolis.hh:72-123 restated              */
/* ------------------------------------------------------------------ */
static double now_s(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return t.tv_sec + 1e-9 * t.tv_nsec;
}

typedef struct {
  /* OccLocation restatement: per (asym, occupant) candidate list of sites */
  int n_asym;
  int n_occ_max;
  long *lists;     /* [n_cand][N] */
  long *count;     /* [n_cand] */
  long *pos;       /* [N] position of site within its list */
  long N;
  long mol_size;   /* mutable sites */
  long *mutable_sites;
} occ_location;

static void occloc_init(occ_location *L, const orc_system *s, const int8_t *occ, int n_occ_max) {
  const long nc = n_cells(s);
  L->N = nc * s->n_sublat;
  L->n_asym = 0;
  for (int b = 0; b < s->n_sublat; ++b)
    if (s->asym[b] + 1 > L->n_asym) L->n_asym = s->asym[b] + 1;
  L->n_occ_max = n_occ_max;
  int ncand = L->n_asym * n_occ_max;
  L->lists = (long *)malloc(sizeof(long) * ncand * L->N);
  L->count = (long *)calloc(ncand, sizeof(long));
  L->pos = (long *)malloc(sizeof(long) * L->N);
  L->mutable_sites = (long *)malloc(sizeof(long) * L->N);
  L->mol_size = 0;
  for (long l = 0; l < L->N; ++l) {
    int b = (int)(l / nc);
    if (s->n_occ[b] < 2) continue;
    L->mutable_sites[L->mol_size++] = l;
    int c = s->asym[b] * n_occ_max + occ[l];
    L->pos[l] = L->count[c];
    L->lists[(long)c * L->N + L->count[c]++] = l;
  }
}

static void occloc_free(occ_location *L) {
  free(L->lists);
  free(L->count);
  free(L->pos);
  free(L->mutable_sites);
}

/* OccLocation::apply restated: move site l from candidate list of its old
 * occupant to the list of the new one */
static void occloc_apply(occ_location *L, const orc_system *s, int8_t *occ, long l, int nw) {
  const long nc = n_cells(s);
  int b = (int)(l / nc);
  int u = s->asym[b];
  int c_old = u * L->n_occ_max + occ[l];
  int c_new = u * L->n_occ_max + nw;
  long p = L->pos[l];
  long last = L->lists[(long)c_old * L->N + L->count[c_old] - 1];
  L->lists[(long)c_old * L->N + p] = last;
  L->pos[last] = p;
  L->count[c_old]--;
  L->pos[l] = L->count[c_new];
  L->lists[(long)c_new * L->N + L->count[c_new]++] = l;
  occ[l] = (int8_t)nw;
}

int orc_run(const orc_system *s, int8_t *occ, const orc_run_params *p, double *samples_E,
            int *samples_counts, orc_run_result *res) {
  const double t0 = now_s();
  const long nc = n_cells(s);
  occ_location L;
  occloc_init(&L, s, occ, p->n_occ_max);
  mt64 g;
  mt64_seed(&g, p->seed);
  const double beta = 1.0 / (p->kb * p->temperature);
  double E = orc_energy(s, occ);
  long long n_attempt = 0, n_accept = 0;
  const long long steps_per_pass = L.mol_size;
  const long long total_passes = p->n_equil_passes + p->n_passes;
  int sites[2], nocc[2];

  /* canonical swap types: (asym u, occupant a) <-> (asym u, occupant b), a<b,
   * only for asym units with >= 2 occupants */
  int n_swaps = 0;
  int swap_u[64], swap_a[64], swap_b[64];
  if (p->kind == 0) {
    for (int u = 0; u < L.n_asym; ++u) {
      int b0 = -1;
      for (int b = 0; b < s->n_sublat; ++b)
        if (s->asym[b] == u) b0 = b;
      if (b0 < 0 || s->n_occ[b0] < 2) continue;
      for (int a = 0; a < s->n_occ[b0]; ++a)
        for (int bb = a + 1; bb < s->n_occ[b0]; ++bb) {
          if (n_swaps < 64) {
            swap_u[n_swaps] = u;
            swap_a[n_swaps] = a;
            swap_b[n_swaps] = bb;
            n_swaps++;
          }
        }
    }
  }
  prepared P;
  prep_init(&P, s);
  int *occ_nb = (int *)malloc(sizeof(int) * (P.nbr_max + 1));
  long long sample_i = 0;
  for (long long pass = 0; pass < total_passes; ++pass) {
    for (long long step = 0; step < steps_per_pass; ++step) {
      double dE;
      int n;
      if (p->kind == 0) {
        /* propose_canonical_event restated: uniform swap type among those
         * with both candidate lists non-empty, then a uniform site of each */
        int avail[64], na = 0;
        for (int t = 0; t < n_swaps; ++t) {
          int ca = swap_u[t] * L.n_occ_max + swap_a[t];
          int cb = swap_u[t] * L.n_occ_max + swap_b[t];
          if (L.count[ca] > 0 && L.count[cb] > 0) avail[na++] = t;
        }
        if (na == 0) break;
        int t = avail[rand_int(&g, (uint64_t)na)];
        int ca = swap_u[t] * L.n_occ_max + swap_a[t];
        int cb = swap_u[t] * L.n_occ_max + swap_b[t];
        long la = L.lists[(long)ca * L.N + rand_int(&g, (uint64_t)L.count[ca])];
        long lb = L.lists[(long)cb * L.N + rand_int(&g, (uint64_t)L.count[cb])];
        sites[0] = (int)la;
        nocc[0] = swap_b[t];
        sites[1] = (int)lb;
        nocc[1] = swap_a[t];
        n = 2;
        dE = prep_delta(&P, s, occ, n, sites, nocc, occ_nb);
      } else {
        /* propose_semigrand_canonical_event restated: uniform mutable site,
         * uniform new occupant among the other allowed ones */
        long l = L.mutable_sites[rand_int(&g, (uint64_t)L.mol_size)];
        int b = (int)(l / nc);
        int cur = occ[l];
        int r = (int)rand_int(&g, (uint64_t)(s->n_occ[b] - 1));
        int nw = r >= cur ? r + 1 : r;
        sites[0] = (int)l;
        nocc[0] = nw;
        n = 1;
        dE = prep_delta(&P, s, occ, n, sites, nocc, occ_nb);
        if (p->chem) dE -= p->chem[b * p->n_occ_max + nw] - p->chem[b * p->n_occ_max + cur];
      }
      /* metropolis_acceptance restated (occupation_metropolis.hh:104-105) */
      int accept = 0;
      if (dE < 0.0)
        accept = 1;
      else
        accept = rand_real(&g) < exp(-dE * beta);
      n_attempt++;
      if (accept) {
        n_accept++;
        /* track formation energy (potential minus chem term for SGC) */
        double dEf = dE;
        if (p->kind == 1 && p->chem) {
          int b = (int)(sites[0] / nc);
          dEf += p->chem[b * p->n_occ_max + nocc[0]] - p->chem[b * p->n_occ_max + occ[sites[0]]];
        }
        E += dEf;
        for (int i = 0; i < n; ++i) occloc_apply(&L, s, occ, sites[i], nocc[i]);
      }
    }
    if (pass >= p->n_equil_passes) {
      double Es = E;
      if (p->full_recompute_samplers) {
        /* clex.formation_energy and potential_energy samplers each evaluate
         * the cluster expansion in full (sampling_functions.cc:141-151,
         * 274-285; CanonicalCalculator.cc:126-128) */
        double e1 = orc_energy(s, occ);
        double e2 = orc_energy(s, occ);
        Es = 0.5 * (e1 + e2);
      }
      if (samples_E) samples_E[sample_i] = Es;
      if (samples_counts) {
        int *cnt = samples_counts + sample_i * p->n_occ_max;
        for (int o = 0; o < p->n_occ_max; ++o) cnt[o] = 0;
        for (long i = 0; i < L.mol_size; ++i) cnt[occ[L.mutable_sites[i]]]++;
      }
      sample_i++;
    }
  }
  free(occ_nb);
  prep_free(&P, s);
  occloc_free(&L);
  if (res) {
    res->n_attempt = n_attempt;
    res->n_accept = n_accept;
    res->E_final = E;
    res->elapsed_s = now_s() - t0;
  }
  return 0;
}

typedef struct {
  const orc_system *s;
  int n_chains, tid, n_threads;
  int8_t *occ;
  const double *temperature, *chem;
  const orc_run_params *p;
  const uint64_t *seeds;
  long long n_attempt;
} batch_arg;

static void *batch_worker(void *v) {
  batch_arg *a = (batch_arg *)v;
  const long N = n_cells(a->s) * a->s->n_sublat;
  for (int c = a->tid; c < a->n_chains; c += a->n_threads) {
    orc_run_params p = *a->p;
    p.temperature = a->temperature[c];
    p.seed = a->seeds[c];
    if (a->chem) p.chem = a->chem + (long)c * a->s->n_sublat * p.n_occ_max;
    orc_run_result r;
    orc_run(a->s, a->occ + (long)c * N, &p, NULL, NULL, &r);
    a->n_attempt += r.n_attempt;
  }
  return NULL;
}

int orc_run_batch(const orc_system *s, int n_chains, int8_t *occ, const double *temperature,
                  const double *chem, const orc_run_params *p, const uint64_t *seeds, int n_threads,
                  long long *n_attempt_total, double *elapsed_s) {
  if (n_threads < 1) n_threads = 1;
  pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * n_threads);
  batch_arg *args = (batch_arg *)calloc(n_threads, sizeof(batch_arg));
  double t0 = now_s();
  for (int t = 0; t < n_threads; ++t) {
    args[t] = (batch_arg){s, n_chains, t, n_threads, occ, temperature, chem, p, seeds, 0};
    pthread_create(&th[t], NULL, batch_worker, &args[t]);
  }
  long long tot = 0;
  for (int t = 0; t < n_threads; ++t) {
    pthread_join(th[t], NULL);
    tot += args[t].n_attempt;
  }
  if (elapsed_s) *elapsed_s = now_s() - t0;
  if (n_attempt_total) *n_attempt_total = tot;
  free(th);
  free(args);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Exact enumeration on tiny supercells                                 */
/* ------------------------------------------------------------------ */
int orc_enumerate_energies(const orc_system *s, int n_occ_max, const int *fixed_counts, double *E_out,
                           long long n_states, long long st0, long long st1) {
  const long nc = n_cells(s);
  const long N = nc * s->n_sublat;
  long long n = 1;
  for (long l = 0; l < N; ++l) {
    n *= s->n_occ[l / nc];
    if (n > (1LL << 26)) return -1;
  }
  if (n != n_states || st0 < 0 || st1 > n_states || st0 > st1) return -2;
  int8_t *occ = (int8_t *)calloc(N, 1);
  for (long long st = st0; st < st1; ++st) {
    long long x = st;
    int cnt[16] = {0};
    for (long l = 0; l < N; ++l) {
      int r = s->n_occ[l / nc];
      occ[l] = (int8_t)(x % r);
      x /= r;
      if (r >= 2) cnt[occ[l]]++;
    }
    int ok = 1;
    if (fixed_counts)
      for (int o = 0; o < n_occ_max; ++o)
        if (cnt[o] != fixed_counts[o]) ok = 0;
    E_out[st] = ok ? orc_energy(s, occ) : NAN;
  }
  free(occ);
  return 0;
}

int orc_enumerate(const orc_system *s, int kind, double temperature, double kb, const double *chem,
                  int n_occ_max, const int *fixed_counts, double *avg, double *probs, long long n_probs) {
  const long nc = n_cells(s);
  const long N = nc * s->n_sublat;
  long long n_states = 1;
  for (long l = 0; l < N; ++l) {
    n_states *= s->n_occ[l / nc];
    if (n_states > (1LL << 26)) return -1;
  }
  if (probs && n_probs != n_states) return -2;
  double *E = (double *)malloc(sizeof(double) * n_states);
  double *W = (double *)malloc(sizeof(double) * n_states);
  int *n1 = (int *)malloc(sizeof(int) * n_states);
  char *valid = (char *)malloc(n_states);
  int8_t *occ = (int8_t *)calloc(N, 1);
  const double beta = 1.0 / (kb * temperature);
  double wmin = INFINITY;
  for (long long st = 0; st < n_states; ++st) {
    long long x = st;
    int cnt[16] = {0};
    for (long l = 0; l < N; ++l) {
      int r = s->n_occ[l / nc];
      occ[l] = (int8_t)(x % r);
      x /= r;
      if (r >= 2) cnt[occ[l]]++;
    }
    valid[st] = 1;
    if (kind == 0) {
      for (int o = 0; o < n_occ_max; ++o)
        if (cnt[o] != fixed_counts[o]) valid[st] = 0;
    }
    n1[st] = cnt[1];
    if (!valid[st]) continue;
    E[st] = orc_energy(s, occ);
    double w = E[st];
    if (kind == 1 && chem) {
      for (long l = 0; l < N; ++l) {
        int b = (int)(l / nc);
        if (s->n_occ[b] >= 2) w -= chem[b * n_occ_max + occ[l]];
      }
    }
    W[st] = w;
    if (w < wmin) wmin = w;
  }
  double Z = 0, sE = 0, sE2 = 0, sn = 0, sn2 = 0, sEn = 0;
  for (long long st = 0; st < n_states; ++st) {
    double p = valid[st] ? exp(-beta * (W[st] - wmin)) : 0.0;
    if (probs) probs[st] = p;
    if (!valid[st]) continue;
    Z += p;
    sE += p * E[st];
    sE2 += p * E[st] * E[st];
    sn += p * n1[st];
    sn2 += p * (double)n1[st] * n1[st];
    sEn += p * E[st] * n1[st];
  }
  if (probs)
    for (long long st = 0; st < n_states; ++st) probs[st] /= Z;
  avg[0] = sE / Z;
  avg[1] = sE2 / Z;
  avg[2] = sn / Z;
  avg[3] = sn2 / Z;
  avg[4] = sEn / Z;
  avg[5] = Z;
  free(E);
  free(W);
  free(n1);
  free(valid);
  free(occ);
  return 0;
}
