// Host side of the clexmc C ABI: owns device memory, prepares per-supercell
// padded geometry (halo copy lists, neighbour offsets, independent-set
// classes) and launches the gfx950 kernels on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <array>
#include <numeric>
#include <random>
#include <set>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "../../include/clexmc.h"
#include "clexmc_device.h"
#include "clexmc_kernels.h"
#include "clexmc_jit.h"

using namespace clexmc;

namespace {

struct HostSupercell {
  DevSupercell dev;  // device pointers filled in
  std::vector<void *> allocs;
  // binary site programs (host copies, for run-time specialisation)
  std::vector<BinSub> bin_sub;
  std::vector<int> bin_off;
  std::vector<GenSub> gen_sub;
  std::vector<int> gen_prog;
  // specialised kernels per (kind, packed, lanes, launch bound 512 / 1024):
  // 0 not tried, 1 loaded, -1 unavailable
  int jit_state[3][2][2][2] = {};
  size_t jit_lds_fail = 0;  // smallest dynamic LDS size at which a specialised launch was refused
  std::vector<hipModule_t> jit_mods;
  hipFunction_t jit_fn[3][2][2][2] = {};  // [kind][packed][lanes][block above 512]
  hipFunction_t jit_nfold = nullptr;  // N-fold kernel of the SGC (unpacked, one-lane) module
  hipFunction_t jit_cube[2] = {nullptr, nullptr};  // register-cube kernel of that module [block above 512]
  // register-cube sweep (binary, one sublattice, classes = the residues of
  // the cells mod 3 listed in block order): eligible, and its layout
  bool cube_ok = false;
  CubeSpec cube{};
  int nb_total = 0;  // N-fold neighbour-list entries over all sublattices
};

}  // namespace

struct clexmc_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // system (host copies)
  int n_sublat = 0, n_occ_max = 0, n_func_max = 0, n_orbits = 0, n_corr = 0, binary = 0;
  std::vector<int> n_occ, asym, orbit_k, orbit_mult, orbit_nfunc, orbit_func_begin;
  std::vector<int> flower_begin, entry_k, entry_site_begin, entry_sites, entry_table, entry_orbit;
  std::vector<int> entry_mul;
  std::vector<double> bin_w;
  int n_flower_tables = 0;
  std::vector<int> entry_home;  // home sublattice of each entry
  std::vector<double> eci;
  int halo = 0;
  DevSystem dsys{};
  DevSystem *d_sys = nullptr;
  std::vector<void *> sys_allocs;
  // supercells
  std::vector<HostSupercell> scs;
  DevSupercell *d_scs = nullptr;
  bool scs_dirty = true;
  // chains
  int n_chains = 0;
  std::vector<int> chain_sc, chain_off;
  // chains grouped by supercell (one launch per group when shapes differ)
  std::vector<std::vector<int>> groups;
  std::vector<int> group_off;
  int *d_group_map = nullptr;
  // fused launches (chains on several supercell shapes, one specialised
  // kernel): per key (kind, launch bound, supercell and image format of each
  // group) the kernel, nullptr when it could not be built or launched
  std::map<std::string, hipFunction_t> fused_fn;
  std::vector<hipModule_t> fused_mods;
  std::vector<hipStream_t> side;    // side streams of the grouped launches (default priority)
  std::vector<hipEvent_t> side_ev;
  long long occ_total = 0;
  ChainParams *d_params = nullptr;
  ChainState *d_state = nullptr;
  int8_t *d_occ = nullptr;
  long long *d_binC = nullptr;
  double *d_sampE = nullptr;
  int *d_sampN = nullptr;
  int *d_sample_count = nullptr;
  int sample_cap = 0;
  // upper bound of the samples any chain holds since the last load or clear
  // (kept on the host so that planning a run needs no device read)
  long long samples_bound = 0;
  int max_lds = 0, n_cu = 0;        // device attributes, read at creation
  unsigned char *noalloc = nullptr; // the device's staging array (clexmc_delta_energy_batch_noalloc)
  int block_size = 0;
  int nfold_threads = 0;  // N-fold kernel block size (0: automatic)
  int packing = 0;                  // LDS image: 0 automatic, 1 int8, 2 packed 2-bit
  bool last_packed = false;
  int last_specialized = 0, last_threads = 0, last_lanes = 1, last_slabs = 1, last_fused = 0, last_cube = 0;
  int fuse = 1;  // chains on several supercell shapes: one fused launch (1) or one launch per shape (0)
  int cube = 1;  // register-cube sweep where the specialised kernel has one (1) or the class loop (0)
  int slabs = 0;  // slabs per chain: 0 automatic, 1 off, >= 2 requested
  SlabRecord *d_slab_rec = nullptr;  // [n_chains][kSlabMax][kSlabRecords] (slab launches)
  int slab_rec_chains = 0;
  int global_swaps = 1;             // canonical serial random-pair swaps per pass
  // semi-grand canonical multiswaps (device tables, kept until replaced)
  int n_mswaps = 0, mswap_per_pass = 0;
  int4 *d_mswap_parts = nullptr;
  int *d_mswap_begin = nullptr, *d_mswap_rev = nullptr;
  int row_pitch = -1;               // padded image row pitch: -1 automatic, else extra bytes per row
  int n_slots = 0;                  // binary path: orbits with >= 1 site
  int lanes_per_site = 0;           // 0 automatic (1; the autotuner may pick 2), 1, 2
  int specialize = 1;               // run-time specialised binary kernels (1 when possible, 0 never)
  std::string jit_log;              // last specialisation compiler log
  hipGraph_t graph = nullptr;       // captured block of passes
  hipGraphExec_t graph_exec = nullptr;
  int graph_replays_left = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0.0;
  // learned proposal
  int n_env = 0;
  std::vector<int4> env_sites;      // [n_sublat][kEnvMax]
  int4 *d_env_sites = nullptr;
  MlpDev *d_mlp = nullptr;
  float *d_mlp_params = nullptr, *d_adam_m = nullptr, *d_adam_v = nullptr;
  long long adam_t = 0;
  bool mlp_initialized = false;     // parameters set by the caller or a training run
  float mlp_eps = 0.05f;
  int proposal = 0;                 // 0 plain, 1 learned (MLP), 2 grouped uniform, 3 informed
  int rounds_per_group = 64;
};

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      h->err = std::string(#expr) + ": " + hipGetErrorString(_e);                  \
      return CLEXMC_ERR_HIP;                                                       \
    }                                                                              \
  } while (0)

#define ARG_CHECK(cond, msg)   \
  do {                         \
    if (!(cond)) {             \
      h->err = (msg);          \
      return CLEXMC_ERR_ARG;   \
    }                          \
  } while (0)

template <typename T>
static hipError_t upload(clexmc_handle *h, std::vector<void *> &allocs, const T *src, size_t n, T **dst) {
  *dst = nullptr;
  if (n == 0) n = 1;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, n * sizeof(T));
  if (e != hipSuccess) return e;
  allocs.push_back(p);
  if (src) {
    e = hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
  } else {
    e = hipMemset(p, 0, n * sizeof(T));
    if (e != hipSuccess) return e;
  }
  *dst = (T *)p;
  return hipSuccess;
}

static void free_all(std::vector<void *> &v) {
  for (void *p : v) hipFree(p);
  v.clear();
}

struct DevBufs {
  std::vector<void *> p;
  ~DevBufs() {
    for (void *x : p) hipFree(x);
  }
  template <typename T>
  hipError_t alloc(T **out, size_t n) {
    void *x = nullptr;
    hipError_t e = hipMalloc(&x, sizeof(T) * std::max<size_t>(n, 1));
    if (e == hipSuccess) p.push_back(x);
    *out = (T *)x;
    return e;
  }
};

extern "C" {

const char *clexmc_version(void) { return "clexmc 0.1 (gfx950)"; }

const char *clexmc_last_error(clexmc_handle *h) { return h ? h->err.c_str() : "null handle"; }

clexmc_status clexmc_create(const clexmc_system_desc *s, int device, clexmc_handle **out) {
  if (!out || !s) return CLEXMC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CLEXMC_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return CLEXMC_ERR_ARG;
  clexmc_handle *h = new clexmc_handle();
  h->device = device;
  *out = h;
  HIP_TRY(hipSetDevice(device));
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&h->ev0));
  HIP_TRY(hipEventCreate(&h->ev1));
  HIP_TRY(hipDeviceGetAttribute(&h->max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
  HIP_TRY(hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, device));
  HIP_TRY(noalloc_scratch(device, &h->noalloc));

  ARG_CHECK(s->n_sublat >= 1 && s->n_sublat <= kMaxSublat, "n_sublat out of range (1..8)");
  ARG_CHECK(s->n_occ_max >= 1 && s->n_occ_max <= kMaxOcc, "n_occ_max out of range (1..4)");
  ARG_CHECK(s->n_func_max >= 0 && s->n_func_max <= kMaxFunc, "n_func_max out of range (0..3)");
  ARG_CHECK(s->n_orbits >= 1 && s->n_corr >= 1, "empty cluster expansion");
  h->n_sublat = s->n_sublat;
  h->n_occ_max = s->n_occ_max;
  h->n_func_max = s->n_func_max;
  h->n_orbits = s->n_orbits;
  h->n_corr = s->n_corr;
  {
    int n_slots = 0;
    for (int a = 0; a < s->n_orbits; ++a)
      if (s->orbit_k[a] >= 1) n_slots++;
    h->binary = s->binary && s->n_orbits <= kMaxOrbits && n_slots <= kMaxSlots;
  }
  h->n_occ.assign(s->n_occ, s->n_occ + s->n_sublat);
  h->asym.resize(s->n_sublat);
  for (int b = 0; b < s->n_sublat; ++b) h->asym[b] = s->asym ? s->asym[b] : s->n_occ[b];
  h->orbit_k.assign(s->orbit_k, s->orbit_k + s->n_orbits);
  h->orbit_mult.assign(s->orbit_mult, s->orbit_mult + s->n_orbits);
  h->orbit_nfunc.assign(s->orbit_nfunc, s->orbit_nfunc + s->n_orbits);
  h->orbit_func_begin.assign(s->orbit_func_begin, s->orbit_func_begin + s->n_orbits);
  h->flower_begin.assign(s->flower_begin, s->flower_begin + s->n_sublat + 1);
  h->entry_k.assign(s->entry_k, s->entry_k + s->n_entries);
  h->entry_site_begin.assign(s->entry_site_begin, s->entry_site_begin + s->n_entries + 1);
  h->entry_sites.assign(s->entry_sites, s->entry_sites + 4 * s->n_entry_sites);
  h->entry_table.assign(s->entry_table, s->entry_table + s->n_entries);
  h->entry_orbit.assign(s->entry_orbit, s->entry_orbit + s->n_entries);
  h->n_flower_tables = s->n_flower_tables;
  h->eci.assign(s->eci, s->eci + s->n_corr);
  ARG_CHECK(h->flower_begin[s->n_sublat] == s->n_entries, "flower_begin inconsistent with n_entries");
  ARG_CHECK(h->entry_site_begin[s->n_entries] == s->n_entry_sites, "entry_site_begin inconsistent");

  DevSystem &d = h->dsys;
  d.n_sublat = s->n_sublat;
  d.n_occ_max = s->n_occ_max;
  d.n_func_max = s->n_func_max;
  d.n_orbits = s->n_orbits;
  d.n_corr = s->n_corr;
  d.binary = h->binary;
  for (int b = 0; b < s->n_sublat; ++b) {
    d.n_occ[b] = s->n_occ[b];
    d.asym[b] = h->asym[b];
    ARG_CHECK(s->n_occ[b] >= 1 && s->n_occ[b] <= kMaxOcc, "n_occ out of range");
    d.n_func[b] = s->n_occ[b] - 1;
    for (int f = 0; f < kMaxFunc; ++f)
      for (int o = 0; o < kMaxOcc; ++o) {
        float v = 0.f;
        if (f < s->n_func_max && o < s->n_occ_max)
          v = (float)s->site_func[(b * s->n_func_max + f) * s->n_occ_max + o];
        d.site_func[b][f][o] = v;
      }
  }
  // home sublattice and radix multipliers of entry sites; halo width
  h->entry_home.assign(s->n_entries, 0);
  std::vector<int> &entry_mul = h->entry_mul;
  entry_mul.assign(std::max(1, s->n_entry_sites), 1);
  int halo = 0;
  for (int b = 0; b < s->n_sublat; ++b)
    for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e) {
      h->entry_home[e] = b;
      int mul = 1;
      for (int q = h->entry_site_begin[e]; q < h->entry_site_begin[e + 1]; ++q) {
        const int bq = h->entry_sites[4 * q];
        ARG_CHECK(bq >= 0 && bq < s->n_sublat, "entry site sublattice out of range");
        entry_mul[q] = mul;
        mul *= s->n_occ[bq];
        for (int c = 1; c < 4; ++c) halo = std::max(halo, std::abs(h->entry_sites[4 * q + c]));
      }
    }
  h->halo = halo;
  // flower tables (FP32)
  std::vector<float> ft(std::max(1, s->n_flower_tables));
  for (int i = 0; i < s->n_flower_tables; ++i) ft[i] = (float)s->flower_tables[i];
  auto &A = h->sys_allocs;
  HIP_TRY(upload(h, A, h->flower_begin.data(), h->flower_begin.size(), (int **)&d.flower_begin));
  HIP_TRY(upload(h, A, h->entry_k.data(), h->entry_k.size(), (int **)&d.entry_k));
  HIP_TRY(upload(h, A, h->entry_site_begin.data(), h->entry_site_begin.size(), (int **)&d.entry_site_begin));
  HIP_TRY(upload(h, A, (const int4 *)h->entry_sites.data(), (size_t)s->n_entry_sites, (int4 **)&d.entry_sites));
  HIP_TRY(upload(h, A, entry_mul.data(), entry_mul.size(), (int **)&d.entry_mul));
  HIP_TRY(upload(h, A, h->entry_table.data(), h->entry_table.size(), (int **)&d.entry_table));
  HIP_TRY(upload(h, A, ft.data(), ft.size(), (float **)&d.flower_tables));
  // binary slots: orbits with >= 1 site, in orbit order
  if (h->binary) {
    d.n_slots = 0;
    h->bin_w.assign(s->n_orbits, 0.0);
    for (int a = 0; a < s->n_orbits; ++a) {
      ARG_CHECK(s->orbit_nfunc[a] == 1, "binary path requires one function per orbit");
      d.bin_w64[a] = s->bin_orbit_weight[a];
      h->bin_w[a] = s->bin_orbit_weight[a];
      if (s->orbit_k[a] >= 1) d.slot_orbit[d.n_slots++] = a;
    }
    h->n_slots = d.n_slots;
  }
  // environment of the learned proposal: sites of the three shortest pair orbits
  {
    std::vector<int> pair_orbits;
    for (int a = 0; a < s->n_orbits && pair_orbits.size() < 3; ++a)
      if (s->orbit_k[a] == 2) pair_orbits.push_back(a);
    h->env_sites.assign((size_t)s->n_sublat * kEnvMax, make_int4(0, 0, 0, 0));
    int n_env = -1;
    for (int b = 0; b < s->n_sublat; ++b) {
      if (s->n_occ[b] < 2) continue;
      int cnt = 0;
      for (int a : pair_orbits)
        for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e) {
          if (h->entry_orbit[e] != a) continue;
          const int q = h->entry_site_begin[e];
          if (cnt < kEnvMax)
            h->env_sites[(size_t)b * kEnvMax + cnt] = make_int4(h->entry_sites[4 * q], h->entry_sites[4 * q + 1],
                                                                h->entry_sites[4 * q + 2], h->entry_sites[4 * q + 3]);
          cnt++;
        }
      cnt = std::min(cnt, kEnvMax);
      n_env = n_env < 0 ? cnt : std::min(n_env, cnt);
    }
    h->n_env = std::max(0, n_env);
    const int np = kHidden * kEnvMax + 2 * kHidden + 1;
    HIP_TRY(upload(h, A, h->env_sites.data(), h->env_sites.size(), &h->d_env_sites));
    HIP_TRY(upload(h, A, (const MlpDev *)nullptr, 1, &h->d_mlp));
    HIP_TRY(upload(h, A, (const float *)nullptr, (size_t)np, &h->d_mlp_params));
    HIP_TRY(upload(h, A, (const float *)nullptr, (size_t)np, &h->d_adam_m));
    HIP_TRY(upload(h, A, (const float *)nullptr, (size_t)np, &h->d_adam_v));
    // zero network (w = eps + 1/2 everywhere) until parameters are set or trained
    HIP_TRY(launch_mlp_publish(h->d_mlp_params, h->d_mlp, h->n_env, h->mlp_eps, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  // correlations
  std::vector<int> func_orbit(s->n_corr, 0);
  for (int a = 0; a < s->n_orbits; ++a)
    for (int f = 0; f < s->orbit_nfunc[a]; ++f) func_orbit[s->orbit_func_begin[a] + f] = a;
  HIP_TRY(upload(h, A, s->orbit_k, (size_t)s->n_orbits, (int **)&d.orbit_k_all));
  HIP_TRY(upload(h, A, s->orbit_mult, (size_t)s->n_orbits, (int **)&d.orbit_mult));
  HIP_TRY(upload(h, A, s->orbit_nfunc, (size_t)s->n_orbits, (int **)&d.orbit_nfunc));
  HIP_TRY(upload(h, A, s->orbit_func_begin, (size_t)s->n_orbits, (int **)&d.orbit_func_begin));
  HIP_TRY(upload(h, A, s->corr_orbit_begin, (size_t)s->n_orbits + 1, (int **)&d.corr_orbit_begin));
  HIP_TRY(upload(h, A, s->corr_elem_begin, (size_t)s->n_orbits + 1, (int **)&d.corr_elem_begin));
  HIP_TRY(upload(h, A, (const int4 *)s->corr_elem_sites, (size_t)s->n_corr_elem_sites, (int4 **)&d.corr_elem_sites));
  HIP_TRY(upload(h, A, s->orbit_table_offset, (size_t)s->n_orbits, (int **)&d.orbit_table_offset));
  HIP_TRY(upload(h, A, s->orbit_tables, (size_t)s->n_orbit_tables, (double **)&d.orbit_tables));
  HIP_TRY(upload(h, A, func_orbit.data(), func_orbit.size(), (int **)&d.func_orbit));
  HIP_TRY(upload(h, A, &h->dsys, 1, &h->d_sys));
  return CLEXMC_OK;
}


// ---------------------------------------------------------------------------
// Independent-set partition computed by the library (clexmc_add_supercell with
// n_classes == 0): periodic colouring by residues of the cell coordinates when
// one exists (equal class sizes), otherwise a greedy colouring; plus
// canonical cross-class pair phases (pairs of s on sublattice b and t on
// sublattice b2, with the same number of occupants, in the cell translated by
// a random v; greedily kept so that no cluster contains sites of two pairs
// and the two sites of a pair share no cluster).  Selection never looks at
// occupations.
// ---------------------------------------------------------------------------
static uint64_t splitmix64(uint64_t &x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static clexmc_status auto_partition(clexmc_handle *h, const int L[3], int n_pair_req, uint64_t seed,
                                    std::vector<int> &cb, std::vector<int> &cs, std::vector<int> &pb,
                                    std::vector<int> &ps) {
  const int nb = h->n_sublat;
  const long nc = (long)L[0] * L[1] * L[2];
  const long N = nc * nb;
  auto lin = [&](int b, int i, int j, int k) {
    i = ((i % L[0]) + L[0]) % L[0];
    j = ((j % L[1]) + L[1]) % L[1];
    k = ((k % L[2]) + L[2]) % L[2];
    return (long)b * nc + i + (long)L[0] * (j + (long)L[1] * k);
  };
  // interaction offsets per sublattice
  std::vector<std::vector<std::array<int, 4>>> offs(nb);
  for (int b = 0; b < nb; ++b) {
    std::set<std::array<int, 4>> u;
    for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e)
      for (int q = h->entry_site_begin[e]; q < h->entry_site_begin[e + 1]; ++q)
        u.insert({h->entry_sites[4 * q], h->entry_sites[4 * q + 1], h->entry_sites[4 * q + 2], h->entry_sites[4 * q + 3]});
    offs[b].assign(u.begin(), u.end());
  }
  // clusters must not wrap onto themselves
  for (int b = 0; b < nb; ++b)
    for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e) {
      std::vector<long> sites = {lin(b, 0, 0, 0)};
      for (int q = h->entry_site_begin[e]; q < h->entry_site_begin[e + 1]; ++q)
        sites.push_back(lin(h->entry_sites[4 * q], h->entry_sites[4 * q + 1], h->entry_sites[4 * q + 2],
                            h->entry_sites[4 * q + 3]));
      std::sort(sites.begin(), sites.end());
      ARG_CHECK(std::adjacent_find(sites.begin(), sites.end()) == sites.end(),
                "supercell too small for the cluster expansion: a cluster wraps onto itself");
    }
  // periodic colouring of the cells: diagonal residues (i mod m1, j mod m2,
  // k mod m3) or a linear form (i + a j + c k) mod M, conflict-free (no
  // same-sublattice interaction offset inside a class) with at most
  // kMaxClassSites cells per class.  Score: the parities (mod 2) of
  // nearest-neighbour offsets among the class translations (with none,
  // period-2 orderings driven by nearest neighbours such as L1_2, L1_0 or B2
  // can leave every class pure).  Choice: among the colourings with a
  // positive score (all, if none has one), the fewest classes, except that
  // one holding every parity is kept when it needs at most 25 % more classes.
  std::set<std::array<int, 3>> par;  // parities of same-sublattice offsets in the shortest pair orbit
  for (int a = 0; a < h->n_orbits && par.empty(); ++a) {
    if (h->orbit_k[a] != 2) continue;
    for (int b = 0; b < nb; ++b)
      for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e) {
        if (h->entry_orbit[e] != a) continue;
        const int q = h->entry_site_begin[e];
        if (h->entry_sites[4 * q] != b) continue;
        par.insert({((h->entry_sites[4 * q + 1] % 2) + 2) % 2, ((h->entry_sites[4 * q + 2] % 2) + 2) % 2,
                    ((h->entry_sites[4 * q + 3] % 2) + 2) % 2});
      }
  }
  std::vector<std::array<int, 3>> same;
  for (int b = 0; b < nb; ++b)
    for (auto const &o : offs[b])
      if (o[0] == b) same.push_back({o[1], o[2], o[3]});
  constexpr long kMaxClassSites = 4096;
  struct Cand {
    int score;
    long n;
    int kind;  // 0: residues (p0, p1, p2); 1: linear (M, a, c) = (p0, p1, p2)
    int p0, p1, p2;
  };
  std::vector<Cand> cands;
  for (int m1 = 1; m1 <= L[0]; ++m1) {
    if (L[0] % m1) continue;
    for (int m2 = 1; m2 <= L[1]; ++m2) {
      if (L[1] % m2) continue;
      for (int m3 = 1; m3 <= L[2]; ++m3) {
        if (L[2] % m3) continue;
        const long n = (long)m1 * m2 * m3;
        if (nc / n > kMaxClassSites) continue;
        bool ok = true;
        for (auto const &d : same)
          if (d[0] % m1 == 0 && d[1] % m2 == 0 && d[2] % m3 == 0) {
            ok = false;
            break;
          }
        if (!ok) continue;
        int score = 0;
        for (auto const &v : par)
          score += ((m1 % 2 || v[0] == 0) && (m2 % 2 || v[1] == 0) && (m3 % 2 || v[2] == 0)) ? 1 : 0;
        cands.push_back({score, n, 0, m1, m2, m3});
      }
    }
  }
  for (int M = 2; M <= L[0]; ++M) {
    if (L[0] % M || nc / M > kMaxClassSites) continue;
    for (int a = 0; a < M; ++a) {
      if (((long)a * L[1]) % M) continue;
      for (int c = 0; c < M; ++c) {
        if (((long)c * L[2]) % M) continue;
        bool ok = true;
        for (auto const &d : same)
          if ((((long)d[0] + (long)a * d[1] + (long)c * d[2]) % M + M) % M == 0) {
            ok = false;
            break;
          }
        if (!ok) continue;
        int score = 0;
        for (auto const &v : par) score += (M % 2 || (v[0] + a * v[1] + c * v[2]) % 2 == 0) ? 1 : 0;
        cands.push_back({score, (long)M, 1, M, a, c});
      }
    }
  }
  int best_kind = -1, bm[3] = {0, 0, 0};
  {
    bool any_mix = false;
    for (auto const &c : cands) any_mix = any_mix || c.score > 0;
    long n_min = -1;
    for (auto const &c : cands)
      if ((c.score > 0 || !any_mix) && (n_min < 0 || c.n < n_min)) n_min = c.n;
    const int full = (int)par.size();
    const Cand *best = nullptr;
    bool best_full = false;
    for (auto const &c : cands) {
      if (!(c.score > 0 || !any_mix)) continue;
      const bool is_full = c.score == full && c.n * 4 <= n_min * 5;
      // full-score colourings within 25 % of the fewest classes come first
      if (!best || (is_full && !best_full) ||
          (is_full == best_full && (c.n < best->n || (c.n == best->n && c.score > best->score)))) {
        best = &c;
        best_full = is_full;
      }
    }
    if (best) {
      best_kind = best->kind;
      bm[0] = best->p0;
      bm[1] = best->p1;
      bm[2] = best->p2;
    }
  }
  cb.assign(1, 0);
  cs.clear();
  if (best_kind >= 0) {
    const long n_lab = best_kind == 0 ? (long)bm[0] * bm[1] * bm[2] : bm[0];
    std::vector<long> lab(nc);
    for (long cell = 0; cell < nc; ++cell) {
      const int i = (int)(cell % L[0]), j = (int)((cell / L[0]) % L[1]), k = (int)(cell / ((long)L[0] * L[1]));
      lab[cell] = best_kind == 0 ? (i % bm[0]) + (long)bm[0] * ((j % bm[1]) + (long)bm[1] * (k % bm[2]))
                                 : ((long)i + (long)bm[1] * j + (long)bm[2] * k) % bm[0];
    }
    for (int b = 0; b < nb; ++b) {
      if (h->n_occ[b] < 2) continue;
      for (long r = 0; r < n_lab; ++r) {
        for (long cell = 0; cell < nc; ++cell)
          if (lab[cell] == r) cs.push_back((int)(b * nc + cell));
        cb.push_back((int)cs.size());
      }
    }
  } else {
    std::vector<std::set<long>> nbr(N);
    for (long l = 0; l < N; ++l) {
      int b = (int)(l / nc);
      if (h->n_occ[b] < 2) continue;
      long c = l % nc;
      int i = (int)(c % L[0]), j = (int)((c / L[0]) % L[1]), k = (int)(c / ((long)L[0] * L[1]));
      for (auto const &o : offs[b]) {
        long t = lin(o[0], i + o[1], j + o[2], k + o[3]);
        if (t != l) {
          nbr[l].insert(t);
          nbr[t].insert(l);
        }
      }
    }
    std::vector<int> color(N, -1);
    int maxc = -1;
    for (long l = 0; l < N; ++l) {
      if (h->n_occ[l / nc] < 2) continue;
      std::set<int> used;
      for (long t : nbr[l]) used.insert(color[t]);
      int c = 0;
      while (used.count(c)) ++c;
      color[l] = c;
      maxc = std::max(maxc, c);
    }
    for (int b = 0; b < nb; ++b)
      for (int c = 0; c <= maxc; ++c) {
        size_t before = cs.size();
        for (long l = b * nc; l < (b + 1) * nc; ++l)
          if (color[l] == c) cs.push_back((int)l);
        if (cs.size() > before) cb.push_back((int)cs.size());
      }
  }
  // pair phases
  pb.assign(1, 0);
  ps.clear();
  std::vector<int> var;
  for (int b = 0; b < nb; ++b)
    if (h->n_occ[b] >= 2) var.push_back(b);
  uint64_t st = seed ^ 0xC0FFEEull;
  int tries = 0;
  std::vector<char> blocked(N);
  std::vector<long> order(nc);
  while ((int)pb.size() - 1 < n_pair_req && tries < 64 * std::max(1, n_pair_req)) {
    ++tries;
    int b = var[splitmix64(st) % var.size()];
    int b2 = var[splitmix64(st) % var.size()];
    int v[3] = {(int)(splitmix64(st) % L[0]), (int)(splitmix64(st) % L[1]), (int)(splitmix64(st) % L[2])};
    if (h->n_occ[b2] != h->n_occ[b] || h->asym[b2] != h->asym[b]) continue;
    if (b2 == b && v[0] == 0 && v[1] == 0 && v[2] == 0) continue;
    bool bad = false;
    for (auto const &o : offs[b])
      if (o[0] == b2 && ((o[1] - v[0]) % L[0]) == 0 && ((o[2] - v[1]) % L[1]) == 0 && ((o[3] - v[2]) % L[2]) == 0) {
        bad = true;
        break;
      }
    if (bad) continue;
    std::fill(blocked.begin(), blocked.end(), 0);
    std::iota(order.begin(), order.end(), 0);
    for (long i = nc - 1; i > 0; --i) std::swap(order[i], order[splitmix64(st) % (uint64_t)(i + 1)]);
    size_t before = ps.size();
    for (long c : order) {
      int i = (int)(c % L[0]), j = (int)((c / L[0]) % L[1]), k = (int)(c / ((long)L[0] * L[1]));
      long s = lin(b, i, j, k), t = lin(b2, i + v[0], j + v[1], k + v[2]);
      if (blocked[s] || blocked[t]) continue;
      ps.push_back((int)s);
      ps.push_back((int)t);
      blocked[s] = blocked[t] = 1;
      for (long x : {s, t}) {
        const int bx = (int)(x / nc);
        long cx = x % nc;
        int ii = (int)(cx % L[0]), jj = (int)((cx / L[0]) % L[1]), kk = (int)(cx / ((long)L[0] * L[1]));
        for (auto const &o : offs[bx]) blocked[lin(o[0], ii + o[1], jj + o[2], kk + o[3])] = 1;
      }
    }
    if (ps.size() > before) pb.push_back((int)(ps.size() / 2));
  }
  return CLEXMC_OK;
}

// Row pitch of the padded int8 image that minimises modelled LDS bank
// conflicts of the neighbour gathers: for the first four waves (64
// consecutive class sites) of every class and every distinct neighbour
// offset of the class's sublattice, each half wave (32 lanes) costs as many
// cycles as the largest number of distinct dwords that fall into one of 32
// four-byte banks.  (This model ranks measured pitches of the FCC 24^3 image
// correctly: 28, 30 and 38 bytes slow, 32 fast; a 64-bank full-wave model
// does not.)  Candidates add 0 .. max(4, P1/4) bytes per row (at most 32)
// and keep the image within 25 % of its unpadded size; ties keep the smaller
// pitch.  Images above 64 KiB (which may need the packed format to fit in
// LDS) keep the unpadded pitch.
static int choose_row_pitch(const clexmc_handle *h, int L1, int L2, int L3, int hw, int P2, int P3,
                            const clexmc_supercell_desc *sc) {
  const int P1_min = L1 + 2 * hw;
  if ((long long)P1_min * P2 * P3 * h->n_sublat > 65536) return P1_min;
  const int n_cells = L1 * L2 * L3;
  // distinct neighbour offsets (sublattice, di, dj, dk) per home sublattice
  std::vector<std::vector<std::array<int, 4>>> offs(h->n_sublat);
  for (int b = 0; b < h->n_sublat; ++b) {
    std::set<std::array<int, 4>> u;
    for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e)
      for (int q = h->entry_site_begin[e]; q < h->entry_site_begin[e + 1]; ++q)
        u.insert({h->entry_sites[4 * q], h->entry_sites[4 * q + 1], h->entry_sites[4 * q + 2], h->entry_sites[4 * q + 3]});
    offs[b].assign(u.begin(), u.end());
  }
  const int max_extra = std::min(32, std::max(4, P1_min / 4));
  int best_p1 = P1_min;
  long long best_cost = -1;
  std::vector<long long> dw(64);
  std::vector<int> bank_count(32);
  for (int extra = 0; extra <= max_extra; ++extra) {
    const int P1 = P1_min + extra;
    if ((long long)P1 * 4 > (long long)P1_min * 5) break;
    const long long P12 = (long long)P1 * P2, S = P12 * P3;
    long long cost = 0;
    for (int c = 0; c < sc->n_classes; ++c) {
      const int t0 = sc->class_begin[c], t1 = std::min(sc->class_begin[c + 1], t0 + 256);
      for (int w0 = t0; w0 < t1; w0 += 64) {
        const int nl = std::min(64, t1 - w0);
        const int l0 = sc->class_sites[w0];
        const int b = l0 / n_cells;
        for (const auto &o : offs[b]) {
          for (int t = 0; t < nl; ++t) {
            const int l = sc->class_sites[w0 + t];
            const int cc = l % n_cells;
            const int i = cc % L1, j = (cc / L1) % L2, k = cc / (L1 * L2);
            const long long a =
                o[0] * S + (long long)(k + o[3] + hw) * P12 + (long long)(j + o[2] + hw) * P1 + (i + o[1] + hw);
            dw[t] = a >> 2;
          }
          // per half wave: largest number of distinct dwords in one bank
          for (int h0 = 0; h0 < nl; h0 += 32) {
            const int h1 = std::min(nl, h0 + 32);
            std::sort(dw.begin() + h0, dw.begin() + h1);
            std::fill(bank_count.begin(), bank_count.end(), 0);
            int worst = 0;
            for (int t = h0; t < h1; ++t) {
              if (t > h0 && dw[t] == dw[t - 1]) continue;
              worst = std::max(worst, ++bank_count[(int)(dw[t] & 31)]);
            }
            cost += worst;
          }
        }
      }
    }
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best_p1 = P1;
    }
  }
  return best_p1;
}

clexmc_status clexmc_add_supercell(clexmc_handle *h, const clexmc_supercell_desc *sc, int *id_out) {
  if (!h || !sc || !id_out) return CLEXMC_ERR_ARG;
  HIP_TRY(hipSetDevice(h->device));
  const int L1 = sc->L[0], L2 = sc->L[1], L3 = sc->L[2];
  ARG_CHECK(L1 >= 1 && L2 >= 1 && L3 >= 1 && L1 <= 4096 && L2 <= 4096 && L3 <= 4096, "bad supercell dims");
  HostSupercell hs;
  DevSupercell &d = hs.dev;
  const int hw = h->halo;
  d.L1 = L1;
  d.L2 = L2;
  d.L3 = L3;
  d.h = hw;
  d.P2 = L2 + 2 * hw;
  d.P3 = L3 + 2 * hw;
  d.n_cells = L1 * L2 * L3;
  d.n_sites = d.n_cells * h->n_sublat;
  const int N = d.n_sites;
  // classes (given, or computed here when n_classes == 0)
  clexmc_supercell_desc auto_desc;
  std::vector<int> a_cb, a_cs, a_pb, a_ps;
  if (sc->n_classes == 0) {
    const int L[3] = {L1, L2, L3};
    clexmc_status pst = auto_partition(h, L, sc->n_pair_phases, 0x5eedull + h->scs.size(), a_cb, a_cs, a_pb, a_ps);
    if (pst != CLEXMC_OK) return pst;
    auto_desc = *sc;
    auto_desc.n_classes = (int)a_cb.size() - 1;
    auto_desc.class_begin = a_cb.data();
    auto_desc.class_sites = a_cs.data();
    auto_desc.n_pair_phases = (int)a_pb.size() - 1;
    auto_desc.pair_phase_begin = a_pb.data();
    auto_desc.pair_sites = a_ps.data();
    sc = &auto_desc;
  }
  // row pitch of the padded image: L1 + 2h, plus unused bytes per row when
  // that spreads the class sites' neighbour gathers over more LDS banks
  d.P1 = L1 + 2 * hw + (h->row_pitch > 0 ? h->row_pitch : 0);
  if (h->row_pitch < 0) d.P1 = choose_row_pitch(h, L1, L2, L3, hw, d.P2, d.P3, sc);
  const long long S = (long long)d.P1 * d.P2 * d.P3;
  const long long pad = S * h->n_sublat;
  ARG_CHECK(pad < (1LL << 30), "padded image too large");
  d.S = (int)S;
  d.pad_bytes = (int)pad;
  auto pad_addr = [&](int b, int i, int j, int k) {
    return (int)(b * S + (long long)(k + hw) * d.P1 * d.P2 + (long long)(j + hw) * d.P1 + (i + hw));
  };
  auto decode = [&](int l, int &b, int &i, int &j, int &k) {
    b = l / d.n_cells;
    int c = l - b * d.n_cells;
    i = c % L1;
    j = (c / L1) % L2;
    k = c / (L1 * L2);
  };
  // all padded images of site l other than its main image
  auto images = [&](int l, std::vector<int> &out) {
    int b, i, j, k;
    decode(l, b, i, j, k);
    out.clear();
    std::vector<int> pis, pjs, pks;
    for (int m = -(hw / L1 + 2); m <= hw / L1 + 2; ++m) {
      int p = i + hw + m * L1;
      if (p >= 0 && p < d.P1) pis.push_back(p);
    }
    for (int m = -(hw / L2 + 2); m <= hw / L2 + 2; ++m) {
      int p = j + hw + m * L2;
      if (p >= 0 && p < d.P2) pjs.push_back(p);
    }
    for (int m = -(hw / L3 + 2); m <= hw / L3 + 2; ++m) {
      int p = k + hw + m * L3;
      if (p >= 0 && p < d.P3) pks.push_back(p);
    }
    const int main = pad_addr(b, i, j, k);
    for (int pk : pks)
      for (int pj : pjs)
        for (int pi : pis) {
          int a = (int)(b * S + (long long)pk * d.P1 * d.P2 + (long long)pj * d.P1 + pi);
          if (a != main) out.push_back(a);
        }
  };
  ARG_CHECK(sc->n_classes >= 1, "no independent-set classes");
  std::vector<int> cb(sc->class_begin, sc->class_begin + sc->n_classes + 1);
  ARG_CHECK(cb[0] == 0, "class_begin[0] must be 0");
  std::vector<int> class_addr(std::max(1, cb.back())), class_lin(std::max(1, cb.back())),
      class_sub(sc->n_classes);
  std::vector<int> halo_begin(1, 0);
  std::vector<int2> halo;
  std::vector<int> halo_site(cb.back() + 1, 0);  // first halo entry of each class site
  std::vector<int> img;
  std::vector<char> seen(N, 0);
  int max_class = 1;
  for (int c = 0; c < sc->n_classes; ++c) {
    ARG_CHECK(cb[c + 1] >= cb[c], "class_begin not monotone");
    max_class = std::max(max_class, cb[c + 1] - cb[c]);
    int b0 = -1;
    for (int t = cb[c]; t < cb[c + 1]; ++t) {
      const int l = sc->class_sites[t];
      ARG_CHECK(l >= 0 && l < N, "class site out of range");
      ARG_CHECK(!seen[l], "site listed in two classes");
      seen[l] = 1;
      int b, i, j, k;
      decode(l, b, i, j, k);
      ARG_CHECK(h->n_occ[b] >= 2, "class contains a fixed site");
      if (b0 < 0) b0 = b;
      ARG_CHECK(b == b0, "a class must hold sites of one sublattice");
      class_addr[t] = pad_addr(b, i, j, k);
      class_lin[t] = l;
      images(l, img);
      halo_site[t] = (int)halo.size();
      for (int a : img) halo.push_back(make_int2(class_addr[t], a));
    }
    class_sub[c] = b0 < 0 ? 0 : b0;
    halo_begin.push_back((int)halo.size());
  }
  halo_site[cb.back()] = (int)halo.size();
  for (int l = 0; l < N; ++l) {
    int b = l / d.n_cells;
    if (h->n_occ[b] >= 2) ARG_CHECK(seen[l], "a variable site is in no class");
  }
  // entry offsets
  std::vector<int> entry_off(std::max<size_t>(1, h->entry_sites.size() / 4));
  for (size_t e = 0; e < h->entry_k.size(); ++e) {
    const int b = h->entry_home[e];
    for (int q = h->entry_site_begin[e]; q < h->entry_site_begin[e + 1]; ++q) {
      const int bq = h->entry_sites[4 * q], di = h->entry_sites[4 * q + 1], dj = h->entry_sites[4 * q + 2],
                dk = h->entry_sites[4 * q + 3];
      entry_off[q] = (int)((long long)(bq - b) * S + (long long)dk * d.P1 * d.P2 + (long long)dj * d.P1 + di);
    }
  }
  // bounds validation of every LDS neighbour access from every class site
  for (int c = 0; c < sc->n_classes; ++c)
    for (int t = cb[c]; t < cb[c + 1]; ++t) {
      const int b = class_sub[c];
      for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e)
        for (int q = h->entry_site_begin[e]; q < h->entry_site_begin[e + 1]; ++q) {
          long long a = (long long)class_addr[t] + entry_off[q];
          ARG_CHECK(a >= 0 && a < pad, "neighbour access outside the padded image");
        }
    }
  // binary site programs: per sublattice, per slot (orbit with >= 1 site), the
  // padded offsets of the other sites of every cluster containing the site,
  // padded to a multiple of 4 clusters with all-zero (self) offsets
  std::vector<BinSub> bin_sub(h->n_sublat);
  std::vector<int> bin_off;
  if (h->binary) {
    for (int b = 0; b < h->n_sublat; ++b) {
      BinSub &B = bin_sub[b];
      memset(&B, 0, sizeof(B));
      for (int j = 0; j < h->dsys.n_slots; ++j) {
        const int a = h->dsys.slot_orbit[j];
        const int ko = h->orbit_k[a] - 1;
        B.slot_k[j] = ko;
        B.slot_w[j] = (float)h->bin_w[a];
        B.slot_begin[j] = (int)bin_off.size();
        int n = 0;
        for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e) {
          if (h->entry_orbit[e] != a) continue;
          ARG_CHECK(h->entry_site_begin[e + 1] - h->entry_site_begin[e] == ko, "entry size mismatch");
          for (int q = h->entry_site_begin[e]; q < h->entry_site_begin[e + 1]; ++q) bin_off.push_back(entry_off[q]);
          n++;
        }
        B.slot_n[j] = n;
        if (ko == 0) {
          B.slot_len[j] = 0;
          B.slot_dummy[j] = 0;
          ARG_CHECK(n <= 1, "point orbit with several clusters per site");
          continue;
        }
        const int len = (n + 3) / 4 * 4;
        for (int e = n; e < len; ++e)
          for (int q = 0; q < ko; ++q) bin_off.push_back(0);
        B.slot_len[j] = len;
        B.slot_dummy[j] = len - n;
        ARG_CHECK(n <= 255, "too many clusters per site for the binary path");
      }
    }
  }
  // environment offsets of the learned proposal
  std::vector<int> env_off((size_t)h->n_sublat * kEnvMax, 0);
  for (int b = 0; b < h->n_sublat; ++b)
    for (int k = 0; k < h->n_env; ++k) {
      const int4 e = h->env_sites[(size_t)b * kEnvMax + k];
      env_off[(size_t)b * kEnvMax + k] =
          (int)((long long)(e.x - b) * S + (long long)e.w * d.P1 * d.P2 + (long long)e.z * d.P1 + e.y);
    }
  // general site programs: per cluster k, table offset, then (offset, radix)
  std::vector<GenSub> gen_sub(h->n_sublat);
  std::vector<int> gen_prog;
  for (int b = 0; b < h->n_sublat; ++b) {
    GenSub &G = gen_sub[b];
    G.begin = (int)gen_prog.size();
    G.n_entries = h->flower_begin[b + 1] - h->flower_begin[b];
    G.nf = std::max(0, h->n_occ[b] - 1);
    G.pad = 0;
    for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e) {
      const int k = h->entry_site_begin[e + 1] - h->entry_site_begin[e];
      gen_prog.push_back(k);
      gen_prog.push_back(h->entry_table[e]);
      for (int q = h->entry_site_begin[e]; q < h->entry_site_begin[e + 1]; ++q) {
        gen_prog.push_back(entry_off[q]);
        gen_prog.push_back(h->entry_mul[q]);
      }
    }
  }
  // pair phases (canonical cross-class mixing)
  std::vector<int> pair_begin(1, 0), pair_sub, pair_halo_begin(1, 0);
  std::vector<int2> pair_addr, pair_halo;
  for (int ph = 0; ph < sc->n_pair_phases; ++ph) {
    const int p0 = sc->pair_phase_begin[ph], p1 = sc->pair_phase_begin[ph + 1];
    ARG_CHECK(p1 >= p0, "pair_phase_begin not monotone");
    for (int t = p0; t < p1; ++t) {
      const int ls = sc->pair_sites[2 * t], lt = sc->pair_sites[2 * t + 1];
      ARG_CHECK(ls >= 0 && ls < N && lt >= 0 && lt < N && ls != lt, "bad pair site");
      int b, i, j, k, b2, i2, j2, k2;
      decode(ls, b, i, j, k);
      decode(lt, b2, i2, j2, k2);
      ARG_CHECK(h->n_occ[b] == h->n_occ[b2] && h->n_occ[b] >= 2 && h->asym[b] == h->asym[b2],
                "pair sites need one asymmetric unit");
      pair_addr.push_back(make_int2(pad_addr(b, i, j, k), pad_addr(b2, i2, j2, k2)));
      pair_sub.push_back(b | (b2 << 8));
      images(ls, img);
      for (int a : img) pair_halo.push_back(make_int2(pad_addr(b, i, j, k), a));
      images(lt, img);
      for (int a : img) pair_halo.push_back(make_int2(pad_addr(b2, i2, j2, k2), a));
    }
    pair_begin.push_back((int)pair_addr.size());
    pair_halo_begin.push_back((int)pair_halo.size());
  }
  d.n_classes = sc->n_classes;
  d.max_class = max_class;
  d.n_pair_phases = sc->n_pair_phases;
  auto &A = hs.allocs;
  HIP_TRY(upload(h, A, cb.data(), cb.size(), (int **)&d.class_begin));
  HIP_TRY(upload(h, A, class_sub.data(), class_sub.size(), (int **)&d.class_sublat));
  HIP_TRY(upload(h, A, class_addr.data(), class_addr.size(), (int **)&d.class_addr));
  HIP_TRY(upload(h, A, bin_sub.data(), bin_sub.size(), (BinSub **)&d.bin_sub));
  HIP_TRY(upload(h, A, bin_off.data(), bin_off.size(), (int **)&d.bin_off));
  hs.bin_sub = bin_sub;
  hs.bin_off = bin_off;
  hs.gen_sub = gen_sub;
  hs.gen_prog = gen_prog;
  HIP_TRY(upload(h, A, gen_sub.data(), gen_sub.size(), (GenSub **)&d.gen_sub));
  HIP_TRY(upload(h, A, gen_prog.data(), gen_prog.size(), (int **)&d.gen_prog));
  HIP_TRY(upload(h, A, env_off.data(), env_off.size(), (int **)&d.env_off));
  {
    // interaction neighbours per sublattice (self first), for rate updates of the N-fold way
    std::vector<int> nb_begin(1, 0);
    std::vector<int4> nb_cell;
    int nb_max = 0;
    for (int b = 0; b < h->n_sublat; ++b) {
      std::set<std::array<int, 4>> u;
      for (int e = h->flower_begin[b]; e < h->flower_begin[b + 1]; ++e)
        for (int q = h->entry_site_begin[e]; q < h->entry_site_begin[e + 1]; ++q)
          u.insert({h->entry_sites[4 * q], h->entry_sites[4 * q + 1], h->entry_sites[4 * q + 2],
                    h->entry_sites[4 * q + 3]});
      u.erase({b, 0, 0, 0});
      nb_cell.push_back(make_int4(b, 0, 0, 0));
      for (auto const &x : u) nb_cell.push_back(make_int4(x[0], x[1], x[2], x[3]));
      nb_begin.push_back((int)nb_cell.size());
      nb_max = std::max(nb_max, nb_begin[b + 1] - nb_begin[b]);
    }
    d.nb_max = nb_max;
    hs.nb_total = (int)nb_cell.size();
    // whether the periodic supercell keeps every list's sites distinct
    d.nb_distinct = 1;
    auto md = [](int x, int m) { return ((x % m) + m) % m; };
    for (int b = 0; b < h->n_sublat && d.nb_distinct; ++b)
      for (int p = nb_begin[b]; p < nb_begin[b + 1] && d.nb_distinct; ++p)
        for (int q = p + 1; q < nb_begin[b + 1]; ++q) {
          const int4 x = nb_cell[p], y = nb_cell[q];
          if (x.x == y.x && md(x.y - y.y, L1) == 0 && md(x.z - y.z, L2) == 0 && md(x.w - y.w, L3) == 0) {
            d.nb_distinct = 0;
            break;
          }
        }
    HIP_TRY(upload(h, A, nb_begin.data(), nb_begin.size(), (int **)&d.nb_begin));
    HIP_TRY(upload(h, A, nb_cell.data(), nb_cell.size(), (int4 **)&d.nb_cell));
    // slabs: reach along k of the clusters (largest |dk| of a flower site) and
    // per class the first list position with cell plane >= K; a class whose
    // sites are not in plane order leaves the supercell without slabs
    int kreach = 0;
    for (int q = 0; q < (int)nb_cell.size(); ++q) kreach = std::max(kreach, std::abs(nb_cell[q].w));
    std::vector<int> kidx((size_t)sc->n_classes * (L3 + 1));
    bool ordered = true;
    for (int c = 0; c < sc->n_classes && ordered; ++c) {
      int *kx = kidx.data() + (size_t)c * (L3 + 1);
      int t = cb[c];
      int prev = -1;
      for (int K = 0; K <= L3; ++K) {
        while (t < cb[c + 1]) {
          const int l = sc->class_sites[t];
          const int kt = (l % d.n_cells) / (L1 * L2);
          if (kt < prev) ordered = false;
          if (kt >= K) break;
          prev = kt;
          ++t;
        }
        kx[K] = t - cb[c];
      }
      if (t != cb[c + 1]) ordered = false;
    }
    d.kreach = ordered ? kreach : -1;
    d.class_kidx = nullptr;
    if (ordered) HIP_TRY(upload(h, A, kidx.data(), kidx.size(), (int **)&d.class_kidx));
  }
  HIP_TRY(upload(h, A, halo_begin.data(), halo_begin.size(), (int **)&d.halo_begin));
  HIP_TRY(upload(h, A, halo_site.data(), halo_site.size(), (int **)&d.halo_site));
  {
    std::vector<int> tpos(N, -1);
    for (int t = 0; t < cb.back(); ++t) tpos[class_lin[t]] = t;
    HIP_TRY(upload(h, A, tpos.data(), tpos.size(), (int **)&d.site_tpos));
  }
  HIP_TRY(upload(h, A, halo.data(), halo.size(), (int2 **)&d.halo));
  HIP_TRY(upload(h, A, entry_off.data(), entry_off.size(), (int **)&d.entry_off));
  HIP_TRY(upload(h, A, pair_begin.data(), pair_begin.size(), (int **)&d.pair_begin));
  HIP_TRY(upload(h, A, pair_addr.data(), pair_addr.size(), (int2 **)&d.pair_addr));
  HIP_TRY(upload(h, A, pair_sub.data(), pair_sub.size(), (int **)&d.pair_sublat));
  HIP_TRY(upload(h, A, pair_halo_begin.data(), pair_halo_begin.size(), (int **)&d.pair_halo_begin));
  HIP_TRY(upload(h, A, pair_halo.data(), pair_halo.size(), (int2 **)&d.pair_halo));
  {
    // register-cube eligibility: every class the cells of one residue
    // (x, y, z) mod 3, site t of a class in block t = bi + n1 (bj + n2 bk)
    const int n1 = L1 / 3, n2 = L2 / 3, n3 = L3 / 3;
    bool ok = h->binary && h->n_sublat == 1 && sc->n_classes == 27 && L1 % 3 == 0 && L2 % 3 == 0 &&
              L3 % 3 == 0 && n1 * n2 * n3 <= kMaxJitThreads;
    CubeSpec cs{};
    bool seen[27] = {};
    for (int c = 0; c < 27 && ok; ++c) {
      if (cb[c + 1] - cb[c] != n1 * n2 * n3) {
        ok = false;
        break;
      }
      const int l0 = sc->class_sites[cb[c]];
      const int r1 = l0 % L1 % 3, r2 = (l0 / L1) % L2 % 3, r3 = l0 / (L1 * L2) % 3;
      if (seen[r1 + 3 * (r2 + 3 * r3)]) ok = false;
      seen[r1 + 3 * (r2 + 3 * r3)] = true;
      cs.res[c][0] = r1;
      cs.res[c][1] = r2;
      cs.res[c][2] = r3;
      for (int t = 0; t < n1 * n2 * n3 && ok; ++t) {
        const int bi = t % n1, bj = (t / n1) % n2, bk = t / (n1 * n2);
        const int want = (3 * bi + r1) + L1 * ((3 * bj + r2) + L2 * (3 * bk + r3));
        if (sc->class_sites[cb[c] + t] != want) ok = false;
      }
    }
    if (ok) {
      cs.nb[0] = n1;
      cs.nb[1] = n2;
      cs.nb[2] = n3;
      cs.h = hw;
      cs.P1 = d.P1;
      cs.P12 = d.P1 * d.P2;
      hs.cube = cs;
    }
    hs.cube_ok = ok;
  }
  h->scs.push_back(std::move(hs));
  h->scs_dirty = true;
  *id_out = (int)h->scs.size() - 1;
  return CLEXMC_OK;
}

static clexmc_status sync_supercells(clexmc_handle *h) {
  if (!h->scs_dirty) return CLEXMC_OK;
  if (h->d_scs) hipFree(h->d_scs);
  h->d_scs = nullptr;
  std::vector<DevSupercell> v;
  for (auto &s : h->scs) v.push_back(s.dev);
  HIP_TRY(hipMalloc(&h->d_scs, sizeof(DevSupercell) * std::max<size_t>(1, v.size())));
  HIP_TRY(hipMemcpy(h->d_scs, v.data(), sizeof(DevSupercell) * v.size(), hipMemcpyHostToDevice));
  h->scs_dirty = false;
  return CLEXMC_OK;
}

static void graph_release(clexmc_handle *h);

static void free_chains(clexmc_handle *h) {
  graph_release(h);
  hipFree(h->d_group_map);
  h->d_group_map = nullptr;
  h->groups.clear();
  h->group_off.clear();
  hipFree(h->d_params);
  hipFree(h->d_state);
  hipFree(h->d_occ);
  hipFree(h->d_binC);
  hipFree(h->d_sampE);
  hipFree(h->d_sampN);
  hipFree(h->d_sample_count);
  hipFree(h->d_slab_rec);
  h->d_slab_rec = nullptr;
  h->slab_rec_chains = 0;
  h->d_params = nullptr;
  h->d_state = nullptr;
  h->d_occ = nullptr;
  h->d_binC = nullptr;
  h->d_sampE = nullptr;
  h->d_sampN = nullptr;
  h->d_sample_count = nullptr;
  h->sample_cap = 0;
  h->n_chains = 0;
}

static clexmc_status run_corr(clexmc_handle *h, int chain0, int n, double *out_host);

clexmc_status clexmc_load_chains(clexmc_handle *h, int n_chains, const int *supercell_id, const int8_t *occ,
                                 const double *temperature, const double *chem, const uint64_t *seeds,
                                 double kb) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(n_chains >= 1 && supercell_id && occ && temperature && seeds, "bad load_chains arguments");
  ARG_CHECK(kb > 0.0, "kb must be positive");
  HIP_TRY(hipSetDevice(h->device));
  clexmc_status st = sync_supercells(h);
  if (st != CLEXMC_OK) return st;
  free_chains(h);
  std::vector<ChainParams> params(n_chains);
  std::vector<ChainState> state(n_chains);
  h->chain_sc.assign(n_chains, 0);
  h->chain_off.assign(n_chains, 0);
  long long off = 0;
  for (int c = 0; c < n_chains; ++c) {
    const int id = supercell_id[c];
    ARG_CHECK(id >= 0 && id < (int)h->scs.size(), "bad supercell id");
    const DevSupercell &d = h->scs[id].dev;
    ChainParams &cp = params[c];
    ChainState &cs = state[c];
    memset(&cp, 0, sizeof(cp));
    memset(&cs, 0, sizeof(cs));
    cp.sc = id;
    cp.occ_offset = (int)off;
    h->chain_sc[c] = id;
    h->chain_off[c] = (int)off;
    ARG_CHECK(temperature[c] >= 0.0, "negative temperature");
    const double beta = temperature[c] > 0.0 ? 1.0 / (kb * temperature[c]) : std::numeric_limits<double>::infinity();
    cp.beta64 = beta;
    cp.beta = std::isfinite(beta) ? (float)beta : std::numeric_limits<float>::infinity();
    for (int b = 0; b < h->n_sublat; ++b)
      for (int o = 0; o < kMaxOcc; ++o)
        cp.chem[b * kMaxOcc + o] =
            (chem && o < h->n_occ_max) ? (float)chem[(long long)c * h->n_sublat * h->n_occ_max + b * h->n_occ_max + o] : 0.f;
    cp.key0 = (uint32_t)(seeds[c] & 0xffffffffu);
    cp.key1 = (uint32_t)(seeds[c] >> 32);
    // validate occupation and count occupants on variable sites
    for (int l = 0; l < d.n_sites; ++l) {
      const int b = l / d.n_cells;
      const int v = occ[off + l];
      ARG_CHECK(v >= 0 && v < h->n_occ[b], "occupant index out of range");
      if (h->n_occ[b] >= 2) cs.counts[v] += 1;
    }
    off += d.n_sites;
  }
  ARG_CHECK(off < (1LL << 31), "total occupation too large");
  h->occ_total = off;
  HIP_TRY(hipMalloc(&h->d_params, sizeof(ChainParams) * n_chains));
  HIP_TRY(hipMalloc(&h->d_state, sizeof(ChainState) * n_chains));
  HIP_TRY(hipMalloc(&h->d_occ, (size_t)off));
  HIP_TRY(hipMemcpy(h->d_occ, occ, (size_t)off, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc(&h->d_binC, sizeof(long long) * (size_t)n_chains * h->n_orbits));
  HIP_TRY(hipMalloc(&h->d_sample_count, sizeof(int) * n_chains));
  HIP_TRY(hipMemset(h->d_sample_count, 0, sizeof(int) * n_chains));
  h->samples_bound = 0;
  HIP_TRY(hipMemcpy(h->d_params, params.data(), sizeof(ChainParams) * n_chains, hipMemcpyHostToDevice));
  h->n_chains = n_chains;
  // initial energies (and exact binary counts) from full correlations
  std::vector<double> corr((size_t)n_chains * h->n_corr);
  st = run_corr(h, 0, n_chains, corr.data());
  if (st != CLEXMC_OK) return st;
  std::vector<long long> C((size_t)n_chains * h->n_orbits, 0);
  for (int c = 0; c < n_chains; ++c) {
    const double *cc = corr.data() + (size_t)c * h->n_corr;
    double E = 0.0;
    for (int i = 0; i < h->n_corr; ++i) E += h->eci[i] * cc[i];
    state[c].E = E;
    if (h->binary) {
      double Eb = 0.0;
      for (int a = 0; a < h->n_orbits; ++a) {
        const long long v = llround(cc[h->orbit_func_begin[a]] * h->orbit_mult[a]);
        C[(size_t)c * h->n_orbits + a] = v;
        Eb += h->bin_w[a] * (double)v;
      }
      state[c].E = Eb;
    }
  }
  HIP_TRY(hipMemcpy(h->d_binC, C.data(), sizeof(long long) * C.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_state, state.data(), sizeof(ChainState) * n_chains, hipMemcpyHostToDevice));
  // chains grouped by supercell, in order of first appearance
  {
    std::vector<int> gid(h->scs.size(), -1);
    for (int c = 0; c < n_chains; ++c) {
      int &g = gid[h->chain_sc[c]];
      if (g < 0) {
        g = (int)h->groups.size();
        h->groups.emplace_back();
      }
      h->groups[g].push_back(c);
    }
    std::vector<int> flat;
    for (auto &g : h->groups) {
      h->group_off.push_back((int)flat.size());
      flat.insert(flat.end(), g.begin(), g.end());
    }
    HIP_TRY(hipMalloc(&h->d_group_map, sizeof(int) * flat.size()));
    HIP_TRY(hipMemcpy(h->d_group_map, flat.data(), sizeof(int) * flat.size(), hipMemcpyHostToDevice));
  }
  return CLEXMC_OK;
}

static clexmc_status run_corr(clexmc_handle *h, int chain0, int n, double *out_host) {
  double *d_out = nullptr;
  HIP_TRY(hipMalloc(&d_out, sizeof(double) * h->n_corr * (size_t)n));
  CorrArgs a;
  a.sys = h->d_sys;
  a.scs = h->d_scs;
  a.params = h->d_params;
  a.occ = h->d_occ;
  a.chain0 = chain0;
  a.n_corr = h->n_corr;
  a.out = d_out;
  hipError_t e = launch_corr(a, h->n_corr, n, h->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out_host, d_out, sizeof(double) * h->n_corr * (size_t)n, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(d_out);
  if (e != hipSuccess) {
    h->err = std::string("correlations: ") + hipGetErrorString(e);
    return CLEXMC_ERR_HIP;
  }
  return CLEXMC_OK;
}

clexmc_status clexmc_correlations(clexmc_handle *h, int chain, double *corr_out) {
  if (!h || !corr_out) return CLEXMC_ERR_ARG;
  ARG_CHECK(chain >= 0 && chain < h->n_chains, "chain out of range");
  HIP_TRY(hipSetDevice(h->device));
  return run_corr(h, chain, 1, corr_out);
}

// Proposal arguments of the delta-energy entry points, checked on the host.
static clexmc_status check_proposals(clexmc_handle *h, int n_prop, const int32_t *chain, const int32_t *sites,
                                     const int32_t *new_occ, int sites_per_prop) {
  ARG_CHECK(chain && sites && new_occ, "null argument");
  ARG_CHECK(sites_per_prop >= 1 && sites_per_prop <= kMaxSitesPerProp, "sites_per_prop out of range (1..4)");
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  for (int p = 0; p < n_prop; ++p) {
    ARG_CHECK(chain[p] >= 0 && chain[p] < h->n_chains, "proposal chain out of range");
    const DevSupercell &d = h->scs[h->chain_sc[chain[p]]].dev;
    for (int i = 0; i < sites_per_prop; ++i) {
      const int l = sites[p * sites_per_prop + i];
      ARG_CHECK(l >= 0 && l < d.n_sites, "proposal site out of range");
      const int b = l / d.n_cells;
      const int v = new_occ[p * sites_per_prop + i];
      ARG_CHECK(v >= 0 && v < h->n_occ[b], "proposal occupant out of range");
    }
  }
  return CLEXMC_OK;
}

// Copies the proposals into the device buffers given, evaluates them and
// copies dE back (the handle's stream, synchronised before returning).
static hipError_t delta_batch_staged(clexmc_handle *h, int n_prop, const int32_t *chain, const int32_t *sites,
                                     const int32_t *new_occ, int sites_per_prop, int include_chem, float *dE_out,
                                     int *d_chain, int *d_sites, int *d_new, float *d_dE) {
  const size_t ns = (size_t)n_prop * sites_per_prop;
  hipError_t e = hipMemcpyAsync(d_chain, chain, sizeof(int) * n_prop, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_sites, sites, sizeof(int) * ns, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_new, new_occ, sizeof(int) * ns, hipMemcpyHostToDevice, h->stream);
  if (e == hipSuccess) {
    DeltaArgs a;
    a.sys = h->d_sys;
    a.scs = h->d_scs;
    a.params = h->d_params;
    a.occ = h->d_occ;
    a.n_prop = n_prop;
    a.sites_per_prop = sites_per_prop;
    a.include_chem = include_chem;
    a.chain = d_chain;
    a.sites = d_sites;
    a.new_occ = d_new;
    a.dE = d_dE;
    e = launch_delta_batch(a, h->stream);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(dE_out, d_dE, sizeof(float) * n_prop, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  return e;
}

clexmc_status clexmc_delta_energy_batch(clexmc_handle *h, int n_prop, const int32_t *chain, const int32_t *sites,
                                        const int32_t *new_occ, int sites_per_prop, int include_chem, float *dE_out) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(n_prop >= 0, "n_prop < 0");
  if (n_prop == 0) return CLEXMC_OK;
  ARG_CHECK(dE_out, "null argument");
  const clexmc_status st = check_proposals(h, n_prop, chain, sites, new_occ, sites_per_prop);
  if (st != CLEXMC_OK) return st;
  HIP_TRY(hipSetDevice(h->device));
  const size_t ns = (size_t)n_prop * sites_per_prop;
  int *d_chain = nullptr, *d_sites = nullptr, *d_new = nullptr;
  float *d_dE = nullptr;
  HIP_TRY(hipMalloc(&d_chain, sizeof(int) * n_prop));
  HIP_TRY(hipMalloc(&d_sites, sizeof(int) * ns));
  HIP_TRY(hipMalloc(&d_new, sizeof(int) * ns));
  HIP_TRY(hipMalloc(&d_dE, sizeof(float) * n_prop));
  const hipError_t e = delta_batch_staged(h, n_prop, chain, sites, new_occ, sites_per_prop, include_chem, dE_out,
                                          d_chain, d_sites, d_new, d_dE);
  hipFree(d_chain);
  hipFree(d_sites);
  hipFree(d_new);
  hipFree(d_dE);
  if (e != hipSuccess) {
    h->err = std::string("delta_energy_batch: ") + hipGetErrorString(e);
    return CLEXMC_ERR_HIP;
  }
  return CLEXMC_OK;
}

clexmc_status clexmc_delta_energy_batch_noalloc(clexmc_handle *h, int n_prop, const int32_t *chain,
                                                const int32_t *sites, const int32_t *new_occ, int sites_per_prop,
                                                int include_chem, float *dE_out) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(n_prop >= 0, "n_prop < 0");
  if (n_prop == 0) return CLEXMC_OK;
  if (n_prop > kNoAllocMaxProposals) {
    h->err = "clexmc_delta_energy_batch_noalloc: " + std::to_string(n_prop) + " proposals exceed the maximum of " +
             std::to_string(kNoAllocMaxProposals) + " per call (use clexmc_delta_energy_batch)";
    return CLEXMC_ERR_ARG;
  }
  ARG_CHECK(dE_out, "null argument");
  const clexmc_status st = check_proposals(h, n_prop, chain, sites, new_occ, sites_per_prop);
  if (st != CLEXMC_OK) return st;
  ARG_CHECK(h->noalloc, "no staging array on this device");
  HIP_TRY(hipSetDevice(h->device));
  // fixed slices of the device's staging array, sized for the maximum call;
  // every slice read by the kernel is written by this call's copies first
  unsigned char *base = h->noalloc;
  int *d_chain = (int *)base;
  int *d_sites = (int *)(base + 4 * (size_t)kNoAllocMaxProposals);
  int *d_new = (int *)(base + (4 + 4 * kMaxSitesPerProp) * (size_t)kNoAllocMaxProposals);
  float *d_dE = (float *)(base + (4 + 8 * kMaxSitesPerProp) * (size_t)kNoAllocMaxProposals);
  const hipError_t e = delta_batch_staged(h, n_prop, chain, sites, new_occ, sites_per_prop, include_chem, dE_out,
                                          d_chain, d_sites, d_new, d_dE);
  if (e != hipSuccess) {
    h->err = std::string("delta_energy_batch_noalloc: ") + hipGetErrorString(e);
    return CLEXMC_ERR_HIP;
  }
  return CLEXMC_OK;
}

clexmc_status clexmc_set_block_size(clexmc_handle *h, int threads) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(threads == 0 || (threads >= 64 && threads <= kMaxJitThreads && threads % 64 == 0),
            "block size must be 0 or a multiple of 64 in [64, 1024]");
  h->block_size = threads;
  return CLEXMC_OK;
}

clexmc_status clexmc_set_nfold_block_size(clexmc_handle *h, int threads) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(threads == 0 || threads == 64 || threads == 128 || threads == 256,
            "N-fold block size must be 0 (automatic), 64, 128 or 256");
  h->nfold_threads = threads;
  return CLEXMC_OK;
}

clexmc_status clexmc_set_slabs(clexmc_handle *h, int slabs) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(slabs >= 0 && slabs <= 64, "slabs must be 0 (automatic), 1 (off) or 2..64");
  h->slabs = slabs;
  return CLEXMC_OK;
}

clexmc_status clexmc_last_run_slabs(clexmc_handle *h, int *slabs) {
  if (!h || !slabs) return CLEXMC_ERR_ARG;
  *slabs = h->last_slabs;
  return CLEXMC_OK;
}

clexmc_status clexmc_set_register_cube(clexmc_handle *h, int mode) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(mode == 0 || mode == 1, "register cube mode must be 0 (class loop) or 1 (where available)");
  h->cube = mode;
  return CLEXMC_OK;
}

clexmc_status clexmc_last_run_cube(clexmc_handle *h, int *cube) {
  if (!h || !cube) return CLEXMC_ERR_ARG;
  *cube = h->last_cube;
  return CLEXMC_OK;
}

clexmc_status clexmc_set_fused_launch(clexmc_handle *h, int mode) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(mode == 0 || mode == 1, "fused launch mode must be 0 (one launch per shape) or 1 (fused)");
  h->fuse = mode;
  return CLEXMC_OK;
}

clexmc_status clexmc_last_run_fused(clexmc_handle *h, int *fused) {
  if (!h || !fused) return CLEXMC_ERR_ARG;
  *fused = h->last_fused;
  return CLEXMC_OK;
}

static void free_mswaps(clexmc_handle *h) {
  hipFree(h->d_mswap_parts);
  hipFree(h->d_mswap_begin);
  hipFree(h->d_mswap_rev);
  h->d_mswap_parts = nullptr;
  h->d_mswap_begin = nullptr;
  h->d_mswap_rev = nullptr;
  h->n_mswaps = 0;
}

clexmc_status clexmc_set_multiswaps(clexmc_handle *h, int n_types, const int *type_begin, const int *parts,
                                    int per_pass) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(n_types >= 0 && n_types <= kMaxMswapTypes, "n_types out of range (0..256)");
  ARG_CHECK(per_pass >= 0 && per_pass <= (1 << 24), "multiswaps per pass out of range");
  HIP_TRY(hipSetDevice(h->device));
  if (n_types == 0) {
    free_mswaps(h);
    h->mswap_per_pass = 0;
    return CLEXMC_OK;
  }
  ARG_CHECK(type_begin && parts && type_begin[0] == 0, "type_begin[0] must be 0");
  // a part: (asymmetric unit, occupant from, occupant to, count)
  std::vector<int4> P;
  for (int t = 0; t < n_types; ++t) {
    const int n_parts = type_begin[t + 1] - type_begin[t];
    ARG_CHECK(n_parts >= 1 && n_parts <= kMaxMswapParts, "a multiswap type needs 1..4 parts");
    int sites = 0;
    for (int q = type_begin[t]; q < type_begin[t + 1]; ++q) {
      const int4 pt = make_int4(parts[4 * q], parts[4 * q + 1], parts[4 * q + 2], parts[4 * q + 3]);
      int n_occ_u = -1;
      for (int b = 0; b < h->n_sublat; ++b)
        if (h->asym[b] == pt.x && h->n_occ[b] >= 2) n_occ_u = h->n_occ[b];
      ARG_CHECK(pt.x >= 0 && pt.x < kMaxSublat && n_occ_u >= 2, "multiswap part: no mutable sites in that unit");
      ARG_CHECK(pt.y >= 0 && pt.z >= 0 && pt.y < n_occ_u && pt.z < n_occ_u && pt.y != pt.z,
                "multiswap part: occupants out of range or equal");
      ARG_CHECK(pt.w >= 1, "multiswap part: count < 1");
      for (int q2 = type_begin[t]; q2 < q; ++q2)
        ARG_CHECK(!(parts[4 * q2] == pt.x && parts[4 * q2 + 1] == pt.y && parts[4 * q2 + 2] == pt.z),
                  "multiswap type repeats a part (merge the counts)");
      sites += pt.w;
      P.push_back(pt);
    }
    ARG_CHECK(sites <= kMaxMswapSites, "a multiswap changes at most 8 sites");
  }
  // the type undoing each type: the same parts with from and to exchanged
  std::vector<int> rev(n_types, -1);
  for (int t = 0; t < n_types; ++t) {
    const int nt = type_begin[t + 1] - type_begin[t];
    for (int r = 0; r < n_types && rev[t] < 0; ++r) {
      if (type_begin[r + 1] - type_begin[r] != nt) continue;
      bool all = true;
      for (int q = type_begin[t]; q < type_begin[t + 1] && all; ++q) {
        bool found = false;
        for (int q2 = type_begin[r]; q2 < type_begin[r + 1] && !found; ++q2)
          found = P[q2].x == P[q].x && P[q2].y == P[q].z && P[q2].z == P[q].y && P[q2].w == P[q].w;
        all = found;
      }
      if (all) rev[t] = r;
    }
    ARG_CHECK(rev[t] >= 0, "the multiswap list must hold the type that undoes each type");
  }
  free_mswaps(h);
  HIP_TRY(hipMalloc(&h->d_mswap_parts, sizeof(int4) * P.size()));
  HIP_TRY(hipMalloc(&h->d_mswap_begin, sizeof(int) * (n_types + 1)));
  HIP_TRY(hipMalloc(&h->d_mswap_rev, sizeof(int) * n_types));
  HIP_TRY(hipMemcpy(h->d_mswap_parts, P.data(), sizeof(int4) * P.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_mswap_begin, type_begin, sizeof(int) * (n_types + 1), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_mswap_rev, rev.data(), sizeof(int) * n_types, hipMemcpyHostToDevice));
  h->n_mswaps = n_types;
  h->mswap_per_pass = per_pass;
  return CLEXMC_OK;
}

clexmc_status clexmc_set_global_swaps(clexmc_handle *h, int per_pass) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(per_pass >= 0 && per_pass <= 65536, "global swaps per pass out of range");
  h->global_swaps = per_pass;
  return CLEXMC_OK;
}

clexmc_status clexmc_set_lds_row_pitch(clexmc_handle *h, int extra) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(extra >= -1 && extra <= 64, "row pitch padding must be -1 (automatic) or 0..64");
  h->row_pitch = extra;
  return CLEXMC_OK;
}

static hipFunction_t jit_kernel(clexmc_handle *h, int sc_id, int kind, bool packed, int lanes, int threads);

// Chains whose N-fold run stopped because their total rate vanished: the
// flags are cleared and the run reported as failed (its blocks past the stop
// are zero).
static clexmc_status nfold_check_stalls(clexmc_handle *h) {
  std::vector<ChainState> st(h->n_chains);
  HIP_TRY(hipMemcpy(st.data(), h->d_state, sizeof(ChainState) * h->n_chains, hipMemcpyDeviceToHost));
  int first = -1, n_bad = 0;
  for (int c = 0; c < h->n_chains; ++c)
    if (st[c].flags & kChainFlagNfoldStalled) {
      if (first < 0) first = c;
      ++n_bad;
      st[c].flags &= ~kChainFlagNfoldStalled;
    }
  if (n_bad == 0) return CLEXMC_OK;
  HIP_TRY(hipMemcpy(h->d_state, st.data(), sizeof(ChainState) * h->n_chains, hipMemcpyHostToDevice));
  h->err = std::to_string(n_bad) + " chain(s), first " + std::to_string(first) +
           ", stopped: the total N-fold event rate fell below 1e-30 (every change has "
           "beta*dE above ~69; FP32 rates cannot resolve them); raise the temperature";
  return CLEXMC_ERR_STATE;
}

static clexmc_status grow_samples(clexmc_handle *h, long long n_new_samples);

// The N-fold run; blocks_out receives kNfoldBlockSums moments per block.
static clexmc_status nfold_run_impl(clexmc_handle *h, int n_events, int events_per_block, double *blocks_out,
                                    int n_time_samples = 0, double sample_dt = 0.0) {
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  ARG_CHECK(n_events >= 0 && events_per_block >= 1 && n_events % events_per_block == 0,
            "n_events must be a non-negative multiple of events_per_block");
  ARG_CHECK(blocks_out || n_events == 0, "blocks_out is NULL");
  if (n_events == 0) return CLEXMC_OK;
  HIP_TRY(hipSetDevice(h->device));
  int max_pad = 0, max_sites = 0, nb_max = 1, nb_total = 1;
  for (int c = 0; c < h->n_chains; ++c) {
    const HostSupercell &hs = h->scs[h->chain_sc[c]];
    const DevSupercell &d = hs.dev;
    max_pad = std::max(max_pad, d.pad_bytes);
    max_sites = std::max(max_sites, d.n_sites);
    nb_max = std::max(nb_max, d.nb_max);
    nb_total = std::max(nb_total, hs.nb_total);
  }
  ARG_CHECK(max_sites <= 262144, "the N-fold kernel takes supercells of at most 262,144 sites");
  ARG_CHECK(h->n_mswaps == 0, "the N-fold kernel's events are single-site changes: clear the multiswaps");
  const size_t lds = nfold_lds_bytes(max_pad, max_sites, nb_max, nb_total);
  // automatic block: the neighbour list's rates one per thread
  const int nf_threads = h->nfold_threads > 0 ? h->nfold_threads : std::min(256, (nb_max + 63) / 64 * 64);
  ARG_CHECK(lds <= 160 * 1024, "supercell too large for the N-fold kernel's LDS (image + one rate per site)");
  NfoldArgs na;
  memset(&na, 0, sizeof(na));
  na.a.sys = h->d_sys;
  na.a.scs = h->d_scs;
  na.a.params = h->d_params;
  na.a.state = h->d_state;
  na.a.occ = h->d_occ;
  na.a.binC = h->d_binC;
  na.a.n_orbits = h->n_orbits;
  na.n_events = n_events;
  na.events_per_block = events_per_block;
  na.n_blocks = n_events / events_per_block;
  std::vector<int> count0;
  if (n_time_samples > 0) {
    const clexmc_status gs = grow_samples(h, n_time_samples);
    if (gs != CLEXMC_OK) return gs;
    count0.resize(h->n_chains);
    HIP_TRY(hipMemcpy(count0.data(), h->d_sample_count, sizeof(int) * h->n_chains, hipMemcpyDeviceToHost));
    na.a.sampE = h->d_sampE;
    na.a.sampN = h->d_sampN;
    na.a.sample_count = h->d_sample_count;
    na.a.sample_cap = h->sample_cap;
    na.n_time_samples = n_time_samples;
    na.sample_dt = sample_dt;
  }
  const size_t nb = (size_t)h->n_chains * na.n_blocks * kNfoldBlockSums;
  double *d_blocks = nullptr;
  HIP_TRY(hipMalloc(&d_blocks, nb * sizeof(double)));
  na.blocks = d_blocks;
  // every chain on one supercell: the run-time specialised N-fold kernel
  // (constant site-program offsets), else the generic one
  hipFunction_t jf = nullptr;
  if (h->specialize) {
    bool one_shape = true;
    for (int c = 1; c < h->n_chains; ++c) one_shape = one_shape && h->chain_sc[c] == h->chain_sc[0];
    if (one_shape) {
      (void)jit_kernel(h, h->chain_sc[0], 1, false, 1, kAotMaxThreads);
      jf = h->scs[h->chain_sc[0]].jit_nfold;
    }
  }
  hipError_t e = hipMemsetAsync(d_blocks, 0, nb * sizeof(double), h->stream);
  h->last_specialized = 0;
  if (e == hipSuccess && jf) {
    NfoldArgs args = na;
    void *params[] = {(void *)&args};
    const hipError_t le =
        hipModuleLaunchKernel(jf, (unsigned)h->n_chains, 1, 1, (unsigned)nf_threads, 1, 1, (unsigned)lds,
                              h->stream, params, nullptr);
    if (le == hipSuccess) {
      h->last_specialized = 1;
    } else {
      (void)hipGetLastError();
      jf = nullptr;
    }
  }
  if (e == hipSuccess && !jf) e = launch_nfold(na, h->binary != 0, h->n_chains, nf_threads, lds, h->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(blocks_out, d_blocks, nb * sizeof(double), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  hipFree(d_blocks);
  HIP_TRY(e);
  const clexmc_status st = nfold_check_stalls(h);
  if (st != CLEXMC_OK || n_time_samples <= 0) return st;
  std::vector<int> count1(h->n_chains);
  HIP_TRY(hipMemcpy(count1.data(), h->d_sample_count, sizeof(int) * h->n_chains, hipMemcpyDeviceToHost));
  for (int c = 0; c < h->n_chains; ++c)
    if (count1[c] - count0[c] != n_time_samples) {
      h->err = "N-fold time sampling: chain " + std::to_string(c) + " took " + std::to_string(count1[c] - count0[c]) +
               " of " + std::to_string(n_time_samples) + " samples within " + std::to_string(n_events) + " events";
      return CLEXMC_ERR_STATE;
    }
  return CLEXMC_OK;
}

clexmc_status clexmc_nfold_run_time(clexmc_handle *h, int n_samples, double period, int max_events) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(n_samples >= 1 && period > 0.0 && max_events >= 1,
            "n_samples and max_events must be positive and the period positive");
  std::vector<double> blk((size_t)h->n_chains * kNfoldBlockSums);
  return nfold_run_impl(h, max_events, max_events, blk.data(), n_samples, period);
}

clexmc_status clexmc_nfold_run_moments(clexmc_handle *h, int n_events, int events_per_block, double *blocks_out) {
  if (!h) return CLEXMC_ERR_ARG;
  return nfold_run_impl(h, n_events, events_per_block, blocks_out);
}

clexmc_status clexmc_nfold_run(clexmc_handle *h, int n_events, int events_per_block, double *blocks_out) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(n_events >= 0 && events_per_block >= 1 && n_events % events_per_block == 0,
            "n_events must be a non-negative multiple of events_per_block");
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  const size_t n_rec = (size_t)h->n_chains * (size_t)(n_events / events_per_block);
  std::vector<double> m(n_rec * kNfoldBlockSums);
  const clexmc_status st = nfold_run_impl(h, n_events, events_per_block, m.data());
  if (st != CLEXMC_OK) return st;
  // (W, sum w E, sum w N1, sum w E^2, sum w N1^2, sum w E N1)
  for (size_t r = 0; r < n_rec; ++r) {
    const double *q = m.data() + r * kNfoldBlockSums;
    double *o = blocks_out + r * 6;
    o[0] = q[0];
    o[1] = q[1];
    o[2] = q[3 + 1];
    o[3] = q[2];
    o[4] = q[7 + 4];
    o[5] = q[17 + 1];
  }
  return CLEXMC_OK;
}

clexmc_status clexmc_set_specialization(clexmc_handle *h, int mode) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(mode == 0 || mode == 1, "specialization must be 0 (off) or 1 (when possible)");
  h->specialize = mode;
  return CLEXMC_OK;
}

clexmc_status clexmc_specialization_log(clexmc_handle *h, char *buf, int cap) {
  if (!h || !buf || cap < 1) return CLEXMC_ERR_ARG;
  const int n = std::min(cap - 1, (int)h->jit_log.size());
  std::memcpy(buf, h->jit_log.data(), (size_t)n);
  buf[n] = '\0';
  return CLEXMC_OK;
}

clexmc_status clexmc_last_run_info(clexmc_handle *h, int *specialized, int *packed, int *threads) {
  if (!h) return CLEXMC_ERR_ARG;
  if (specialized) *specialized = h->last_specialized;
  if (packed) *packed = h->last_packed ? 1 : 0;
  if (threads) *threads = h->last_threads;
  return CLEXMC_OK;
}

clexmc_status clexmc_last_run_lanes(clexmc_handle *h, int *lanes_per_site) {
  if (!h || !lanes_per_site) return CLEXMC_ERR_ARG;
  *lanes_per_site = h->last_lanes;
  return CLEXMC_OK;
}

clexmc_status clexmc_set_lanes_per_site(clexmc_handle *h, int lanes) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(lanes >= 0 && lanes <= 2, "lanes per site must be 0 (automatic), 1 or 2");
  h->lanes_per_site = lanes;
  return CLEXMC_OK;
}

clexmc_status clexmc_set_image_packing(clexmc_handle *h, int mode) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(mode >= 0 && mode <= 2, "packing must be 0 (automatic), 1 (int8) or 2 (2-bit)");
  h->packing = mode;
  return CLEXMC_OK;
}

// Launch plan of the Metropolis (or grouped-proposal) kernel for the loaded
// chains: block size, LDS image format and size, kernel arguments.
struct RunPlan {
  int kind = 0, threads = 64, weights = 0;
  bool learned = false, packed = false;
  size_t lds = 0;
  hipFunction_t jit = nullptr;  // specialised kernel (one supercell shape)
  int lanes = 1;                // lanes per site in the class sweeps (binary path: 1 or 2)
  int n_blocks = 0;             // chains of this launch
  int sc_id = 0;                // supercell of a single-shape launch
  bool one_shape = false;       // every chain of the launch on supercell sc_id
  int slabs = 1;                // workgroups per chain (slab decomposition along k)
  bool threads_auto = false;    // block size chosen by the planner (not requested)
  bool cube = false;            // jit is the register-cube kernel
  unsigned long long pass_start = 0;  // the chains' pass counter before the run
  RunArgs a;
  LearnArgs la;
};

// The specialised kernel for (supercell, kind, packing, lanes per site, block
// size above 512 or not), compiled on first use; nullptr when specialisation
// is off or unavailable.
static hipFunction_t jit_kernel(clexmc_handle *h, int sc_id, int kind, bool packed, int lanes, int threads) {
  HostSupercell &hs = h->scs[sc_id];
  const int k = kind < 0 ? 0 : (kind > 2 ? 2 : kind), p = packed ? 1 : 0, l = lanes == 2 ? 1 : 0,
            w = lanes * threads > kAotMaxThreads;
  int &state = hs.jit_state[k][p][l][w];
  hipFunction_t &fn = hs.jit_fn[k][p][l][w];
  if (state == 0) {
    state = -1;
    std::vector<int> variable(h->n_sublat);
    for (int b = 0; b < h->n_sublat; ++b) variable[b] = h->n_occ[b] >= 2;
    const std::string src = jit_source(h->n_sublat, variable, h->binary != 0, hs.bin_sub, hs.bin_off, h->n_slots,
                                       hs.gen_sub, hs.gen_prog, kind, packed, lanes, w ? kMaxJitThreads : kAotMaxThreads,
                                       hs.cube_ok ? &hs.cube : nullptr);
    hipModule_t mod = nullptr;
    const hipError_t je = jit_compile_load(src, &mod, &h->jit_log);
    if (je != hipSuccess) h->jit_log += std::string("\n[compile/load: ") + hipGetErrorString(je) + "]";
    if (je == hipSuccess) {
      if (hipModuleGetFunction(&fn, mod, jit_kernel_name(kind, packed, lanes)) == hipSuccess) {
        hs.jit_mods.push_back(mod);
        state = 1;
        if (k == 1 && p == 0 && l == 0 && hipModuleGetFunction(&hs.jit_nfold, mod, kJitNfoldName) != hipSuccess)
          hs.jit_nfold = nullptr;
        if (k == 1 && p == 0 && l == 0 && hs.cube_ok &&
            hipModuleGetFunction(&hs.jit_cube[w], mod, kJitCubeName) != hipSuccess)
          hs.jit_cube[w] = nullptr;
      } else {
        h->jit_log += "\n[kernel lookup failed]";
        hipModuleUnload(mod);
      }
    }
  }
  return state == 1 ? fn : nullptr;
}

// Sample storage for n_new_samples more samples per chain.  The count in use
// is the host's upper bound (no device read); capacity at least doubles when
// it grows and starts at kMinSampleCap samples per chain, so short runs
// after a warm-up do not reallocate.
static constexpr long long kMinSampleCap = 256;
static clexmc_status grow_samples(clexmc_handle *h, long long n_new_samples) {
  const long long need = h->samples_bound + n_new_samples + 2;
  h->samples_bound += n_new_samples + 1;
  ARG_CHECK(need < (1ll << 31), "sample storage beyond 2^31 samples per chain; clear the samples");
  if (need > h->sample_cap) {
    const int cap = (int)std::min<long long>((1ll << 31) - 1,
                                             std::max({need, 2ll * h->sample_cap, kMinSampleCap}));
    double *nE = nullptr;
    int *nN = nullptr;
    HIP_TRY(hipMalloc(&nE, sizeof(double) * (size_t)h->n_chains * cap));
    HIP_TRY(hipMalloc(&nN, sizeof(int) * (size_t)h->n_chains * cap * kMaxOcc));
    if (h->sample_cap > 0) {
      HIP_TRY(hipMemcpy2D(nE, sizeof(double) * cap, h->d_sampE, sizeof(double) * h->sample_cap,
                          sizeof(double) * h->sample_cap, h->n_chains, hipMemcpyDeviceToDevice));
      HIP_TRY(hipMemcpy2D(nN, sizeof(int) * kMaxOcc * cap, h->d_sampN, sizeof(int) * kMaxOcc * h->sample_cap,
                          sizeof(int) * kMaxOcc * h->sample_cap, h->n_chains, hipMemcpyDeviceToDevice));
    }
    hipFree(h->d_sampE);
    hipFree(h->d_sampN);
    h->d_sampE = nE;
    h->d_sampN = nN;
    h->sample_cap = cap;
  }
  return CLEXMC_OK;
}

// threads_override / packing_override: 0 = the handle's setting
// subset: chain indices of one supercell group (nullptr: all chains)
// fused_spec: the caller supplies a fused specialised kernel for several
// plans (no per-plan kernel is built; blocks above 512 threads are kept)
static clexmc_status plan_run(clexmc_handle *h, int kind, int n_passes, int sample_period, int canonical_rounds,
                              int total_passes, RunPlan &P, int threads_override = 0, int packing_override = 0,
                              const std::vector<int> *subset = nullptr, int lanes_override = 0,
                              bool fused_spec = false) {
  ARG_CHECK(canonical_rounds >= 0 && canonical_rounds <= 4096, "canonical_rounds out of range (0..4096)");
  int max_pad = 0, max_class = 1;
  const int n_sel = subset ? (int)subset->size() : h->n_chains;
  for (int i = 0; i < n_sel; ++i) {
    const int c = subset ? (*subset)[i] : i;
    const DevSupercell &d = h->scs[h->chain_sc[c]].dev;
    max_pad = std::max(max_pad, d.pad_bytes);
    max_class = std::max(max_class, d.max_class);
  }
  // one lane per class site up to 1024 threads; blocks above 512 threads need
  // the specialised kernel (the ahead-of-time kernels are bounded at 512)
  int threads = threads_override ? threads_override : h->block_size;
  P.threads_auto = threads == 0;
  if (threads == 0) threads = std::min(kMaxJitThreads, std::max(64, ((max_class + 63) / 64) * 64));
  const int packing = packing_override ? packing_override : h->packing;
  const int n_tab = h->binary ? 0 : h->n_flower_tables;
  const int max_lds = h->max_lds;
  // int8 image when it fits (or is forced), else 2 bits per site (<= 4 occupants)
  bool packed = packing == 2;
  size_t lds = metropolis_lds_bytes(max_pad, n_tab, h->binary != 0, h->n_func_max, max_class, packed);
  if (packing == 0 && (long long)lds > (long long)max_lds - 64 && h->n_occ_max <= 4) {
    packed = true;
    lds = metropolis_lds_bytes(max_pad, n_tab, h->binary != 0, h->n_func_max, max_class, packed);
  }
  ARG_CHECK(!packed || h->n_occ_max <= 4, "2-bit packed images need at most 4 occupants per site");
  P.learned = h->proposal != 0;
  if (P.learned) {
    if (!h->binary) {
      h->err = "the learned proposal is implemented for binary (+-1) cluster expansions";
      return CLEXMC_ERR_UNSUPPORTED;
    }
    packed = false;
    lds = learned_lds_bytes(max_pad);
  }
  if ((long long)lds > (long long)max_lds - 64) {
    h->err = "chain image does not fit in LDS (" + std::to_string(lds) + " > " + std::to_string(max_lds) + " bytes)";
    return CLEXMC_ERR_UNSUPPORTED;
  }
  clexmc_status st = grow_samples(h, (long long)total_passes / sample_period);
  if (st != CLEXMC_OK) return st;
  // semi-grand canonical runs with multiswaps: the fixed-composition
  // (canonical) sweep followed by the multiswap attempts of every pass
  const bool mswap_run = kind == 1 && h->n_mswaps > 0;
  if (mswap_run && P.learned) {
    h->err = "multiswaps run with the plain proposal only";
    return CLEXMC_ERR_UNSUPPORTED;
  }
  kind = mswap_run ? 2 : kind;
  P.kind = kind;
  P.lanes = 1;
  const int lanes_req = lanes_override ? lanes_override : h->lanes_per_site;
  if (h->binary && !P.learned && !mswap_run && lanes_req == 2 && 2 * threads <= 1024) P.lanes = 2;
  P.threads = threads;
  P.packed = packed;
  P.lds = lds;
  P.weights = h->proposal == 1 ? 1 : (h->proposal == 3 ? 2 : 0);
  RunArgs &a = P.a;
  a.sys = h->d_sys;
  a.scs = h->d_scs;
  a.params = h->d_params;
  a.state = h->d_state;
  a.occ = h->d_occ;
  a.binC = h->d_binC;
  a.sampE = h->d_sampE;
  a.sampN = h->d_sampN;
  a.sample_count = h->d_sample_count;
  a.sample_cap = h->sample_cap;
  a.n_passes = n_passes;
  a.sample_period = sample_period;
  a.rounds = canonical_rounds <= 0 ? 0 : canonical_rounds;  // 0: automatic per independent set
  a.n_orbits = h->n_orbits;
  a.n_tab = n_tab;
  a.global_swaps = h->global_swaps;
  a.chain_map = nullptr;
  a.slabs = 1;
  a.slab_rec = h->d_slab_rec;
  a.n_mswaps = mswap_run ? h->n_mswaps : 0;
  a.mswap_per_pass = h->mswap_per_pass;
  a.mswap_parts = h->d_mswap_parts;
  a.mswap_begin = h->d_mswap_begin;
  a.mswap_rev = h->d_mswap_rev;
  P.n_blocks = n_sel;
  P.la.mlp = h->d_mlp;
  P.la.rounds_per_group = h->rounds_per_group;
  // specialised kernel when every chain uses the same supercell
  P.jit = nullptr;
  {
    const int c0 = subset ? (*subset)[0] : 0;
    bool one_shape = true;
    for (int i = 1; i < n_sel; ++i) one_shape = one_shape && h->chain_sc[subset ? (*subset)[i] : i] == h->chain_sc[c0];
    P.sc_id = h->chain_sc[c0];
    P.one_shape = one_shape;
    const size_t fail = h->scs[h->chain_sc[c0]].jit_lds_fail;
    if (h->specialize && !P.learned && one_shape && !fused_spec && (fail == 0 || lds < fail))
      P.jit = jit_kernel(h, h->chain_sc[c0], kind, packed, P.lanes, P.threads);
  }
  if (!P.jit && !fused_spec) P.threads = std::min(P.threads, kAotMaxThreads);
  // the register-cube sweep: the specialised semi-grand canonical kernel's
  // module carries it, one lane per 3x3x3 block
  P.cube = false;
  if (P.jit && P.one_shape && kind == 1 && !packed && P.lanes == 1 && h->cube) {
    const HostSupercell &hs = h->scs[P.sc_id];
    const hipFunction_t cf = hs.jit_cube[P.threads > kAotMaxThreads ? 1 : 0];
    if (hs.cube_ok && cf && P.threads == hs.cube.nb[0] * hs.cube.nb[1] * hs.cube.nb[2]) {
      P.jit = cf;
      P.cube = true;
    }
  }
  return CLEXMC_OK;
}

// The runtime refused a specialised launch (dynamic LDS of a module kernel
// above its limit): remember the LDS size and fall back to the ahead-of-time
// kernel, whose block size is bounded at 512 threads.
static void drop_jit(clexmc_handle *h, RunPlan &P, int sc_id) {
  (void)hipGetLastError();
  h->scs[sc_id].jit_lds_fail = P.lds;
  P.jit = nullptr;
  P.cube = false;
  P.threads = std::min(P.threads, kAotMaxThreads);
}

// attribute setting (not a stream operation) and the launch, separately, so
// that a launch can be recorded into a graph
static hipError_t plan_attributes(const RunPlan &P, clexmc_handle *h) {
  if (P.jit) return hipSuccess;
  if (P.lanes == 2)
    return launch_metropolis_split(P.a, P.kind, P.packed, P.n_blocks, P.threads * 2, P.lds, h->stream, false);
  if (P.learned) return launch_learned(P.a, P.la, P.kind, P.weights, h->n_chains, P.threads, P.lds, h->stream, false);
  return launch_metropolis(P.a, h->binary != 0, P.kind, P.packed, h->n_chains, P.threads, P.lds, h->stream, false);
}

static hipError_t plan_launch_one(const RunPlan &P, const RunArgs &args_in, clexmc_handle *h, hipStream_t st) {
  const int grid = P.n_blocks * std::max(1, args_in.slabs);
  if (P.jit) {
    RunArgs args = args_in;
    void *params[] = {(void *)&args};
    return hipModuleLaunchKernel(P.jit, (unsigned)grid, 1, 1, (unsigned)(P.threads * P.lanes), 1, 1, (unsigned)P.lds, st,
                                 params, nullptr);
  }
  if (P.lanes == 2) return launch_metropolis_split(args_in, P.kind, P.packed, grid, P.threads * 2, P.lds, st, true);
  if (P.learned) return launch_learned(args_in, P.la, P.kind, P.weights, P.n_blocks, P.threads, P.lds, h->stream, true);
  return launch_metropolis(args_in, h->binary != 0, P.kind, P.packed, grid, P.threads, P.lds, st, true);
}

// Slab plans run as a sequence of launches of at most kSlabEpochPasses passes
// (a launch may hold several sample points); each is followed by
// slab_post_kernel (samples, chain totals, pass and draw counters).  The slab
// boundaries move between launches.

static std::vector<int> slab_epochs(int n_passes) {
  std::vector<int> lens;
  int done = 0;
  while (done < n_passes) {
    const int len = std::min(n_passes - done, kSlabEpochPasses);
    lens.push_back(len);
    done += len;
  }
  return lens;
}

// `len` passes of the plan (its next launch of an epoch sequence)
static hipError_t plan_launch_window(const RunPlan &P, clexmc_handle *h, hipStream_t s, int len) {
  RunArgs a = P.a;
  a.n_passes = len;
  a.slabs = P.slabs;
  hipError_t e = plan_launch_one(P, a, h, s);
  if (e == hipSuccess && P.slabs > 1) e = launch_slab_post(a, P.n_blocks, s);
  return e;
}

static hipError_t plan_launch(const RunPlan &P, clexmc_handle *h, hipStream_t st = nullptr) {
  hipStream_t s = st ? st : h->stream;
  if (P.slabs <= 1) return plan_launch_one(P, P.a, h, s);
  for (int len : slab_epochs(P.a.n_passes)) {
    const hipError_t e = plan_launch_window(P, h, s, len);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Workgroups per chain.  Automatic: when a batch of large chains does not
// fill the device, chains are cut along k into slabs so that each workgroup
// holds about total_sites / compute_units sites (at least kMinSlabSites),
// with slabs at least three cluster reaches thick (two thirds or more of the
// planes updated per launch), at most kMaxSlabs.  Requested (>= 2): that many
// slabs where every slab keeps an updated plane.  Single-shape, plain
// Metropolis launches with one lane per site only.
static constexpr int kMinSlabSites = 8192;
static constexpr int kMaxSlabs = 16;

static hipError_t choose_slabs(clexmc_handle *h, const std::vector<RunPlan *> &plans) {
  const int n_cu = std::max(1, h->n_cu);
  long long total = 0;
  for (int c = 0; c < h->n_chains; ++c) total += h->scs[h->chain_sc[c]].dev.n_sites;
  const long long per_wg = std::max<long long>(kMinSlabSites, total / std::max(1, n_cu));
  for (RunPlan *P : plans) {
    P->slabs = 1;
    P->a.slabs = 1;
    if (h->slabs == 1 || P->learned || P->lanes != 1 || !P->one_shape || P->a.n_mswaps > 0) continue;
    const DevSupercell &d = h->scs[P->sc_id].dev;
    if (d.kreach < 0 || !d.class_kidx) continue;
    int S;
    if (h->slabs >= 2) {
      S = std::min(h->slabs, d.L3 / (d.kreach + 1));
    } else {
      const int cap = std::min(kMaxSlabs, d.kreach > 0 ? d.L3 / (3 * d.kreach) : d.L3);
      S = (int)std::min<long long>(cap, (d.n_sites + per_wg / 2) / per_wg);
    }
    if (S < 2) continue;
    if (P->cube) {
      // the register-cube kernel runs whole chains only: the class loop's
      // kernel for slabs
      P->jit = jit_kernel(h, P->sc_id, P->kind, P->packed, P->lanes, P->threads);
      P->cube = false;
    }
    if (!h->d_slab_rec || h->slab_rec_chains < h->n_chains) {
      if (h->d_slab_rec) (void)hipFree(h->d_slab_rec);
      h->d_slab_rec = nullptr;
      h->slab_rec_chains = 0;
      const hipError_t e =
          hipMalloc(&h->d_slab_rec, sizeof(SlabRecord) * (size_t)h->n_chains * kSlabMax * kSlabRecords);
      if (e != hipSuccess) return e;
      h->slab_rec_chains = h->n_chains;
    }
    P->a.slab_rec = h->d_slab_rec;
    P->slabs = S;
    // a slab's class ranges are a fraction of the class: 512-thread blocks
    // (measured faster than 1024 for slabs of the HCP and BCC workloads)
    if (P->threads_auto && P->threads > kAotMaxThreads) {
      P->threads = kAotMaxThreads;
      if (P->jit) P->jit = jit_kernel(h, P->sc_id, P->kind, P->packed, P->lanes, P->threads);
    }
  }
  return hipSuccess;
}

// The fused kernel for these group plans (one specialised field per supercell
// shape), compiled on first use; nullptr when it cannot be built.
static hipFunction_t fused_kernel(clexmc_handle *h, const std::vector<RunPlan> &plans, int kind, int threads) {
  std::string key = std::to_string(kind) + (threads > kAotMaxThreads ? "w" : "n");
  for (const RunPlan &P : plans) key += ":" + std::to_string(P.sc_id) + (P.packed ? "p" : "u");
  auto it = h->fused_fn.find(key);
  if (it != h->fused_fn.end()) return it->second;
  hipFunction_t fn = nullptr;
  std::vector<int> variable(h->n_sublat);
  for (int b = 0; b < h->n_sublat; ++b) variable[b] = h->n_occ[b] >= 2;
  std::vector<JitShape> shapes;
  for (const RunPlan &P : plans) {
    const HostSupercell &hs = h->scs[P.sc_id];
    shapes.push_back(JitShape{&hs.bin_sub, &hs.bin_off, &hs.gen_sub, &hs.gen_prog, P.packed});
  }
  const std::string src = jit_source_fused(h->n_sublat, variable, h->binary != 0, h->n_slots, shapes, kind, threads);
  hipModule_t mod = nullptr;
  const hipError_t je = jit_compile_load(src, &mod, &h->jit_log);
  if (je != hipSuccess) {
    h->jit_log += std::string("\n[fused compile/load: ") + hipGetErrorString(je) + "]";
  } else if (hipModuleGetFunction(&fn, mod, kJitFusedName) == hipSuccess) {
    h->fused_mods.push_back(mod);
  } else {
    fn = nullptr;
    h->jit_log += "\n[fused kernel lookup failed]";
    hipModuleUnload(mod);
  }
  h->fused_fn[key] = fn;
  return fn;
}

// Chains on different supercells in ONE launch of a fused specialised kernel:
// each workgroup dispatches on its group (the field of its supercell, its
// image format, chain map and slabs), so every group runs at once and the
// hardware spreads all of them over the compute units; slab groups get their
// slab_post_kernel after each launch.  Returns false (nothing launched) when
// the fused path does not apply or its kernel is unavailable.
static bool run_fused(clexmc_handle *h, int kind, int n_passes, int sample_period, int canonical_rounds,
                      clexmc_status *status) {
  const int ng = (int)h->groups.size();
  *status = CLEXMC_OK;
  if (ng < 2 || ng > kMaxFusedGroups || h->proposal != 0 || h->lanes_per_site == 2 || h->block_size != 0)
    return false;
  std::vector<RunPlan> plans(ng);
  for (int g = 0; g < ng; ++g) {
    const clexmc_status st = plan_run(h, kind, n_passes, sample_period, canonical_rounds, n_passes, plans[g], 0, 0,
                                      &h->groups[g], 0, true);
    if (st != CLEXMC_OK) {
      *status = st;
      return true;
    }
    if (plans[g].lanes != 1) return false;
    plans[g].a.chain_map = h->d_group_map + h->group_off[g];
  }
  std::vector<RunPlan *> pp;
  for (auto &P : plans) pp.push_back(&P);
  const hipError_t se = choose_slabs(h, pp);
  if (se != hipSuccess) {
    h->err = std::string("HIP: ") + hipGetErrorString(se);
    *status = CLEXMC_ERR_HIP;
    return true;
  }
  int threads = 64, max_slabs = 1;
  size_t lds = 0;
  bool any_packed = false;
  FusedGroups t{};
  int nb = 0;
  for (int g = 0; g < ng; ++g) {
    const RunPlan &P = plans[g];
    threads = std::max(threads, P.threads);
    lds = std::max(lds, P.lds);
    max_slabs = std::max(max_slabs, P.slabs);
    any_packed = any_packed || P.packed;
    t.block_begin[g] = nb;
    t.slabs[g] = std::max(1, P.slabs);
    t.chain_map[g] = P.a.chain_map;
    nb += P.n_blocks * std::max(1, P.slabs);
  }
  t.block_begin[ng] = nb;
  if (threads > kMaxJitThreads || (long long)lds > (long long)h->max_lds - 64) return false;
  const hipFunction_t fn = fused_kernel(h, plans, plans[0].kind, threads);
  if (!fn) return false;
  // every plan shares the final sample buffers and the slab records
  RunArgs a = plans[0].a;
  a.sampE = h->d_sampE;
  a.sampN = h->d_sampN;
  a.sample_cap = h->sample_cap;
  a.slab_rec = h->d_slab_rec;
  a.chain_map = nullptr;
  a.slabs = 1;
  std::vector<int> lens;
  if (max_slabs > 1)
    lens = slab_epochs(n_passes);
  else
    lens.push_back(n_passes);
#define FUSED_TRY(x)                                          \
  do {                                                        \
    const hipError_t e_ = (x);                                \
    if (e_ != hipSuccess) {                                   \
      h->err = std::string("HIP: ") + hipGetErrorString(e_); \
      *status = CLEXMC_ERR_HIP;                               \
      return true;                                            \
    }                                                         \
  } while (0)
  FUSED_TRY(hipEventRecord(h->ev0, h->stream));
  bool first = true;
  for (int len : lens) {
    a.n_passes = len;
    void *params[] = {(void *)&a, (void *)&t};
    const hipError_t le = hipModuleLaunchKernel(fn, (unsigned)nb, 1, 1, (unsigned)threads, 1, 1, (unsigned)lds,
                                                h->stream, params, nullptr);
    if (le != hipSuccess && first) {
      // refused before anything ran (for example its LDS size): not again
      (void)hipGetLastError();
      for (auto &kv : h->fused_fn)
        if (kv.second == fn) kv.second = nullptr;
      return false;
    }
    FUSED_TRY(le);
    first = false;
    for (int g = 0; g < ng; ++g) {
      if (plans[g].slabs <= 1) continue;
      RunArgs ag = a;
      ag.chain_map = t.chain_map[g];
      ag.slabs = plans[g].slabs;
      FUSED_TRY(launch_slab_post(ag, plans[g].n_blocks, h->stream));
    }
  }
  FUSED_TRY(hipEventRecord(h->ev1, h->stream));
  FUSED_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  FUSED_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
#undef FUSED_TRY
  h->last_ms = ms;
  h->last_specialized = 1;
  h->last_packed = any_packed;
  h->last_threads = threads;
  h->last_lanes = 1;
  h->last_slabs = max_slabs;
  h->last_fused = 1;
  h->last_cube = 0;
  return true;
}

// Chains on different supercells: one launch per supercell group, each with
// the kernel specialised to its shape and its own block size and LDS image,
// on side streams so that the groups run concurrently; the handle's stream
// waits for all of them.
static clexmc_status run_grouped(clexmc_handle *h, int kind, int n_passes, int sample_period, int canonical_rounds) {
  {
    clexmc_status st = CLEXMC_OK;
    if (h->fuse && run_fused(h, kind, n_passes, sample_period, canonical_rounds, &st)) return st;
  }
  h->last_fused = 0;
  h->last_cube = 0;
  const int ng = (int)h->groups.size();
  std::vector<RunPlan> plans(ng);
  for (int g = 0; g < ng; ++g) {
    clexmc_status st = plan_run(h, kind, n_passes, sample_period, canonical_rounds, n_passes, plans[g], 0, 0,
                                &h->groups[g]);
    if (st != CLEXMC_OK) return st;
    plans[g].a.chain_map = h->d_group_map + h->group_off[g];
  }
  // sample buffers may have grown while planning: every plan uses the final ones
  std::vector<RunPlan *> pp;
  for (auto &P : plans) {
    P.a.sampE = h->d_sampE;
    P.a.sampN = h->d_sampN;
    P.a.sample_cap = h->sample_cap;
    pp.push_back(&P);
  }
  HIP_TRY(choose_slabs(h, pp));
  for (auto &P : plans) HIP_TRY(plan_attributes(P, h));
  unsigned long long pass_start = 0;
  HIP_TRY(hipMemcpy(&pass_start, &h->d_state[0].pass_base, sizeof(pass_start), hipMemcpyDeviceToHost));
  int max_slabs = 1;
  for (auto &P : plans) {
    P.pass_start = pass_start;
    max_slabs = std::max(max_slabs, P.slabs);
  }
  // streams: as many as the runtime's hardware queues (GPU_MAX_HW_QUEUES,
  // 4 unless the process sets it; at most 8), so that streams do not share a
  // queue; groups of whole chains run one persistent launch, slab groups their
  // sequence of launches.  Groups go to streams longest first, each to the
  // stream with the least work so far (work = sites x passes of the group).
  int n_queues = 4;
  if (const char *q = std::getenv("GPU_MAX_HW_QUEUES")) n_queues = std::max(1, std::min(8, std::atoi(q)));
  const int ns = std::min(ng, n_queues);
  std::vector<int> stream_of(ng, 0);
  {
    std::vector<std::pair<long long, int>> work(ng);
    for (int g = 0; g < ng; ++g) {
      long long w = 0;
      for (int c : h->groups[g]) w += h->scs[h->chain_sc[c]].dev.n_sites;
      work[g] = {w, g};
    }
    std::sort(work.begin(), work.end(), [](auto const &a, auto const &b) { return a.first > b.first; });
    std::vector<long long> load(ns, 0);
    for (auto const &wg : work) {
      const int s_min = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      stream_of[wg.second] = s_min;
      load[s_min] += wg.first;
    }
  }
  while ((int)h->side.size() < ns) {
    hipStream_t s;
    hipEvent_t e;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    h->side.push_back(s);
    h->side_ev.push_back(e);
  }
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  for (int s = 0; s < ns; ++s) HIP_TRY(hipStreamWaitEvent(h->side[s], h->ev0, 0));
  bool all_spec = true, any_packed = false;
  for (int g = 0; g < ng; ++g) {
    RunPlan &P = plans[g];
    hipStream_t st = h->side[stream_of[g]];
    hipError_t le = plan_launch(P, h, st);
    if (le != hipSuccess && P.jit) {
      drop_jit(h, P, P.sc_id);
      HIP_TRY(plan_attributes(P, h));
      le = plan_launch(P, h, st);
    }
    HIP_TRY(le);
    all_spec = all_spec && P.jit != nullptr;
    any_packed = any_packed || P.packed;
  }
  for (int s = 0; s < ns; ++s) {
    HIP_TRY(hipEventRecord(h->side_ev[s], h->side[s]));
    HIP_TRY(hipStreamWaitEvent(h->stream, h->side_ev[s], 0));
  }
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->last_specialized = all_spec ? 1 : 0;
  h->last_packed = any_packed;
  h->last_threads = plans[0].threads;
  h->last_lanes = plans[0].lanes;
  h->last_slabs = max_slabs;
  return CLEXMC_OK;
}

// Multiswap runs: a chain whose pool counts disagreed with its image (a
// search found no site; never expected) reports an error; flags are cleared.
static clexmc_status mswap_check_flags(clexmc_handle *h) {
  std::vector<ChainState> st(h->n_chains);
  HIP_TRY(hipMemcpy(st.data(), h->d_state, sizeof(ChainState) * h->n_chains, hipMemcpyDeviceToHost));
  int first = -1;
  for (int c = 0; c < h->n_chains; ++c)
    if (st[c].flags & kChainFlagMswapCounts) {
      if (first < 0) first = c;
      st[c].flags &= ~kChainFlagMswapCounts;
    }
  if (first < 0) return CLEXMC_OK;
  HIP_TRY(hipMemcpy(h->d_state, st.data(), sizeof(ChainState) * h->n_chains, hipMemcpyHostToDevice));
  h->err = "multiswap site search found no candidate (chain " + std::to_string(first) +
           "): occupant counts per asymmetric unit disagree with the occupation";
  return CLEXMC_ERR_STATE;
}

static clexmc_status run_passes_impl(clexmc_handle *h, int kind, int n_passes, int sample_period,
                                     int canonical_rounds);

clexmc_status clexmc_run_passes(clexmc_handle *h, int kind, int n_passes, int sample_period, int canonical_rounds) {
  if (!h) return CLEXMC_ERR_ARG;
  const clexmc_status st = run_passes_impl(h, kind, n_passes, sample_period, canonical_rounds);
  if (st != CLEXMC_OK || kind != 1 || h->n_mswaps == 0 || n_passes == 0) return st;
  return mswap_check_flags(h);
}

static clexmc_status run_passes_impl(clexmc_handle *h, int kind, int n_passes, int sample_period,
                                     int canonical_rounds) {
  ARG_CHECK(kind == 0 || kind == 1, "kind must be 0 (canonical) or 1 (semi-grand canonical)");
  ARG_CHECK(n_passes >= 0 && sample_period >= 1, "bad n_passes / sample_period");
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  if (n_passes == 0) return CLEXMC_OK;
  HIP_TRY(hipSetDevice(h->device));
  if (h->specialize && h->proposal == 0 && h->groups.size() > 1)
    return run_grouped(h, kind, n_passes, sample_period, canonical_rounds);
  RunPlan P;
  clexmc_status st = plan_run(h, kind, n_passes, sample_period, canonical_rounds, n_passes, P);
  if (st != CLEXMC_OK) return st;
  HIP_TRY(choose_slabs(h, {&P}));
  if (P.slabs > 1) HIP_TRY(hipMemcpy(&P.pass_start, &h->d_state[0].pass_base, sizeof(P.pass_start), hipMemcpyDeviceToHost));
  h->last_packed = P.packed;
  h->last_specialized = P.jit ? 1 : 0;
  h->last_threads = P.threads;
  h->last_lanes = P.lanes;
  h->last_slabs = P.slabs;
  h->last_fused = 0;
  h->last_cube = P.cube && P.slabs <= 1 ? 1 : 0;
  HIP_TRY(plan_attributes(P, h));
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  hipError_t le = plan_launch(P, h);
  if (le != hipSuccess && P.jit) {
    drop_jit(h, P, h->chain_sc[0]);
    h->last_specialized = 0;
    h->last_threads = P.threads;
    HIP_TRY(plan_attributes(P, h));
    le = plan_launch(P, h);
  }
  HIP_TRY(le);
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  return CLEXMC_OK;
}

// ---------------------------------------------------------------------------
// HIP graph of a block of passes: captured once, replayed many times.  The
// kernel reads each chain's pass and draw counters from device memory, so
// replays continue the chains exactly like successive clexmc_run_passes calls
// of the same block length.  The sample buffer is sized for max_replays at
// capture time (the graph holds its address).
// ---------------------------------------------------------------------------
static void graph_release(clexmc_handle *h) {
  if (h->graph_exec) hipGraphExecDestroy(h->graph_exec);
  if (h->graph) hipGraphDestroy(h->graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
  h->graph_replays_left = 0;
}

clexmc_status clexmc_graph_capture(clexmc_handle *h, int kind, int passes_per_replay, int max_replays,
                                   int sample_period, int canonical_rounds) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(kind == 0 || kind == 1, "kind must be 0 (canonical) or 1 (semi-grand canonical)");
  ARG_CHECK(passes_per_replay >= 1 && max_replays >= 1 && sample_period >= 1, "bad graph arguments");
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  HIP_TRY(hipSetDevice(h->device));
  graph_release(h);
  RunPlan P;
  clexmc_status st = plan_run(h, kind, passes_per_replay, sample_period, canonical_rounds,
                              passes_per_replay * max_replays, P);
  if (st != CLEXMC_OK) return st;
  HIP_TRY(plan_attributes(P, h));
  HIP_TRY(hipStreamSynchronize(h->stream));
  hipError_t e = hipSuccess;
  hipGraph_t g = nullptr;
  for (int attempt = 0; attempt < 2; ++attempt) {
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    e = plan_launch(P, h);
    g = nullptr;
    hipError_t e2 = hipStreamEndCapture(h->stream, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess || !P.jit) break;
    // specialised launch refused: capture the generic kernel instead
    if (g) hipGraphDestroy(g);
    g = nullptr;
    drop_jit(h, P, h->chain_sc[0]);
    HIP_TRY(plan_attributes(P, h));
  }
  if (e != hipSuccess) {
    if (g) hipGraphDestroy(g);
    h->err = std::string("graph capture: ") + hipGetErrorString(e);
    return CLEXMC_ERR_HIP;
  }
  h->graph = g;
  e = hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    graph_release(h);
    h->err = std::string("graph instantiate: ") + hipGetErrorString(e);
    return CLEXMC_ERR_HIP;
  }
  h->graph_replays_left = max_replays;
  return CLEXMC_OK;
}

clexmc_status clexmc_graph_replay(clexmc_handle *h, int n_replays) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(h->graph_exec != nullptr, "no captured graph");
  ARG_CHECK(n_replays >= 1 && n_replays <= h->graph_replays_left, "replays beyond the captured sample capacity");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  for (int r = 0; r < n_replays; ++r) HIP_TRY(hipGraphLaunch(h->graph_exec, h->stream));
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  h->graph_replays_left -= n_replays;
  return CLEXMC_OK;
}

clexmc_status clexmc_graph_release(clexmc_handle *h) {
  if (!h) return CLEXMC_ERR_ARG;
  graph_release(h);
  return CLEXMC_OK;
}

// ---------------------------------------------------------------------------
// Autotuner: times short trial launches over block sizes (and LDS image
// formats where the int8 image fits) on the loaded chains, restoring the
// chains' device state after every trial, and keeps the fastest setting.
// Results are cached per (kind, proposal, CE shape, chain count, largest
// padded supercell, largest class) for the life of the process.
// ---------------------------------------------------------------------------
static std::map<std::string, std::array<int, 3>> g_tune_cache;  // threads, packing, lanes per site

clexmc_status clexmc_autotune(clexmc_handle *h, int kind, int trial_passes, int canonical_rounds, int *threads_out,
                              int *packing_out, int *lanes_out, double *ms_per_pass_out) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(kind == 0 || kind == 1, "kind must be 0 (canonical) or 1 (semi-grand canonical)");
  ARG_CHECK(trial_passes >= 1, "trial_passes >= 1 required");
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  HIP_TRY(hipSetDevice(h->device));
  int max_pad = 0, max_class = 1;
  long long occ_bytes = 0;
  for (int c = 0; c < h->n_chains; ++c) {
    const DevSupercell &d = h->scs[h->chain_sc[c]].dev;
    max_pad = std::max(max_pad, d.pad_bytes);
    max_class = std::max(max_class, d.max_class);
    occ_bytes += d.n_sites;
  }
  const std::string key = std::to_string(kind) + "/" + std::to_string(h->proposal) + "/" + std::to_string(h->n_corr) +
                          "/" + std::to_string(h->n_sublat) + "/" + std::to_string(h->n_chains) + "/" +
                          std::to_string(max_pad) + "/" + std::to_string(max_class) + "/" +
                          std::to_string(trial_passes);
  auto it = g_tune_cache.find(key);
  double best_ms = -1.0;
  int best_t = 0, best_p = 0, best_l = 1;
  if (it != g_tune_cache.end()) {
    best_t = it->second[0];
    best_p = it->second[1];
    best_l = it->second[2];
  } else {
    // snapshot of the chains' device state
    DevBufs B;
    int8_t *s_occ;
    ChainState *s_state;
    long long *s_binC;
    int *s_cnt;
    HIP_TRY(B.alloc(&s_occ, (size_t)occ_bytes));
    HIP_TRY(B.alloc(&s_state, (size_t)h->n_chains));
    HIP_TRY(B.alloc(&s_binC, (size_t)h->n_chains * h->n_orbits));
    HIP_TRY(B.alloc(&s_cnt, (size_t)h->n_chains));
    HIP_TRY(hipMemcpy(s_occ, h->d_occ, (size_t)occ_bytes, hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(s_state, h->d_state, sizeof(ChainState) * h->n_chains, hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(s_binC, h->d_binC, sizeof(long long) * h->n_chains * h->n_orbits, hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(s_cnt, h->d_sample_count, sizeof(int) * h->n_chains, hipMemcpyDeviceToDevice));
    // block sizes: powers of two from 64, and one lane per site of the
    // largest class (up to 1024 threads, specialised kernels only)
    const int t_class = std::min(kMaxJitThreads, std::max(64, ((max_class + 63) / 64) * 64));
    std::vector<int> t_list;
    for (int t = 64; t <= t_class; t *= 2) t_list.push_back(t);
    if (t_list.back() != t_class) t_list.push_back(t_class);
    for (int lanes = 1; lanes <= (h->binary ? 2 : 1); ++lanes)
    for (int packing = 1; packing <= 2; ++packing) {
      for (int t : t_list) {
        RunPlan P;
        if (plan_run(h, kind, trial_passes, trial_passes + 1, canonical_rounds, trial_passes, P, t, packing, nullptr,
                     lanes) != CLEXMC_OK)
          continue;
        if (packing == 2 && (P.learned || !P.packed)) continue;
        if (P.lanes != lanes || P.threads != t) continue;
        HIP_TRY(choose_slabs(h, {&P}));  // trials run as the real runs will (slabs included)
        if (P.slabs > 1)
          HIP_TRY(hipMemcpy(&P.pass_start, &h->d_state[0].pass_base, sizeof(P.pass_start), hipMemcpyDeviceToHost));
        HIP_TRY(plan_attributes(P, h));
        HIP_TRY(hipEventRecord(h->ev0, h->stream));
        HIP_TRY(plan_launch(P, h));
        HIP_TRY(hipEventRecord(h->ev1, h->stream));
        HIP_TRY(hipEventSynchronize(h->ev1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        HIP_TRY(hipMemcpy(h->d_occ, s_occ, (size_t)occ_bytes, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(h->d_state, s_state, sizeof(ChainState) * h->n_chains, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(h->d_binC, s_binC, sizeof(long long) * h->n_chains * h->n_orbits, hipMemcpyDeviceToDevice));
        HIP_TRY(hipMemcpy(h->d_sample_count, s_cnt, sizeof(int) * h->n_chains, hipMemcpyDeviceToDevice));
        if (best_ms < 0.0 || ms < best_ms) {
          best_ms = ms;
          best_t = t;
          best_p = packing;
          best_l = lanes;
        }
      }
    }
    ARG_CHECK(best_t > 0, "no launch configuration fits");
    g_tune_cache[key] = {best_t, best_p, best_l};
  }
  h->block_size = best_t;
  h->packing = best_p;
  h->lanes_per_site = best_l;
  if (threads_out) *threads_out = best_t;
  if (packing_out) *packing_out = best_p;
  if (lanes_out) *lanes_out = best_l;
  if (ms_per_pass_out) *ms_per_pass_out = best_ms < 0.0 ? -1.0 : best_ms / trial_passes;
  return CLEXMC_OK;
}

clexmc_status clexmc_lds_probe(clexmc_handle *h, int mode, double *gbps) {
  if (!h || !gbps) return CLEXMC_ERR_ARG;
  ARG_CHECK(mode == 0 || mode == 1, "mode must be 0 (byte gathers) or 1 (16-byte reads)");
  HIP_TRY(hipSetDevice(h->device));
  int n_cu = 0;
  HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, h->device));
  const int blocks = 4 * n_cu, iters = 4096;
  DevBufs B;
  unsigned *sink;
  HIP_TRY(B.alloc(&sink, (size_t)blocks * 512));
  HIP_TRY(launch_lds_probe(mode, 16, blocks, sink, h->stream));  // warm-up
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  HIP_TRY(launch_lds_probe(mode, iters, blocks, sink, h->stream));
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  const double per_read = mode == 0 ? 1.0 : 16.0;
  const double bytes = (double)blocks * 512.0 * iters * 8.0 * per_read;
  *gbps = bytes / (ms * 1e-3) / 1e9;
  return CLEXMC_OK;
}

clexmc_status clexmc_reanchor_energies(clexmc_handle *h, double *max_drift) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  HIP_TRY(hipSetDevice(h->device));
  std::vector<double> corr((size_t)h->n_chains * h->n_corr);
  clexmc_status st = run_corr(h, 0, h->n_chains, corr.data());
  if (st != CLEXMC_OK) return st;
  std::vector<ChainState> state(h->n_chains);
  HIP_TRY(hipMemcpy(state.data(), h->d_state, sizeof(ChainState) * h->n_chains, hipMemcpyDeviceToHost));
  double drift = 0.0;
  for (int c = 0; c < h->n_chains; ++c) {
    const double *cc = corr.data() + (size_t)c * h->n_corr;
    double E = 0.0;
    for (int i = 0; i < h->n_corr; ++i) E += h->eci[i] * cc[i];
    if (h->binary) {
      E = 0.0;
      for (int a = 0; a < h->n_orbits; ++a)
        E += h->bin_w[a] * (double)llround(cc[h->orbit_func_begin[a]] * h->orbit_mult[a]);
    }
    drift = std::max(drift, std::fabs(state[c].E - E));
    state[c].E = E;
  }
  HIP_TRY(hipMemcpy(h->d_state, state.data(), sizeof(ChainState) * h->n_chains, hipMemcpyHostToDevice));
  if (max_drift) *max_drift = drift;
  return CLEXMC_OK;
}

clexmc_status clexmc_last_run_ms(clexmc_handle *h, double *ms) {
  if (!h || !ms) return CLEXMC_ERR_ARG;
  *ms = h->last_ms;
  return CLEXMC_OK;
}

clexmc_status clexmc_read_samples(clexmc_handle *h, int *n_samples, double *E, int32_t *counts) {
  if (!h || !n_samples) return CLEXMC_ERR_ARG;
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  HIP_TRY(hipSetDevice(h->device));
  std::vector<int> cnt(h->n_chains);
  HIP_TRY(hipMemcpy(cnt.data(), h->d_sample_count, sizeof(int) * h->n_chains, hipMemcpyDeviceToHost));
  int ns = 0;
  for (int v : cnt) ns = std::max(ns, v);
  const int want = *n_samples;
  *n_samples = ns;
  if (!E && !counts) return CLEXMC_OK;
  ARG_CHECK(want >= ns, "output arrays too small for the recorded samples");
  if (ns == 0) return CLEXMC_OK;
  if (E)
    HIP_TRY(hipMemcpy2D(E, sizeof(double) * ns, h->d_sampE, sizeof(double) * h->sample_cap, sizeof(double) * ns,
                        h->n_chains, hipMemcpyDeviceToHost));
  if (counts)
    HIP_TRY(hipMemcpy2D(counts, sizeof(int) * kMaxOcc * ns, h->d_sampN, sizeof(int) * kMaxOcc * h->sample_cap,
                        sizeof(int) * kMaxOcc * ns, h->n_chains, hipMemcpyDeviceToHost));
  return CLEXMC_OK;
}

clexmc_status clexmc_clear_samples(clexmc_handle *h) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipMemset(h->d_sample_count, 0, sizeof(int) * h->n_chains));
  h->samples_bound = 0;
  return CLEXMC_OK;
}

clexmc_status clexmc_read_state(clexmc_handle *h, int chain, int8_t *occ, double *E_running, long long *n_attempt,
                                long long *n_accept) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(chain >= 0 && chain < h->n_chains, "chain out of range");
  HIP_TRY(hipSetDevice(h->device));
  ChainState ch;
  HIP_TRY(hipMemcpy(&ch, h->d_state + chain, sizeof(ChainState), hipMemcpyDeviceToHost));
  const DevSupercell &d = h->scs[h->chain_sc[chain]].dev;
  if (occ) HIP_TRY(hipMemcpy(occ, h->d_occ + h->chain_off[chain], (size_t)d.n_sites, hipMemcpyDeviceToHost));
  if (E_running) *E_running = ch.E;
  if (n_attempt) *n_attempt = ch.n_attempt;
  if (n_accept) *n_accept = ch.n_accept;
  return CLEXMC_OK;
}

clexmc_status clexmc_read_binary_counts(clexmc_handle *h, int chain, long long *counts_out) {
  if (!h || !counts_out) return CLEXMC_ERR_ARG;
  ARG_CHECK(h->binary, "binary path not active");
  ARG_CHECK(chain >= 0 && chain < h->n_chains, "chain out of range");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipMemcpy(counts_out, h->d_binC + (size_t)chain * h->n_orbits, sizeof(long long) * h->n_orbits,
                    hipMemcpyDeviceToHost));
  return CLEXMC_OK;
}

clexmc_status clexmc_philox_device(clexmc_handle *h, int n, const uint32_t *ctr, const uint32_t *key,
                                   uint32_t *out) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(n >= 1 && ctr && key && out, "bad philox arguments");
  HIP_TRY(hipSetDevice(h->device));
  uint4 *dc = nullptr, *dout = nullptr;
  uint2 *dk = nullptr;
  HIP_TRY(hipMalloc(&dc, sizeof(uint4) * n));
  HIP_TRY(hipMalloc(&dk, sizeof(uint2) * n));
  HIP_TRY(hipMalloc(&dout, sizeof(uint4) * n));
  hipError_t e = hipMemcpy(dc, ctr, sizeof(uint4) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dk, key, sizeof(uint2) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_philox(n, dc, dk, dout, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(out, dout, sizeof(uint4) * n, hipMemcpyDeviceToHost);
  hipFree(dc);
  hipFree(dk);
  hipFree(dout);
  if (e != hipSuccess) {
    h->err = std::string("philox: ") + hipGetErrorString(e);
    return CLEXMC_ERR_HIP;
  }
  return CLEXMC_OK;
}


// ---------------------------------------------------------------------------
// learned proposal
// ---------------------------------------------------------------------------
static const int kNParams = kHidden * kEnvMax + 2 * kHidden + 1;

clexmc_status clexmc_set_proposal(clexmc_handle *h, int proposal, int rounds_per_group) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(proposal >= 0 && proposal <= 3,
            "proposal must be 0 (plain), 1 (learned), 2 (grouped uniform) or 3 (grouped informed)");
  ARG_CHECK(rounds_per_group >= 1 && rounds_per_group <= 4096, "rounds_per_group out of range");
  h->proposal = proposal;
  h->rounds_per_group = rounds_per_group;
  return CLEXMC_OK;
}

clexmc_status clexmc_mlp_info(clexmc_handle *h, int *n_env, int *n_params, int *env_sites) {
  if (!h) return CLEXMC_ERR_ARG;
  if (n_env) *n_env = h->n_env;
  if (n_params) *n_params = kNParams;
  if (env_sites)
    for (size_t i = 0; i < h->env_sites.size(); ++i) {
      env_sites[4 * i] = h->env_sites[i].x;
      env_sites[4 * i + 1] = h->env_sites[i].y;
      env_sites[4 * i + 2] = h->env_sites[i].z;
      env_sites[4 * i + 3] = h->env_sites[i].w;
    }
  return CLEXMC_OK;
}

clexmc_status clexmc_mlp_set_params(clexmc_handle *h, const float *params, float eps) {
  if (!h || !params) return CLEXMC_ERR_ARG;
  ARG_CHECK(eps > 0.0f && eps < 1e3f, "eps must be positive");
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipMemcpy(h->d_mlp_params, params, sizeof(float) * kNParams, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(h->d_adam_m, 0, sizeof(float) * kNParams));
  HIP_TRY(hipMemset(h->d_adam_v, 0, sizeof(float) * kNParams));
  h->adam_t = 0;
  h->mlp_initialized = true;
  h->mlp_eps = eps;
  HIP_TRY(launch_mlp_publish(h->d_mlp_params, h->d_mlp, h->n_env, eps, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return CLEXMC_OK;
}

clexmc_status clexmc_mlp_get_params(clexmc_handle *h, float *params) {
  if (!h || !params) return CLEXMC_ERR_ARG;
  HIP_TRY(hipSetDevice(h->device));
  HIP_TRY(hipMemcpy(params, h->d_mlp_params, sizeof(float) * kNParams, hipMemcpyDeviceToHost));
  return CLEXMC_OK;
}

clexmc_status clexmc_mlp_forward(clexmc_handle *h, int n, const int8_t *X, float *out) {
  if (!h || !X || !out) return CLEXMC_ERR_ARG;
  ARG_CHECK(n >= 1, "n >= 1 required");
  HIP_TRY(hipSetDevice(h->device));
  int8_t *dX = nullptr;
  float *dO = nullptr;
  HIP_TRY(hipMalloc(&dX, (size_t)n * kEnvMax));
  HIP_TRY(hipMalloc(&dO, sizeof(float) * 2 * (size_t)n));
  hipError_t e = hipMemcpy(dX, X, (size_t)n * kEnvMax, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_mlp_infer(h->d_mlp, n, dX, dO, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess) e = hipMemcpy(out, dO, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost);
  hipFree(dX);
  hipFree(dO);
  if (e != hipSuccess) {
    h->err = std::string("mlp_forward: ") + hipGetErrorString(e);
    return CLEXMC_ERR_HIP;
  }
  return CLEXMC_OK;
}

clexmc_status clexmc_wave_pick_device(clexmc_handle *h, int n_rows, const float *v, const float *u, float *scan_out,
                                      int *pick_out) {
  if (!h || !v || !u || !scan_out || !pick_out) return CLEXMC_ERR_ARG;
  ARG_CHECK(n_rows >= 1 && n_rows <= 65535, "n_rows out of range");
  HIP_TRY(hipSetDevice(h->device));
  DevBufs B;
  float *dv, *du, *ds;
  int *dp;
  HIP_TRY(B.alloc(&dv, (size_t)n_rows * 64));
  HIP_TRY(B.alloc(&du, (size_t)n_rows));
  HIP_TRY(B.alloc(&ds, (size_t)n_rows * 64));
  HIP_TRY(B.alloc(&dp, (size_t)n_rows));
  HIP_TRY(hipMemcpy(dv, v, sizeof(float) * 64 * (size_t)n_rows, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(du, u, sizeof(float) * (size_t)n_rows, hipMemcpyHostToDevice));
  HIP_TRY(launch_wave_pick(n_rows, dv, du, ds, dp, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(scan_out, ds, sizeof(float) * 64 * (size_t)n_rows, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(pick_out, dp, sizeof(int) * (size_t)n_rows, hipMemcpyDeviceToHost));
  return CLEXMC_OK;
}

static TrainArgs train_args(clexmc_handle *h, int n, const int8_t *dX, const float *dy, const float *dt, float *dH,
                            float *ddl, float *dloss, float *dgrad) {
  TrainArgs t;
  t.n = n;
  t.n_env = h->n_env;
  t.X = dX;
  t.y = dy;
  t.t = dt;
  t.W1 = h->d_mlp_params;
  t.b1 = h->d_mlp_params + kHidden * kEnvMax;
  t.W2 = h->d_mlp_params + kHidden * kEnvMax + kHidden;
  t.b2 = h->d_mlp_params + kHidden * kEnvMax + 2 * kHidden;
  t.H = dH;
  t.dl = ddl;
  t.loss = dloss;
  t.grad = dgrad;
  return t;
}



clexmc_status clexmc_mlp_gradients(clexmc_handle *h, int n, const int8_t *X, const float *y, const float *t,
                                   float *grad_out, float *loss_out) {
  if (!h || !X || !y || !grad_out) return CLEXMC_ERR_ARG;
  ARG_CHECK(n >= 1, "n >= 1 required");
  HIP_TRY(hipSetDevice(h->device));
  DevBufs B;
  int8_t *dX;
  float *dy, *dt = nullptr, *dH, *ddl, *dloss, *dg;
  HIP_TRY(B.alloc(&dX, (size_t)n * kEnvMax));
  HIP_TRY(B.alloc(&dy, n));
  if (t) {
    HIP_TRY(B.alloc(&dt, n));
    HIP_TRY(hipMemcpy(dt, t, sizeof(float) * n, hipMemcpyHostToDevice));
  }
  HIP_TRY(B.alloc(&dH, (size_t)n * kHidden));
  HIP_TRY(B.alloc(&ddl, n));
  HIP_TRY(B.alloc(&dloss, n));
  HIP_TRY(B.alloc(&dg, kNParams));
  HIP_TRY(hipMemcpy(dX, X, (size_t)n * kEnvMax, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dy, y, sizeof(float) * n, hipMemcpyHostToDevice));
  HIP_TRY(launch_mlp_train_step(train_args(h, n, dX, dy, dt, dH, ddl, dloss, dg), h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(grad_out, dg, sizeof(float) * kNParams, hipMemcpyDeviceToHost));
  if (loss_out) {
    std::vector<float> l(n);
    HIP_TRY(hipMemcpy(l.data(), dloss, sizeof(float) * n, hipMemcpyDeviceToHost));
    double acc = 0.0;
    for (float v : l) acc += v;
    *loss_out = (float)(acc / n);
  }
  return CLEXMC_OK;
}

static clexmc_status collect_device(clexmc_handle *h, int n_per_chain, uint32_t seed, int include_chem, int8_t *dX,
                                    float *dy, float *dt) {
  CollectArgs c;
  c.sys = h->d_sys;
  c.scs = h->d_scs;
  c.params = h->d_params;
  c.occ = h->d_occ;
  c.env_sites = h->d_env_sites;
  c.n_env = h->n_env;
  c.n_per_chain = n_per_chain;
  c.seed = seed;
  c.include_chem = include_chem;
  c.X = dX;
  c.y = dy;
  c.t = dt;
  HIP_TRY(launch_mlp_collect(c, h->n_chains, h->stream));
  return CLEXMC_OK;
}

clexmc_status clexmc_mlp_collect(clexmc_handle *h, int n_per_chain, uint32_t seed, int include_chem, int8_t *X_out,
                                 float *y_out, float *t_out) {
  if (!h || !X_out || !y_out) return CLEXMC_ERR_ARG;
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  ARG_CHECK(h->binary, "training data collection is implemented for binary cluster expansions");
  ARG_CHECK(n_per_chain >= 1, "n_per_chain >= 1 required");
  HIP_TRY(hipSetDevice(h->device));
  const size_t n = (size_t)n_per_chain * h->n_chains;
  DevBufs B;
  int8_t *dX;
  float *dy, *dt;
  HIP_TRY(B.alloc(&dX, n * kEnvMax));
  HIP_TRY(B.alloc(&dy, n));
  HIP_TRY(B.alloc(&dt, n));
  clexmc_status st = collect_device(h, n_per_chain, seed, include_chem, dX, dy, dt);
  if (st != CLEXMC_OK) return st;
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(X_out, dX, n * kEnvMax, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(y_out, dy, sizeof(float) * n, hipMemcpyDeviceToHost));
  if (t_out) HIP_TRY(hipMemcpy(t_out, dt, sizeof(float) * n, hipMemcpyDeviceToHost));
  return CLEXMC_OK;
}

clexmc_status clexmc_mlp_train(clexmc_handle *h, int n_per_chain, int n_steps, float lr, uint32_t seed,
                               int include_chem, float *loss_out) {
  if (!h) return CLEXMC_ERR_ARG;
  ARG_CHECK(h->n_chains > 0, "no chains loaded");
  ARG_CHECK(h->binary, "training is implemented for binary cluster expansions");
  ARG_CHECK(n_per_chain >= 1 && n_steps >= 1 && lr > 0.0f, "bad training arguments");
  HIP_TRY(hipSetDevice(h->device));
  const size_t n = (size_t)n_per_chain * h->n_chains;
  ARG_CHECK(n < (1u << 24), "training batch too large");
  DevBufs B;
  int8_t *dX;
  float *dy, *dt, *dH, *ddl, *dloss, *dg;
  HIP_TRY(B.alloc(&dX, n * kEnvMax));
  HIP_TRY(B.alloc(&dy, n));
  HIP_TRY(B.alloc(&dt, n));
  HIP_TRY(B.alloc(&dH, n * kHidden));
  HIP_TRY(B.alloc(&ddl, n));
  HIP_TRY(B.alloc(&dloss, n));
  HIP_TRY(B.alloc(&dg, kNParams));
  if (!h->mlp_initialized) {
    // He initialisation of a fresh network (a zero network has dead ReLUs):
    // W1 ~ N(0, 2/n_env) on the used inputs, b1 = 0.1, W2 ~ N(0, 1/64), b2 = 0
    std::mt19937_64 rng(0x5EEDull ^ ((uint64_t)seed << 20));
    std::normal_distribution<float> nw1(0.0f, std::sqrt(2.0f / (float)std::max(1, h->n_env)));
    std::normal_distribution<float> nw2(0.0f, 0.125f);
    std::vector<float> p0(kNParams, 0.0f);
    for (int j = 0; j < kHidden; ++j)
      for (int k = 0; k < h->n_env; ++k) p0[(size_t)j * kEnvMax + k] = nw1(rng);
    for (int j = 0; j < kHidden; ++j) p0[kHidden * kEnvMax + j] = 0.1f;
    for (int j = 0; j < kHidden; ++j) p0[kHidden * kEnvMax + kHidden + j] = nw2(rng);
    HIP_TRY(hipMemcpy(h->d_mlp_params, p0.data(), sizeof(float) * kNParams, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(h->d_adam_m, 0, sizeof(float) * kNParams));
    HIP_TRY(hipMemset(h->d_adam_v, 0, sizeof(float) * kNParams));
    h->adam_t = 0;
    h->mlp_initialized = true;
  }
  clexmc_status st = collect_device(h, n_per_chain, seed, include_chem, dX, dy, dt);
  if (st != CLEXMC_OK) return st;
  std::vector<float> l(n);
  for (int it = 0; it < n_steps; ++it) {
    HIP_TRY(launch_mlp_train_step(train_args(h, (int)n, dX, dy, dt, dH, ddl, dloss, dg), h->stream));
    h->adam_t += 1;
    AdamArgs ad;
    ad.n_params = kNParams;
    ad.p = h->d_mlp_params;
    ad.m = h->d_adam_m;
    ad.v = h->d_adam_v;
    ad.grad = dg;
    ad.lr = lr;
    ad.b1m = 0.9f;
    ad.b2m = 0.999f;
    ad.b1t = (float)std::pow(0.9, (double)h->adam_t);
    ad.b2t = (float)std::pow(0.999, (double)h->adam_t);
    HIP_TRY(launch_mlp_adam(ad, h->stream));
    if (loss_out) {
      HIP_TRY(hipMemcpyAsync(l.data(), dloss, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
      double acc = 0.0;
      for (float v : l) acc += v;
      loss_out[it] = (float)(acc / n);
    }
  }
  HIP_TRY(launch_mlp_publish(h->d_mlp_params, h->d_mlp, h->n_env, h->mlp_eps, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return CLEXMC_OK;
}

void clexmc_destroy(clexmc_handle *h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  free_chains(h);
  for (auto &s : h->scs) {
    free_all(s.allocs);
    for (auto m : s.jit_mods) hipModuleUnload(m);
  }
  for (auto m : h->fused_mods) hipModuleUnload(m);
  free_all(h->sys_allocs);
  free_mswaps(h);
  if (h->d_scs) hipFree(h->d_scs);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  for (auto e : h->side_ev) hipEventDestroy(e);
  for (auto s : h->side) {
    hipStreamSynchronize(s);
    hipStreamDestroy(s);
  }
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}

}  // extern "C"
