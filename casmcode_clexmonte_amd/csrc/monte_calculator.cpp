// C++ calculator layer over the clexmc C ABI; see monte_calculator.hh.
#include "monte_calculator.hh"

#include <algorithm>
#include <functional>
#include <array>
#include <chrono>
#include <cstring>
#include <cmath>
#include <fstream>
#include <numeric>
#include <set>
#include <sstream>

namespace CASM {
namespace clexmonte {
namespace gpu {

namespace {

const std::set<std::string> kIntArrays = {
    "n_sublat", "n_occ", "n_occ_max", "n_func_max", "n_orbits", "orbit_k", "orbit_mult", "orbit_nfunc",
    "orbit_func_begin", "corr_orbit_begin", "corr_elem_begin", "corr_elem_sites", "orbit_table_offset",
    "n_corr", "flower_begin", "entry_orbit", "entry_k", "entry_site_begin", "entry_sites", "entry_table",
    "binary", "bin_seg_begin", "bin_segs", "n_components", "occ_component", "n_independent"};

// Integrated autocorrelation time with Sokal's window (c = 5)
double tau_int(std::vector<double> const &x) {
  const size_t n = x.size();
  if (n < 4) return 0.5;
  double m = std::accumulate(x.begin(), x.end(), 0.0) / n;
  double v = 0.0;
  for (double a : x) v += (a - m) * (a - m);
  v /= n;
  if (v <= 0.0) return 0.5;
  double tau = 0.5;
  for (size_t M = 1; M < n; ++M) {
    double c = 0.0;
    for (size_t i = 0; i + M < n; ++i) c += (x[i] - m) * (x[i + M] - m);
    c /= (n * v);
    tau += c;
    if ((double)M >= 5.0 * tau) break;
  }
  return std::max(tau, 0.5);
}

double mean_of(std::vector<double> const &x) {
  return x.empty() ? NAN : std::accumulate(x.begin(), x.end(), 0.0) / x.size();
}

double cov_of(std::vector<double> const &x, std::vector<double> const &y) {
  if (x.empty()) return NAN;
  double mx = mean_of(x), my = mean_of(y), c = 0.0;
  for (size_t i = 0; i < x.size(); ++i) c += (x[i] - mx) * (y[i] - my);
  return c / x.size();
}

std::vector<double> column(std::vector<std::vector<double>> const &s, size_t i) {
  std::vector<double> out;
  out.reserve(s.size());
  for (auto const &r : s) out.push_back(r[i]);
  return out;
}

}  // namespace

void check(clexmc_handle *h, clexmc_status st, std::string const &what) {
  if (st != CLEXMC_OK) {
    std::string msg = h ? clexmc_last_error(h) : "";
    throw std::runtime_error("Error in " + what + ": " + msg + " (status " + std::to_string((int)st) + ")");
  }
}

// ---------------------------------------------------------------------------
// System
// ---------------------------------------------------------------------------
std::shared_ptr<System> System::from_flat_file(std::string const &path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("Error in System::from_flat_file: cannot open " + path);
  auto sys = std::make_shared<System>();
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty()) continue;
    std::istringstream ss(line);
    std::string name;
    size_t n = 0;
    ss >> name >> n;
    if (kIntArrays.count(name)) {
      std::vector<int> v(n);
      for (size_t i = 0; i < n; ++i) ss >> v[i];
      sys->m_i[name] = v;
    } else {
      std::vector<double> v(n);
      for (size_t i = 0; i < n; ++i) ss >> v[i];
      sys->m_d[name] = v;
    }
    if (!ss) throw std::runtime_error("Error in System::from_flat_file: malformed entry '" + name + "'");
  }
  sys->_finalize();
  return sys;
}

void System::_finalize() {
  auto gi = [&](const char *k) -> std::vector<int> & {
    auto it = m_i.find(k);
    if (it == m_i.end()) throw std::runtime_error(std::string("Error in System: missing '") + k + "'");
    return it->second;
  };
  auto gd = [&](const char *k) -> std::vector<double> & {
    auto it = m_d.find(k);
    if (it == m_d.end()) throw std::runtime_error(std::string("Error in System: missing '") + k + "'");
    return it->second;
  };
  n_sublat = gi("n_sublat")[0];
  n_occ_max = gi("n_occ_max")[0];
  n_occ = gi("n_occ");
  n_corr = gi("n_corr")[0];
  n_orbits = gi("n_orbits")[0];
  orbit_func_begin = gi("orbit_func_begin");
  orbit_mult = gi("orbit_mult");
  eci = gd("eci");
  n_components = gi("n_components")[0];
  n_independent = gi("n_independent")[0];
  occ_component = gi("occ_component");
  comp_origin = gd("comp_origin");
  comp_R = gd("comp_R");
  // pseudo-inverse of R (n_components x n_independent) by normal equations
  const int nc = n_components, ni = n_independent;
  m_Rpinv.assign((size_t)ni * nc, 0.0);
  if (ni > 0) {
    std::vector<double> G((size_t)ni * ni, 0.0);
    for (int a = 0; a < ni; ++a)
      for (int b = 0; b < ni; ++b)
        for (int c = 0; c < nc; ++c) G[a * ni + b] += comp_R[a * nc + c] * comp_R[b * nc + c];
    // invert G (small) by Gauss-Jordan
    std::vector<double> Ginv((size_t)ni * ni, 0.0);
    for (int i = 0; i < ni; ++i) Ginv[i * ni + i] = 1.0;
    for (int c = 0; c < ni; ++c) {
      int p = c;
      for (int r = c + 1; r < ni; ++r)
        if (std::fabs(G[r * ni + c]) > std::fabs(G[p * ni + c])) p = r;
      if (std::fabs(G[p * ni + c]) < 1e-14) throw std::runtime_error("Error in System: singular composition axes");
      for (int k = 0; k < ni; ++k) {
        std::swap(G[c * ni + k], G[p * ni + k]);
        std::swap(Ginv[c * ni + k], Ginv[p * ni + k]);
      }
      double d = G[c * ni + c];
      for (int k = 0; k < ni; ++k) {
        G[c * ni + k] /= d;
        Ginv[c * ni + k] /= d;
      }
      for (int r = 0; r < ni; ++r) {
        if (r == c) continue;
        double f = G[r * ni + c];
        for (int k = 0; k < ni; ++k) {
          G[r * ni + k] -= f * G[c * ni + k];
          Ginv[r * ni + k] -= f * Ginv[c * ni + k];
        }
      }
    }
    for (int a = 0; a < ni; ++a)
      for (int c = 0; c < nc; ++c) {
        double v = 0.0;
        for (int b = 0; b < ni; ++b) v += Ginv[a * ni + b] * comp_R[b * nc + c];
        m_Rpinv[a * nc + c] = v;
      }
  }
}

clexmc_system_desc System::desc() const {
  clexmc_system_desc d;
  std::memset(&d, 0, sizeof(d));
  auto I = [&](const char *k) -> const int * {
    auto it = m_i.find(k);
    return it == m_i.end() ? nullptr : it->second.data();
  };
  auto N = [&](const char *k) -> int {
    auto it = m_i.find(k);
    if (it != m_i.end()) return (int)it->second.size();
    auto jt = m_d.find(k);
    return jt == m_d.end() ? 0 : (int)jt->second.size();
  };
  auto D = [&](const char *k) -> const double * {
    auto it = m_d.find(k);
    return it == m_d.end() ? nullptr : it->second.data();
  };
  d.n_sublat = n_sublat;
  d.n_occ = I("n_occ");
  d.n_occ_max = n_occ_max;
  d.n_func_max = m_i.at("n_func_max")[0];
  d.site_func = D("site_func");
  d.n_orbits = n_orbits;
  d.orbit_k = I("orbit_k");
  d.orbit_mult = I("orbit_mult");
  d.orbit_nfunc = I("orbit_nfunc");
  d.orbit_func_begin = I("orbit_func_begin");
  d.corr_orbit_begin = I("corr_orbit_begin");
  d.corr_elem_begin = I("corr_elem_begin");
  d.corr_elem_sites = I("corr_elem_sites");
  d.n_corr_elem_sites = N("corr_elem_sites") / 4;
  d.orbit_table_offset = I("orbit_table_offset");
  d.orbit_tables = D("orbit_tables");
  d.n_orbit_tables = N("orbit_tables");
  d.n_corr = n_corr;
  d.eci = eci.data();
  d.flower_begin = I("flower_begin");
  d.n_entries = N("entry_orbit");
  d.entry_orbit = I("entry_orbit");
  d.entry_k = I("entry_k");
  d.entry_site_begin = I("entry_site_begin");
  d.entry_sites = I("entry_sites");
  d.n_entry_sites = N("entry_sites") / 4;
  d.entry_table = I("entry_table");
  d.flower_tables = D("flower_tables");
  d.n_flower_tables = N("flower_tables");
  d.binary = m_i.at("binary")[0];
  if (d.binary) {
    d.bin_seg_begin = I("bin_seg_begin");
    d.bin_segs = I("bin_segs");
    d.n_bin_segs = N("bin_segs") / 3;
    d.bin_orbit_weight = D("bin_orbit_weight");
  }
  d.asym = I("asym");  // optional; NULL: units by occupant count
  return d;
}

std::vector<double> System::mol_composition(Configuration const &config) const {
  std::vector<double> n(n_components, 0.0);
  const Index nc = config.n_unitcells();
  for (Index l = 0; l < (Index)config.occupation.size(); ++l) {
    int b = (int)(l / nc);
    int c = occ_component[b * n_occ_max + config.occupation[l]];
    n[c] += 1.0;
  }
  for (auto &v : n) v /= (double)nc;
  return n;
}

std::vector<double> System::param_composition(std::vector<double> const &mol) const {
  std::vector<double> x(n_independent, 0.0);
  for (int a = 0; a < n_independent; ++a)
    for (int c = 0; c < n_components; ++c) x[a] += m_Rpinv[a * n_components + c] * (mol[c] - comp_origin[c]);
  return x;
}

std::vector<double> System::chem_weights(std::vector<double> const &mu) const {
  std::vector<double> wc(n_components, 0.0);
  for (int c = 0; c < n_components; ++c)
    for (int a = 0; a < n_independent; ++a) wc[c] += mu[a] * m_Rpinv[a * n_components + c];
  std::vector<double> w((size_t)n_sublat * n_occ_max, 0.0);
  for (int b = 0; b < n_sublat; ++b)
    for (int o = 0; o < n_occ[b]; ++o) w[b * n_occ_max + o] = wc[occ_component[b * n_occ_max + o]];
  return w;
}

void OccLocation::initialize(System const &system, Configuration const &config) {
  const Index nc = config.n_unitcells();
  m_mol_size = 0;
  m_counts.assign(system.n_occ_max, 0);
  for (Index l = 0; l < (Index)config.occupation.size(); ++l) {
    int b = (int)(l / nc);
    if (system.n_occ[b] < 2) continue;
    m_mol_size++;
    m_counts[config.occupation[l]]++;
  }
}

// ---------------------------------------------------------------------------
// BaseMonteCalculator
// ---------------------------------------------------------------------------
BaseMonteCalculator::~BaseMonteCalculator() { release_handle(); }

void BaseMonteCalculator::release_handle() {
  if (m_handle) clexmc_destroy(m_handle);
  m_handle = nullptr;
  m_sc_ids.clear();
}

void BaseMonteCalculator::reset(std::map<std::string, std::string> const &_params, std::shared_ptr<System> _system) {
  if (!_system) throw std::runtime_error("Error in " + calculator_name + "::reset: system==nullptr");
  params = _params;
  system = _system;
  release_handle();
  auto it = params.find("device");
  device = it == params.end() ? 0 : std::stoi(it->second);
  _reset();
}

clexmc_handle *BaseMonteCalculator::handle() {
  if (!system) throw std::runtime_error("Error in " + calculator_name + ": system==nullptr (call reset)");
  if (!m_handle) {
    clexmc_system_desc d = system->desc();
    clexmc_handle *h = nullptr;
    clexmc_status st = clexmc_create(&d, device, &h);
    if (st != CLEXMC_OK) {
      std::string msg = h ? clexmc_last_error(h) : "no HIP device";
      if (h) clexmc_destroy(h);
      throw std::runtime_error("Error in " + calculator_name + ": clexmc_create failed: " + msg);
    }
    m_handle = h;
    try {
      _configure_handle(h);
    } catch (...) {
      release_handle();
      throw;
    }
  }
  return m_handle;
}

namespace {

// Potential over chain 0 of a calculator's handle
class GpuPotential : public BaseMontePotential {
 public:
  GpuPotential(clexmc_handle *h, std::shared_ptr<System> sys, Index n_unitcells, bool sgc,
               std::vector<double> param_chem_pot)
      : m_h(h), m_sys(sys), m_n(n_unitcells), m_sgc(sgc), m_mu(param_chem_pot) {}

  double per_supercell() override {
    std::vector<double> corr(m_sys->n_corr);
    check(m_h, clexmc_correlations(m_h, 0, corr.data()), "BaseMontePotential::per_supercell");
    double E = 0.0;
    for (int i = 0; i < m_sys->n_corr; ++i) E += m_sys->eci[i] * corr[i];
    if (m_sgc) {
      Configuration c;
      // composition from the device state
      clexmc_status st;
      long long na, nacc;
      double Er;
      std::vector<int8_t> occ(m_occ_size);
      st = clexmc_read_state(m_h, 0, occ.data(), &Er, &na, &nacc);
      check(m_h, st, "BaseMontePotential::per_supercell");
      c.L[0] = m_L[0];
      c.L[1] = m_L[1];
      c.L[2] = m_L[2];
      c.occupation.assign(occ.begin(), occ.end());
      auto x = m_sys->param_composition(m_sys->mol_composition(c));
      for (size_t a = 0; a < x.size(); ++a) E -= (double)m_n * m_mu[a] * x[a];
    }
    return E;
  }
  double per_unitcell() override { return per_supercell() / (double)m_n; }

  double occ_delta_per_supercell(std::vector<Index> const &linear_site_index,
                                 std::vector<int> const &new_occ) override {
    if (linear_site_index.size() != new_occ.size() || linear_site_index.empty() || linear_site_index.size() > 4)
      throw std::runtime_error("Error in occ_delta_per_supercell: 1..4 sites with matching new_occ required");
    std::vector<int32_t> sites(linear_site_index.begin(), linear_site_index.end());
    std::vector<int32_t> nw(new_occ.begin(), new_occ.end());
    int32_t chain = 0;
    float dE = 0.f;
    check(m_h,
          clexmc_delta_energy_batch(m_h, 1, &chain, sites.data(), nw.data(), (int)sites.size(), m_sgc ? 1 : 0, &dE),
          "occ_delta_per_supercell");
    return (double)dE;
  }

  int m_L[3] = {1, 1, 1};
  size_t m_occ_size = 0;

 private:
  clexmc_handle *m_h;
  std::shared_ptr<System> m_sys;
  Index m_n;
  bool m_sgc;
  std::vector<double> m_mu;
};

std::vector<double> required_vector(State const &s, std::string const &k) {
  auto it = s.conditions.vector_values.find(k);
  if (it == s.conditions.vector_values.end()) throw std::runtime_error("Error: missing vector condition '" + k + "'");
  return it->second;
}

double required_scalar(State const &s, std::string const &k) {
  auto it = s.conditions.scalar_values.find(k);
  if (it == s.conditions.scalar_values.end()) throw std::runtime_error("Error: missing scalar condition '" + k + "'");
  return it->second;
}

}  // namespace

void BaseMonteCalculator::_load_chains(int kind, std::vector<State *> const &states,
                                       std::vector<std::vector<double>> const &chem, RunManager const &run_manager,
                                       std::vector<double> &temps) {
  clexmc_handle *h = handle();
  SamplingFixtureParams const &P = run_manager.params;
  const int n = (int)states.size();
  if (n == 0) throw std::runtime_error("Error in " + calculator_name + "::run: no states");
  std::vector<int> ids(n);
  std::vector<int8_t> occ;
  std::vector<double> chem_all;
  temps.assign(n, 0.0);
  std::vector<uint64_t> seeds(n);
  for (int c = 0; c < n; ++c) {
    Configuration const &cfg = states[c]->configuration;
    std::array<int, 3> key = {cfg.L[0], cfg.L[1], cfg.L[2]};
    auto it = m_sc_ids.find(key);
    if (it == m_sc_ids.end()) {
      clexmc_supercell_desc d;
      std::memset(&d, 0, sizeof(d));
      d.L[0] = key[0];
      d.L[1] = key[1];
      d.L[2] = key[2];
      d.n_classes = 0;  // library-computed independent sets
      d.n_pair_phases = kind == 0 ? P.n_pair_phases : 0;
      int id = -1;
      check(h, clexmc_add_supercell(h, &d, &id), calculator_name + "::run (supercell)");
      it = m_sc_ids.emplace(key, id).first;
    }
    ids[c] = it->second;
    if ((Index)cfg.occupation.size() != cfg.n_unitcells() * system->n_sublat)
      throw std::runtime_error("Error in " + calculator_name + "::run: occupation size mismatch");
    for (int v : cfg.occupation) occ.push_back((int8_t)v);
    temps[c] = required_scalar(*states[c], "temperature");
    seeds[c] = run_manager.seed + (uint64_t)c * 0x9E3779B97F4A7C15ull;
    for (double w : chem[c]) chem_all.push_back(w);
  }
  check(h, clexmc_load_chains(h, n, ids.data(), occ.data(), temps.data(), chem_all.data(), seeds.data(), KB),
        calculator_name + "::run (load)");
}

void BaseMonteCalculator::_run_chains(int kind, std::vector<State *> const &states,
                                      std::vector<std::vector<double>> const &chem, RunManager &run_manager,
                                      std::vector<Results> &results) {
  clexmc_handle *h = handle();
  SamplingFixtureParams const &P = run_manager.params;
  const int n = (int)states.size();
  std::vector<double> temps;
  _load_chains(kind, states, chem, run_manager, temps);
  auto t0 = std::chrono::steady_clock::now();
  double kms = 0.0, ms = 0.0;
  if (P.n_equilibration_passes > 0) {
    check(h, clexmc_run_passes(h, kind, (int)P.n_equilibration_passes, 1, P.canonical_rounds), calculator_name + "::run");
    clexmc_last_run_ms(h, &ms);
    kms += ms;
    check(h, clexmc_clear_samples(h), calculator_name + "::run");
    // sampling starts from energies recomputed in FP64 (table-path CEs accumulate FP32 dE)
    check(h, clexmc_reanchor_energies(h, nullptr), calculator_name + "::run");
  }
  // "corr.<key>" sampler (sampling_functions.cc:121-135): launches end at the
  // sampled passes (g + 1 a multiple of the period, g counted from the load)
  // and every chain's FP64 correlations per unit cell are read there
  auto wants = [&](const char *name) {
    return std::find(P.sampler_names.begin(), P.sampler_names.end(), name) != P.sampler_names.end();
  };
  // "corr.<key>" (all correlations) and "clex.<key>.sparse_corr" (those of the
  // non-zero ECIs, sampling_functions.cc:185-219) read the correlations
  const bool want_all_corr = wants("corr.formation_energy");
  const bool want_sparse = wants("clex.formation_energy.sparse_corr");
  const bool want_corr = want_all_corr || want_sparse;
  std::vector<std::vector<std::vector<double>>> corr(n);
  if (!want_corr) {
    check(h, clexmc_run_passes(h, kind, (int)P.n_passes, P.sample_period, P.canonical_rounds), calculator_name + "::run");
    clexmc_last_run_ms(h, &ms);
    kms += ms;
  } else {
    const long long per = std::max(1, P.sample_period);
    long long g = std::max<Index>(P.n_equilibration_passes, 0);
    const long long end = g + P.n_passes;
    std::vector<double> cv(system->n_corr);
    while (g < end) {
      const long long due = g + (per - 1 - g % per);
      const long long stop = std::min(end, due + 1);
      check(h, clexmc_run_passes(h, kind, (int)(stop - g), (int)per, P.canonical_rounds), calculator_name + "::run");
      clexmc_last_run_ms(h, &ms);
      kms += ms;
      g = stop;
      if (stop != due + 1) continue;
      for (int c = 0; c < n; ++c) {
        check(h, clexmc_correlations(h, c, cv.data()), calculator_name + "::run (corr)");
        const double nu = (double)states[c]->configuration.n_unitcells();
        std::vector<double> row(cv.size());
        for (size_t i = 0; i < cv.size(); ++i) row[i] = cv[i] / nu;
        corr[c].push_back(std::move(row));
      }
    }
  }
  check(h, clexmc_reanchor_energies(h, nullptr), calculator_name + "::run");
  double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  int ns = 0;
  check(h, clexmc_read_samples(h, &ns, nullptr, nullptr), calculator_name + "::run (samples)");
  std::vector<double> E((size_t)n * std::max(ns, 1));
  std::vector<int32_t> cnt((size_t)n * std::max(ns, 1) * 4);
  int ns2 = std::max(ns, 1);
  check(h, clexmc_read_samples(h, &ns2, E.data(), cnt.data()), calculator_name + "::run (samples)");
  results.assign(n, Results());
  const auto &sys = *system;
  for (int c = 0; c < n; ++c) {
    State &st = *states[c];
    Results &R = results[c];
    const Index nu = st.configuration.n_unitcells();
    const double T = temps[c];
    std::vector<double> mu;
    if (kind == 1) mu = required_vector(st, "param_chem_pot");
    // fixed-sublattice component counts
    std::vector<double> fixed(sys.n_components, 0.0);
    for (int b = 0; b < sys.n_sublat; ++b)
      if (sys.n_occ[b] < 2) fixed[sys.occ_component[b * sys.n_occ_max]] += 1.0;
    // variable sublattices share their occupant list (as in all supported systems)
    int bvar = 0;
    while (bvar < sys.n_sublat && sys.n_occ[bvar] < 2) ++bvar;
    for (int q = 0; q < ns; ++q) {
      const double e = E[(size_t)c * ns + q] / (double)nu;
      std::vector<double> mol(fixed);
      for (int o = 0; o < sys.n_occ_max && bvar < sys.n_sublat; ++o) {
        int comp = sys.occ_component[bvar * sys.n_occ_max + o];
        if (comp >= 0) mol[comp] += (double)cnt[((size_t)c * ns + q) * 4 + o] / (double)nu;
      }
      auto x = sys.param_composition(mol);
      double pe = e;
      if (kind == 1)
        for (size_t a = 0; a < x.size(); ++a) pe -= mu[a] * x[a];
      R.samplers["clex.formation_energy"].push_back({e});
      R.samplers["potential_energy"].push_back({pe});
      R.samplers["mol_composition"].push_back(mol);
      R.samplers["param_composition"].push_back(x);
      R.samplers["temperature"].push_back({T});
      if (kind == 1) R.samplers["param_chem_pot"].push_back(mu);
    }
    if (want_corr) {
      if ((int)corr[c].size() < ns) throw std::runtime_error("Error in " + calculator_name + "::run: corr samples");
      if (want_all_corr) R.samplers["corr.formation_energy"].assign(corr[c].begin(), corr[c].begin() + ns);
      if (want_sparse) {
        auto &S = R.samplers["clex.formation_energy.sparse_corr"];
        for (int q = 0; q < ns; ++q) {
          std::vector<double> row;
          for (int i = 0; i < sys.n_corr; ++i)
            if (sys.eci[i] != 0.0) row.push_back(corr[c][q][i]);
          S.push_back(std::move(row));
        }
      }
    }
    // statistics (mean, autocorrelation-corrected standard error)
    for (auto const &kv : R.samplers) {
      if (kv.second.empty()) continue;
      for (size_t i = 0; i < kv.second[0].size(); ++i) {
        auto col = column(kv.second, i);
        double m = mean_of(col), v = cov_of(col, col);
        double err = std::sqrt(std::max(0.0, v) * 2.0 * tau_int(col) / col.size());
        R.statistics[kv.first + "(" + std::to_string(i) + ")"] = {m, err};
      }
    }
    // analysis functions (analysis_functions.cc:17-173)
    auto pe = ns ? column(R.samplers["potential_energy"], 0) : std::vector<double>();
    R.analysis["heat_capacity"] = {cov_of(pe, pe) * (double)nu / (KB * T * T)};
    if (kind == 1 && ns) {
      auto const &molS = R.samplers["mol_composition"];
      auto const &xS = R.samplers["param_composition"];
      std::vector<double> mol_susc, param_susc, mol_tc, param_tc;
      for (size_t j = 0; j < molS[0].size(); ++j)
        for (size_t i = 0; i < molS[0].size(); ++i)
          mol_susc.push_back(cov_of(column(molS, i), column(molS, j)) * (double)nu / (KB * T));
      for (size_t j = 0; j < xS[0].size(); ++j)
        for (size_t i = 0; i < xS[0].size(); ++i)
          param_susc.push_back(cov_of(column(xS, i), column(xS, j)) * (double)nu / (KB * T));
      for (size_t j = 0; j < molS[0].size(); ++j) mol_tc.push_back(cov_of(pe, column(molS, j)) * (double)nu / (KB * T));
      for (size_t j = 0; j < xS[0].size(); ++j) param_tc.push_back(cov_of(pe, column(xS, j)) * (double)nu / (KB * T));
      R.analysis["mol_susc"] = mol_susc;
      R.analysis["param_susc"] = param_susc;
      R.analysis["mol_thermochem_susc"] = mol_tc;
      R.analysis["param_thermochem_susc"] = param_tc;
    }
    // final state back into the caller's configuration
    std::vector<int8_t> o(st.configuration.occupation.size());
    double Er;
    long long na, nacc;
    check(h, clexmc_read_state(h, c, o.data(), &Er, &na, &nacc), calculator_name + "::run (state)");
    st.configuration.occupation.assign(o.begin(), o.end());
    R.n_accept = nacc;
    R.n_reject = na - nacc;
    R.elapsed_s = elapsed;
    R.kernel_ms = kms;
  }
}

std::shared_ptr<MonteCalculator> make_monte_calculator(std::map<std::string, std::string> const &params,
                                                       std::shared_ptr<System> system,
                                                       std::unique_ptr<BaseMonteCalculator> base) {
  base->reset(params, system);
  return std::make_shared<MonteCalculator>(std::move(base));
}

// ---------------------------------------------------------------------------
// GpuCanonicalCalculator
// ---------------------------------------------------------------------------
std::vector<std::string> GpuCanonicalCalculator::standard_sampling_function_names() const {
  return {"temperature",           "mol_composition", "param_composition", "clex.formation_energy",
          "clex.formation_energy.sparse_corr", "corr.formation_energy", "potential_energy"};
}
std::vector<std::string> GpuCanonicalCalculator::standard_analysis_function_names() const { return {"heat_capacity"}; }

SamplingFixtureParams GpuCanonicalCalculator::make_default_sampling_fixture_params() const {
  SamplingFixtureParams p;
  p.sampler_names = {"clex.formation_energy", "potential_energy", "mol_composition", "param_composition"};
  return p;
}

void GpuCanonicalCalculator::_reset() {
  auto it = params.find("mol_composition_tol");
  mol_composition_tol = it == params.end() ? 1e-5 : std::stod(it->second);
}

std::vector<std::string> GpuCanonicalCalculator::validate_conditions(State const &state) const {
  std::vector<std::string> err;
  if (!state.conditions.scalar_values.count("temperature")) err.push_back("Missing required scalar condition 'temperature'");
  return err;
}

std::vector<std::string> GpuCanonicalCalculator::validate_state(State const &state) const {
  auto err = validate_conditions(state);
  if (!err.empty()) return err;
  auto mol = system->mol_composition(state.configuration);
  auto it = state.conditions.vector_values.find("mol_composition");
  if (it != state.conditions.vector_values.end()) {
    if (it->second.size() != mol.size()) {
      err.push_back("mol_composition condition has the wrong size");
    } else {
      for (size_t i = 0; i < mol.size(); ++i)
        if (std::fabs(mol[i] - it->second[i]) > mol_composition_tol)
          err.push_back("Calculated composition is not consistent with conditions composition.");
    }
  }
  auto jt = state.conditions.vector_values.find("param_composition");
  if (jt != state.conditions.vector_values.end()) {
    auto x = system->param_composition(mol);
    for (size_t i = 0; i < x.size() && i < jt->second.size(); ++i)
      if (std::fabs(x[i] - jt->second[i]) > mol_composition_tol)
        err.push_back("Calculated param_composition is not consistent with conditions.");
  }
  return err;
}

void GpuCanonicalCalculator::set_state_and_potential(State &state, OccLocation *occ_location) {
  if (!system) throw std::runtime_error("Error in GpuCanonicalCalculator::run: system==nullptr");
  auto err = validate_state(state);
  if (!err.empty()) throw std::runtime_error("Error in GpuCanonicalCalculator::run: Invalid initial state: " + err[0]);
  clexmc_handle *h = handle();
  std::array<int, 3> key = {state.configuration.L[0], state.configuration.L[1], state.configuration.L[2]};
  auto it = m_sc_ids.find(key);
  if (it == m_sc_ids.end()) {
    clexmc_supercell_desc d;
    std::memset(&d, 0, sizeof(d));
    d.L[0] = key[0];
    d.L[1] = key[1];
    d.L[2] = key[2];
    d.n_pair_phases = 8;
    int id = -1;
    check(h, clexmc_add_supercell(h, &d, &id), "GpuCanonicalCalculator::set_state_and_potential");
    it = m_sc_ids.emplace(key, id).first;
  }
  std::vector<int8_t> occ(state.configuration.occupation.begin(), state.configuration.occupation.end());
  double T = required_scalar(state, "temperature");
  uint64_t seed = 1;
  int id = it->second;
  check(h, clexmc_load_chains(h, 1, &id, occ.data(), &T, nullptr, &seed, KB),
        "GpuCanonicalCalculator::set_state_and_potential");
  auto pot = std::make_shared<GpuPotential>(h, system, state.configuration.n_unitcells(), false, std::vector<double>());
  pot->m_L[0] = key[0];
  pot->m_L[1] = key[1];
  pot->m_L[2] = key[2];
  pot->m_occ_size = occ.size();
  potential = pot;
  if (occ_location) occ_location->initialize(*system, state.configuration);
}

void GpuCanonicalCalculator::run(State &state, OccLocation &occ_location, RunManager &run_manager) {
  set_state_and_potential(state, &occ_location);
  std::vector<State *> v = {&state};
  std::vector<std::vector<double>> chem(1, std::vector<double>((size_t)system->n_sublat * system->n_occ_max, 0.0));
  std::vector<Results> res;
  _run_chains(0, v, chem, run_manager, res);
  run_manager.results = res[0];
  occ_location.initialize(*system, state.configuration);
}

void GpuCanonicalCalculator::run(int current_state, std::vector<State> &states, std::vector<OccLocation> &occ_locations,
                                 RunManager &run_manager) {
  std::vector<State *> v;
  for (auto &s : states) {
    auto err = validate_state(s);
    if (!err.empty()) throw std::runtime_error("Error in GpuCanonicalCalculator::run: Invalid state: " + err[0]);
    v.push_back(&s);
  }
  std::vector<std::vector<double>> chem(states.size(),
                                       std::vector<double>((size_t)system->n_sublat * system->n_occ_max, 0.0));
  _run_chains(0, v, chem, run_manager, run_manager.multistate_results);
  if (current_state >= 0 && current_state < (int)states.size()) run_manager.results = run_manager.multistate_results[current_state];
  occ_locations.resize(states.size());
  for (size_t i = 0; i < states.size(); ++i) occ_locations[i].initialize(*system, states[i].configuration);
}

// ---------------------------------------------------------------------------
// GpuSemiGrandCanonicalCalculator
// ---------------------------------------------------------------------------
std::vector<std::string> GpuSemiGrandCanonicalCalculator::standard_sampling_function_names() const {
  return {"temperature",           "mol_composition",       "param_composition",
          "param_chem_pot",        "clex.formation_energy", "clex.formation_energy.sparse_corr",
          "corr.formation_energy", "potential_energy"};
}
std::vector<std::string> GpuSemiGrandCanonicalCalculator::standard_analysis_function_names() const {
  return {"heat_capacity", "mol_susc", "param_susc", "mol_thermochem_susc", "param_thermochem_susc"};
}

SamplingFixtureParams GpuSemiGrandCanonicalCalculator::make_default_sampling_fixture_params() const {
  SamplingFixtureParams p;
  p.sampler_names = {"clex.formation_energy", "potential_energy", "mol_composition", "param_composition"};
  return p;
}

// Multiswap types from every multiset of 2 .. max_total single semi-grand
// canonical changes (unit, occupant from, occupant to) of the system's
// mutable asymmetric units, at most 4 distinct changes and 8 sites each.
static void make_multiswaps(System const &sys, int max_total, std::vector<int> &begin, std::vector<int> &parts) {
  const clexmc_system_desc d = sys.desc();
  std::map<int, int> unit_occ;  // unit -> occupants
  for (int b = 0; b < sys.n_sublat; ++b) {
    if (sys.n_occ[b] < 2) continue;
    const int u = d.asym ? d.asym[b] : sys.n_occ[b];
    unit_occ[u] = std::max(unit_occ[u], sys.n_occ[b]);
  }
  std::vector<std::array<int, 3>> singles;
  for (auto const &kv : unit_occ)
    for (int a = 0; a < kv.second; ++a)
      for (int c = 0; c < kv.second; ++c)
        if (a != c) singles.push_back({kv.first, a, c});
  begin.assign(1, 0);
  parts.clear();
  const int ns = (int)singles.size();
  std::vector<int> idx;
  // multisets as non-decreasing index sequences
  std::function<void(int, int)> rec = [&](int start, int left) {
    if (left == 0) {
      if (idx.size() < 2) return;
      std::map<int, int> cnt;
      for (int i : idx) ++cnt[i];
      if (cnt.size() > 4 || idx.size() > 8) return;
      for (auto const &kv : cnt) {
        parts.insert(parts.end(), {singles[kv.first][0], singles[kv.first][1], singles[kv.first][2], kv.second});
      }
      begin.push_back((int)parts.size() / 4);
      return;
    }
    for (int i = start; i < ns; ++i) {
      idx.push_back(i);
      rec(i, left - 1);
      idx.pop_back();
    }
  };
  for (int n = 2; n <= max_total; ++n) rec(0, n);
}

void GpuSemiGrandCanonicalCalculator::_configure_handle(clexmc_handle *h) {
  auto it = params.find("semigrand_canonical_multiswaps");
  if (it == params.end()) return;
  if (calculator_name == "GpuNfoldCalculator")
    throw std::runtime_error("Error in GpuNfoldCalculator: the nfold method takes single-site changes only");
  const int max_total = std::stoi(it->second);
  if (max_total < 2 || max_total > 8)
    throw std::runtime_error("Error in " + calculator_name + ": semigrand_canonical_multiswaps must be 2..8");
  std::vector<int> begin, parts;
  make_multiswaps(*system, max_total, begin, parts);
  auto pp = params.find("multiswaps_per_pass");
  const int per_pass = pp == params.end() ? 0 : std::stoi(pp->second);
  check(h, clexmc_set_multiswaps(h, (int)begin.size() - 1, begin.data(), parts.data(), per_pass),
        calculator_name + " (multiswaps)");
}

std::vector<std::string> GpuSemiGrandCanonicalCalculator::validate_conditions(State const &state) const {
  std::vector<std::string> err;
  if (!state.conditions.scalar_values.count("temperature")) err.push_back("Missing required scalar condition 'temperature'");
  auto it = state.conditions.vector_values.find("param_chem_pot");
  if (it == state.conditions.vector_values.end())
    err.push_back("Missing required vector condition 'param_chem_pot'");
  else if ((int)it->second.size() != system->n_independent)
    err.push_back("param_chem_pot size error");
  return err;
}

std::vector<std::string> GpuSemiGrandCanonicalCalculator::validate_state(State const &state) const {
  return validate_conditions(state);
}

void GpuSemiGrandCanonicalCalculator::set_state_and_potential(State &state, OccLocation *occ_location) {
  if (!system) throw std::runtime_error("Error in GpuSemiGrandCanonicalCalculator::run: system==nullptr");
  auto err = validate_state(state);
  if (!err.empty())
    throw std::runtime_error("Error in GpuSemiGrandCanonicalCalculator::run: Invalid initial state: " + err[0]);
  clexmc_handle *h = handle();
  std::array<int, 3> key = {state.configuration.L[0], state.configuration.L[1], state.configuration.L[2]};
  auto it = m_sc_ids.find(key);
  if (it == m_sc_ids.end()) {
    clexmc_supercell_desc d;
    std::memset(&d, 0, sizeof(d));
    d.L[0] = key[0];
    d.L[1] = key[1];
    d.L[2] = key[2];
    int id = -1;
    check(h, clexmc_add_supercell(h, &d, &id), "GpuSemiGrandCanonicalCalculator::set_state_and_potential");
    it = m_sc_ids.emplace(key, id).first;
  }
  std::vector<int8_t> occ(state.configuration.occupation.begin(), state.configuration.occupation.end());
  double T = required_scalar(state, "temperature");
  auto mu = required_vector(state, "param_chem_pot");
  auto w = system->chem_weights(mu);
  uint64_t seed = 1;
  int id = it->second;
  check(h, clexmc_load_chains(h, 1, &id, occ.data(), &T, w.data(), &seed, KB),
        "GpuSemiGrandCanonicalCalculator::set_state_and_potential");
  auto pot = std::make_shared<GpuPotential>(h, system, state.configuration.n_unitcells(), true, mu);
  pot->m_L[0] = key[0];
  pot->m_L[1] = key[1];
  pot->m_L[2] = key[2];
  pot->m_occ_size = occ.size();
  potential = pot;
  if (occ_location) occ_location->initialize(*system, state.configuration);
}

void GpuSemiGrandCanonicalCalculator::run(State &state, OccLocation &occ_location, RunManager &run_manager) {
  set_state_and_potential(state, &occ_location);
  std::vector<State *> v = {&state};
  std::vector<std::vector<double>> chem = {system->chem_weights(required_vector(state, "param_chem_pot"))};
  std::vector<Results> res;
  _run_chains(1, v, chem, run_manager, res);
  run_manager.results = res[0];
  occ_location.initialize(*system, state.configuration);
}

void GpuSemiGrandCanonicalCalculator::run(int current_state, std::vector<State> &states,
                                          std::vector<OccLocation> &occ_locations, RunManager &run_manager) {
  std::vector<State *> v;
  std::vector<std::vector<double>> chem;
  for (auto &s : states) {
    auto err = validate_state(s);
    if (!err.empty()) throw std::runtime_error("Error in GpuSemiGrandCanonicalCalculator::run: Invalid state: " + err[0]);
    v.push_back(&s);
    chem.push_back(system->chem_weights(required_vector(s, "param_chem_pot")));
  }
  _run_chains(1, v, chem, run_manager, run_manager.multistate_results);
  if (current_state >= 0 && current_state < (int)states.size()) run_manager.results = run_manager.multistate_results[current_state];
  occ_locations.resize(states.size());
  for (size_t i = 0; i < states.size(); ++i) occ_locations[i].initialize(*system, states[i].configuration);
}

// ---------------------------------------------------------------------------
// GpuNfoldCalculator
// ---------------------------------------------------------------------------
void BaseMonteCalculator::_run_nfold(std::vector<State *> const &states, std::vector<std::vector<double>> const &chem,
                                     RunManager &run_manager, std::vector<Results> &results) {
  const std::string where = calculator_name + "::run";
  SamplingFixtureParams const &P = run_manager.params;
  const auto &sys = *system;
  const int n = (int)states.size();
  if (n == 0) throw std::runtime_error("Error in " + where + ": no states");
  int n_var_sub = 0, bvar = -1;
  for (int b = 0; b < sys.n_sublat; ++b) {
    if (sys.n_occ[b] >= 2) {
      ++n_var_sub;
      if (bvar < 0) bvar = b;
    }
  }
  if (bvar < 0) throw std::runtime_error("Error in " + where + ": no variable sites");
  for (const char *name : {"corr.formation_energy", "clex.formation_energy.sparse_corr"})
    if (std::find(P.sampler_names.begin(), P.sampler_names.end(), name) != P.sampler_names.end())
      throw std::runtime_error("Error in " + where + ": nfold runs sample energies and compositions only");
  for (auto *st : states)
    if (!std::equal(st->configuration.L, st->configuration.L + 3, states[0]->configuration.L))
      throw std::runtime_error("Error in " + where + ": nfold states must share one supercell");
  const Index nu = states[0]->configuration.n_unitcells();
  const long long n_var = (long long)nu * n_var_sub;
  const long long n_events = (long long)P.n_passes * n_var;
  if (P.n_passes < 1 || n_events > 2147483647LL)
    throw std::runtime_error("Error in " + where + ": n_passes x mutable sites must lie in [1, 2^31)");
  clexmc_handle *h = handle();
  std::vector<double> temps;
  _load_chains(1, states, chem, run_manager, temps);
  auto t0 = std::chrono::steady_clock::now();
  double kms = 0.0, ms = 0.0;
  if (P.n_equilibration_passes > 0) {
    check(h, clexmc_run_passes(h, 1, (int)P.n_equilibration_passes, (int)P.n_equilibration_passes, 0), where);
    clexmc_last_run_ms(h, &ms);
    kms += ms;
  }
  std::vector<long long> att0(n);
  std::vector<int8_t> o(states[0]->configuration.occupation.size());
  double Er;
  long long na, nacc;
  for (int c = 0; c < n; ++c) {
    check(h, clexmc_read_state(h, c, o.data(), &Er, &na, &nacc), where + " (state)");
    att0[c] = na;
  }
  const int nb = (int)P.n_passes;
  constexpr int M = 21;  // clexmc_nfold_run_moments record: W, wE, wE^2, wN_o, wN_oN_q (o <= q), wEN_o
  std::vector<double> blocks((size_t)n * nb * M);
  check(h, clexmc_nfold_run_moments(h, (int)n_events, (int)n_var, blocks.data()), where + " (nfold)");
  double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // compositions as affine functions of the occupant counts N_o (sites of the
  // variable sublattices holding occupant o)
  std::vector<double> fixed(sys.n_components, 0.0);
  for (int b = 0; b < sys.n_sublat; ++b)
    if (sys.n_occ[b] < 2) fixed[sys.occ_component[b * sys.n_occ_max]] += 1.0;
  auto mol_of = [&](const double (&N)[4]) {
    std::vector<double> mol(fixed);
    for (int q = 0; q < sys.n_occ[bvar]; ++q) {
      int comp = sys.occ_component[bvar * sys.n_occ_max + q];
      if (comp >= 0) mol[comp] += N[q] / (double)nu;
    }
    return mol;
  };
  const double zero[4] = {0.0, 0.0, 0.0, 0.0};
  const std::vector<double> m0 = mol_of(zero), x0 = sys.param_composition(m0);
  std::vector<std::vector<double>> dm(4), dx(4);  // change per unit of N_o
  for (int o = 0; o < 4; ++o) {
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    e[o] = 1.0;
    const std::vector<double> mo = mol_of(e), xo = sys.param_composition(mo);
    for (size_t i = 0; i < m0.size(); ++i) dm[o].push_back(mo[i] - m0[i]);
    for (size_t i = 0; i < x0.size(); ++i) dx[o].push_back(xo[i] - x0[i]);
  }
  results.assign(n, Results());
  for (int c = 0; c < n; ++c) {
    State &st = *states[c];
    Results &R = results[c];
    const double T = temps[c];
    const auto mu = required_vector(st, "param_chem_pot");
    // q = a + b E + g . N for every sampled component (E per supercell)
    struct Aff {
      double a, b, g[4];
    };
    std::map<std::string, std::vector<Aff>> aff;
    Aff pe{0.0, 1.0 / (double)nu, {0.0, 0.0, 0.0, 0.0}};
    for (size_t i = 0; i < x0.size(); ++i) {
      pe.a -= mu[i] * x0[i];
      for (int o = 0; o < 4; ++o) pe.g[o] -= mu[i] * dx[o][i];
    }
    aff["clex.formation_energy"] = {Aff{0.0, 1.0 / (double)nu, {0.0, 0.0, 0.0, 0.0}}};
    aff["potential_energy"] = {pe};
    for (size_t i = 0; i < m0.size(); ++i)
      aff["mol_composition"].push_back(Aff{m0[i], 0.0, {dm[0][i], dm[1][i], dm[2][i], dm[3][i]}});
    for (size_t i = 0; i < x0.size(); ++i)
      aff["param_composition"].push_back(Aff{x0[i], 0.0, {dx[0][i], dx[1][i], dx[2][i], dx[3][i]}});
    const double *bk = blocks.data() + (size_t)c * nb * M;
    double S[M] = {};
    std::vector<double> Eb(nb);
    std::vector<std::array<double, 4>> Nb(nb);
    for (int k = 0; k < nb; ++k) {
      for (int j = 0; j < M; ++j) S[j] += bk[k * M + j];
      Eb[k] = bk[k * M + 1] / bk[k * M];
      for (int o = 0; o < 4; ++o) Nb[k][o] = bk[k * M + 3 + o] / bk[k * M];
    }
    const double E1 = S[1] / S[0];
    double N1[4], vEN[4], vNN[4][4];
    for (int o = 0; o < 4; ++o) N1[o] = S[3 + o] / S[0];
    const double vEE = S[2] / S[0] - E1 * E1;
    for (int o = 0; o < 4; ++o) vEN[o] = S[17 + o] / S[0] - E1 * N1[o];
    for (int o = 0, t = 0; o < 4; ++o)
      for (int q = o; q < 4; ++q, ++t) vNN[o][q] = vNN[q][o] = S[7 + t] / S[0] - N1[o] * N1[q];
    auto mean = [&](Aff const &p) {
      double v = p.a + p.b * E1;
      for (int o = 0; o < 4; ++o) v += p.g[o] * N1[o];
      return v;
    };
    auto cov = [&](Aff const &p, Aff const &q) {
      double v = p.b * q.b * vEE;
      for (int o = 0; o < 4; ++o) v += (p.b * q.g[o] + p.g[o] * q.b) * vEN[o];
      for (int o = 0; o < 4; ++o)
        for (int r = 0; r < 4; ++r) v += p.g[o] * q.g[r] * vNN[o][r];
      return v;
    };
    // one sample per pass: the pass's residence-weighted means
    for (auto const &kv : aff) {
      auto &ser = R.samplers[kv.first];
      ser.assign(nb, std::vector<double>(kv.second.size()));
      for (int k = 0; k < nb; ++k)
        for (size_t i = 0; i < kv.second.size(); ++i) {
          const Aff &f = kv.second[i];
          double v = f.a + f.b * Eb[k];
          for (int o = 0; o < 4; ++o) v += f.g[o] * Nb[k][o];
          ser[k][i] = v;
        }
      for (size_t i = 0; i < kv.second.size(); ++i) {
        auto col = column(ser, i);
        double v = cov_of(col, col);
        double err = std::sqrt(std::max(0.0, v) * 2.0 * tau_int(col) / col.size());
        R.statistics[kv.first + "(" + std::to_string(i) + ")"] = {mean(kv.second[i]), err};
      }
    }
    R.samplers["temperature"].assign(nb, {T});
    R.samplers["param_chem_pot"].assign(nb, mu);
    // analysis functions (analysis_functions.cc:17-173) on the weighted moments
    const auto &Mc = aff["mol_composition"];
    const auto &X = aff["param_composition"];
    const double kT = KB * T;
    R.analysis["heat_capacity"] = {cov(pe, pe) * (double)nu / (kT * T)};
    std::vector<double> mol_susc, param_susc, mol_tc, param_tc;
    for (size_t j = 0; j < Mc.size(); ++j)
      for (size_t i = 0; i < Mc.size(); ++i) mol_susc.push_back(cov(Mc[i], Mc[j]) * (double)nu / kT);
    for (size_t j = 0; j < X.size(); ++j)
      for (size_t i = 0; i < X.size(); ++i) param_susc.push_back(cov(X[i], X[j]) * (double)nu / kT);
    for (size_t j = 0; j < Mc.size(); ++j) mol_tc.push_back(cov(pe, Mc[j]) * (double)nu / kT);
    for (size_t j = 0; j < X.size(); ++j) param_tc.push_back(cov(pe, X[j]) * (double)nu / kT);
    R.analysis["mol_susc"] = mol_susc;
    R.analysis["param_susc"] = param_susc;
    R.analysis["mol_thermochem_susc"] = mol_tc;
    R.analysis["param_thermochem_susc"] = param_tc;
    check(h, clexmc_read_state(h, c, o.data(), &Er, &na, &nacc), where + " (state)");
    st.configuration.occupation.assign(o.begin(), o.end());
    R.n_accept = na - att0[c];  // every event changes the state
    R.n_reject = 0;
    R.elapsed_s = elapsed;
    R.kernel_ms = kms;
  }
}

void GpuNfoldCalculator::run(State &state, OccLocation &occ_location, RunManager &run_manager) {
  set_state_and_potential(state, &occ_location);
  std::vector<State *> v = {&state};
  std::vector<std::vector<double>> chem = {system->chem_weights(required_vector(state, "param_chem_pot"))};
  std::vector<Results> res;
  _run_nfold(v, chem, run_manager, res);
  run_manager.results = res[0];
  occ_location.initialize(*system, state.configuration);
}

void GpuNfoldCalculator::run(int current_state, std::vector<State> &states, std::vector<OccLocation> &occ_locations,
                             RunManager &run_manager) {
  std::vector<State *> v;
  std::vector<std::vector<double>> chem;
  for (auto &s : states) {
    auto err = validate_state(s);
    if (!err.empty()) throw std::runtime_error("Error in GpuNfoldCalculator::run: Invalid state: " + err[0]);
    v.push_back(&s);
    chem.push_back(system->chem_weights(required_vector(s, "param_chem_pot")));
  }
  _run_nfold(v, chem, run_manager, run_manager.multistate_results);
  if (current_state >= 0 && current_state < (int)states.size()) run_manager.results = run_manager.multistate_results[current_state];
  occ_locations.resize(states.size());
  for (size_t i = 0; i < states.size(); ++i) occ_locations[i].initialize(*system, states[i].configuration);
}

}  // namespace gpu
}  // namespace clexmonte
}  // namespace CASM

extern "C" {

CASM::clexmonte::gpu::BaseMonteCalculator *make_GpuCanonicalCalculator() {
  return new CASM::clexmonte::gpu::GpuCanonicalCalculator();
}

CASM::clexmonte::gpu::BaseMonteCalculator *make_GpuSemiGrandCanonicalCalculator() {
  return new CASM::clexmonte::gpu::GpuSemiGrandCanonicalCalculator();
}

CASM::clexmonte::gpu::BaseMonteCalculator *make_GpuNfoldCalculator() {
  return new CASM::clexmonte::gpu::GpuNfoldCalculator();
}
}
