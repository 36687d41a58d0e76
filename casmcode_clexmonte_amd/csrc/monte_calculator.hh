// C++ host layer mirroring the reference's Monte Carlo calculator plugin
// surface, implemented on top of the clexmc C ABI (include/clexmc.h).
//
// Mirrored reference interfaces (paths relative to the reference root):
//  - BaseMontePotential: include/casm/clexmonte/monte_calculator/
//      BaseMonteCalculator.hh:26-44 (per_supercell, per_unitcell,
//      occ_delta_per_supercell);
//  - BaseMonteCalculator: BaseMonteCalculator.hh:47-264 (reset,
//      set_state_and_potential, run, multi-state run, validate_*,
//      standard sampling / analysis function names, clone);
//  - MonteCalculator wrapper: include/casm/clexmonte/monte_calculator/
//      MonteCalculator.hh:44-403;
//  - extern "C" make_<Name>() factories: src/casm/clexmonte/monte_calculator/
//      CanonicalCalculator.cc:580-586, SemiGrandCanonicalCalculator.cc:517-523;
//  - conditions: include/casm/clexmonte/state/Conditions.hh:17-50,358-405.
// libcasm types (Eigen, jsonParser, RunManager, OccLocation, ...) are not
// available here; the types below keep the reference's names and argument
// meaning with plain C++ containers.
#pragma once

#include <array>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/clexmc.h"

namespace CASM {
namespace clexmonte {
namespace gpu {

typedef long Index;

/// Boltzmann constant (eV/K) used by CASM (CASM::KB)
constexpr double KB = 8.6173303e-05;

/// Conditions: scalar and vector values ("temperature", "param_chem_pot",
/// "mol_composition", "param_composition"), as monte::ValueMap.
struct ValueMap {
  std::map<std::string, double> scalar_values;
  std::map<std::string, std::vector<double>> vector_values;
};

/// Configuration of a diagonal supercell L1 x L2 x L3 with occupant indices
/// per site, linear index l = b * n_cells + i + L1 * (j + L2 * k).
struct Configuration {
  int L[3] = {1, 1, 1};
  std::vector<int> occupation;
  Index n_unitcells() const { return (Index)L[0] * L[1] * L[2]; }
};

/// Monte Carlo state: configuration + conditions (monte::State)
struct State {
  Configuration configuration;
  ValueMap conditions;
};

/// Flattened cluster expansion and composition axes (the role of the
/// reference's System, src/casm/clexmonte/system/System.cc:41-72).
class System {
 public:
  /// Read the text file written by casmcode_clexmonte_amd.system.save_flat
  static std::shared_ptr<System> from_flat_file(std::string const &path);

  clexmc_system_desc desc() const;

  int n_sublat = 0, n_occ_max = 0, n_components = 0, n_independent = 0;
  std::vector<int> n_occ;
  std::vector<int> occ_component;   // [n_sublat][n_occ_max], -1 unused
  std::vector<double> comp_origin;  // [n_components]
  std::vector<double> comp_R;       // [n_independent][n_components] (end - origin)
  std::vector<double> eci;
  int n_corr = 0, n_orbits = 0;
  std::vector<int> orbit_func_begin, orbit_mult;

  /// number of each component per unit cell (mean_num_each_component)
  std::vector<double> mol_composition(Configuration const &config) const;
  /// parametric composition x = R^+ (n - n0)
  std::vector<double> param_composition(std::vector<double> const &mol) const;
  /// per (sublattice, occupant) semi-grand weights w = mu^T R^+
  std::vector<double> chem_weights(std::vector<double> const &param_chem_pot) const;

 private:
  std::map<std::string, std::vector<double>> m_d;
  std::map<std::string, std::vector<int>> m_i;
  std::vector<double> m_Rpinv;  // [n_independent][n_components]
  void _finalize();
};

/// Occupant location tracker (monte::OccLocation): here only the quantities
/// the Metropolis loop needs on the host side (number of mutable sites =
/// steps per pass, occupant counts).
class OccLocation {
 public:
  void initialize(System const &system, Configuration const &config);
  Index mol_size() const { return m_mol_size; }
  std::vector<Index> const &counts() const { return m_counts; }

 private:
  Index m_mol_size = 0;
  std::vector<Index> m_counts;
};

/// Sampling fixture parameters and results (the parts of monte::RunManager /
/// SamplingFixture used by the calculators).  Completion: a fixed number of
/// equilibration and sampled passes.
struct SamplingFixtureParams {
  std::vector<std::string> sampler_names;
  Index n_equilibration_passes = 0;
  Index n_passes = 1000;
  int sample_period = 1;
  int canonical_rounds = 0;  // 0: automatic (about one swap attempt per mutable site and pass)
  int n_pair_phases = 32;
};

struct Results {
  std::map<std::string, std::vector<std::vector<double>>> samplers;  // name -> samples x components
  std::map<std::string, std::vector<double>> analysis;
  std::map<std::string, std::pair<double, double>> statistics;  // name(i) -> (mean, error)
  Index n_accept = 0, n_reject = 0;
  double elapsed_s = 0.0;
  double kernel_ms = 0.0;
};

struct RunManager {
  uint64_t seed = 0;  // key of the counter-based generator
  SamplingFixtureParams params;
  Results results;
  std::vector<Results> multistate_results;
};

class MonteCalculator;

/// \brief Implements a potential (BaseMonteCalculator.hh:26-44)
class BaseMontePotential {
 public:
  virtual ~BaseMontePotential() {}
  virtual double per_supercell() = 0;
  virtual double per_unitcell() = 0;
  virtual double occ_delta_per_supercell(std::vector<Index> const &linear_site_index,
                                         std::vector<int> const &new_occ) = 0;
};

/// \brief Calculator plugin base (BaseMonteCalculator.hh:47-264)
class BaseMonteCalculator {
 public:
  BaseMonteCalculator(std::string _calculator_name, bool _is_multistate_method)
      : calculator_name(_calculator_name), is_multistate_method(_is_multistate_method) {}
  virtual ~BaseMonteCalculator();

  std::string calculator_name;
  std::map<std::string, std::string> params;
  std::shared_ptr<System> system;
  bool time_sampling_allowed = false;
  bool update_atoms = false;
  bool save_atom_info = false;
  bool is_multistate_method = false;
  int device = 0;

  /// Set parameters and system, then reset the derived calculator
  void reset(std::map<std::string, std::string> const &_params, std::shared_ptr<System> _system);

  virtual std::vector<std::string> standard_sampling_function_names() const = 0;
  virtual std::vector<std::string> standard_analysis_function_names() const = 0;
  virtual SamplingFixtureParams make_default_sampling_fixture_params() const = 0;

  virtual std::vector<std::string> validate_conditions(State const &state) const = 0;
  virtual std::vector<std::string> validate_state(State const &state) const = 0;

  std::shared_ptr<BaseMontePotential> potential;

  /// Validate and set the current state, construct the potential
  virtual void set_state_and_potential(State &state, OccLocation *occ_location) = 0;

  /// Single run evolving `state` (occupation_metropolis_v2 driven by run)
  virtual void run(State &state, OccLocation &occ_location, RunManager &run_manager) = 0;

  /// Multi-state run: every state is an independent chain of one persistent
  /// launch (the reference declares this hook, BaseMonteCalculator.hh:245-247)
  virtual void run(int current_state, std::vector<State> &states, std::vector<OccLocation> &occ_locations,
                   RunManager &run_manager) = 0;

  std::unique_ptr<BaseMonteCalculator> clone() const { return std::unique_ptr<BaseMonteCalculator>(_clone()); }

 protected:
  /// one clexmc handle per calculator
  clexmc_handle *handle();
  void release_handle();
  // copies share no device handle (each clone creates its own on first use)
 public:
  BaseMonteCalculator(BaseMonteCalculator const &o)
      : calculator_name(o.calculator_name), params(o.params), system(o.system),
        time_sampling_allowed(o.time_sampling_allowed), update_atoms(o.update_atoms),
        save_atom_info(o.save_atom_info), is_multistate_method(o.is_multistate_method), device(o.device) {}
  BaseMonteCalculator &operator=(BaseMonteCalculator const &) = delete;

 protected:
  virtual void _reset() {}
  virtual BaseMonteCalculator *_clone() const = 0;
  /// method-specific settings of a newly created handle
  virtual void _configure_handle(clexmc_handle *) {}

  /// run n chains through the C ABI; kind 0 canonical, 1 semi-grand
  void _run_chains(int kind, std::vector<State *> const &states, std::vector<std::vector<double>> const &chem,
                   RunManager &run_manager, std::vector<Results> &results);
  /// load the states as chains of the handle (one per state); temperatures out
  void _load_chains(int kind, std::vector<State *> const &states, std::vector<std::vector<double>> const &chem,
                    RunManager const &run_manager, std::vector<double> &temps);
  /// N-fold way (rejection-free) semi-grand canonical runs with
  /// residence-time-weighted statistics; binary expansions, one supercell shape
  void _run_nfold(std::vector<State *> const &states, std::vector<std::vector<double>> const &chem,
                  RunManager &run_manager, std::vector<Results> &results);

  clexmc_handle *m_handle = nullptr;
  std::map<std::array<int, 3>, int> m_sc_ids;
};

void check(clexmc_handle *h, clexmc_status st, std::string const &what);

/// Canonical Metropolis on the GPU (CanonicalCalculator.cc:143-575)
class GpuCanonicalCalculator : public BaseMonteCalculator {
 public:
  GpuCanonicalCalculator() : BaseMonteCalculator("GpuCanonicalCalculator", false) {}
  std::vector<std::string> standard_sampling_function_names() const override;
  std::vector<std::string> standard_analysis_function_names() const override;
  SamplingFixtureParams make_default_sampling_fixture_params() const override;
  std::vector<std::string> validate_conditions(State const &state) const override;
  std::vector<std::string> validate_state(State const &state) const override;
  void set_state_and_potential(State &state, OccLocation *occ_location) override;
  void run(State &state, OccLocation &occ_location, RunManager &run_manager) override;
  void run(int current_state, std::vector<State> &states, std::vector<OccLocation> &occ_locations,
           RunManager &run_manager) override;
  double mol_composition_tol = 1e-5;

 protected:
  void _reset() override;
  BaseMonteCalculator *_clone() const override { return new GpuCanonicalCalculator(*this); }
};

/// Semi-grand canonical Metropolis on the GPU (SemiGrandCanonicalCalculator.cc:216-516)
class GpuSemiGrandCanonicalCalculator : public BaseMonteCalculator {
 public:
  GpuSemiGrandCanonicalCalculator() : BaseMonteCalculator("GpuSemiGrandCanonicalCalculator", false) {}
  explicit GpuSemiGrandCanonicalCalculator(std::string name) : BaseMonteCalculator(std::move(name), false) {}
  std::vector<std::string> standard_sampling_function_names() const override;
  std::vector<std::string> standard_analysis_function_names() const override;
  SamplingFixtureParams make_default_sampling_fixture_params() const override;
  std::vector<std::string> validate_conditions(State const &state) const override;
  std::vector<std::string> validate_state(State const &state) const override;
  void set_state_and_potential(State &state, OccLocation *occ_location) override;
  void run(State &state, OccLocation &occ_location, RunManager &run_manager) override;
  void run(int current_state, std::vector<State> &states, std::vector<OccLocation> &occ_locations,
           RunManager &run_manager) override;

 protected:
  BaseMonteCalculator *_clone() const override { return new GpuSemiGrandCanonicalCalculator(*this); }
  /// params "semigrand_canonical_multiswaps" (the largest number of single
  /// semi-grand canonical changes to combine, as the system's integer input,
  /// System_json_io.cc:974-983) and "multiswaps_per_pass": composition then
  /// changes through multiswaps only (clexmc_set_multiswaps)
  void _configure_handle(clexmc_handle *h) override;
};

/// N-fold way (rejection-free) semi-grand canonical runs on the GPU (the
/// reference's nfold method, include/casm/clexmonte/nfold/nfold_impl.hh:18-98):
/// the semi-grand canonical conditions, potential and samplers; equilibration
/// by Metropolis passes, then n_passes x mol_size events, one sample per pass
/// (the pass's residence-weighted means).  Statistics and analysis are
/// residence-time weighted.  Binary expansions; all states on one supercell.
class GpuNfoldCalculator : public GpuSemiGrandCanonicalCalculator {
 public:
  GpuNfoldCalculator() : GpuSemiGrandCanonicalCalculator("GpuNfoldCalculator") {}
  void run(State &state, OccLocation &occ_location, RunManager &run_manager) override;
  void run(int current_state, std::vector<State> &states, std::vector<OccLocation> &occ_locations,
           RunManager &run_manager) override;

 protected:
  BaseMonteCalculator *_clone() const override { return new GpuNfoldCalculator(*this); }
};

/// MonteCalculator wrapper (MonteCalculator.hh:44-403)
class MonteCalculator {
 public:
  MonteCalculator(std::unique_ptr<BaseMonteCalculator> base) : m_calc(std::move(base)) {}
  std::string const &calculator_name() const { return m_calc->calculator_name; }
  std::shared_ptr<System> system() const { return m_calc->system; }
  BaseMontePotential &potential() {
    if (!m_calc->potential) throw std::runtime_error("Error in MonteCalculator::potential: not set");
    return *m_calc->potential;
  }
  void set_state_and_potential(State &state, OccLocation *occ_location) {
    m_calc->set_state_and_potential(state, occ_location);
  }
  void run(State &state, OccLocation &occ_location, RunManager &run_manager) {
    m_calc->run(state, occ_location, run_manager);
  }
  void run(int current_state, std::vector<State> &states, std::vector<OccLocation> &occ_locations,
           RunManager &run_manager) {
    m_calc->run(current_state, states, occ_locations, run_manager);
  }
  BaseMonteCalculator &base() { return *m_calc; }

 private:
  std::unique_ptr<BaseMonteCalculator> m_calc;
};

/// make_monte_calculator: reset a calculator with params and system
std::shared_ptr<MonteCalculator> make_monte_calculator(std::map<std::string, std::string> const &params,
                                                       std::shared_ptr<System> system,
                                                       std::unique_ptr<BaseMonteCalculator> base);

}  // namespace gpu
}  // namespace clexmonte
}  // namespace CASM

extern "C" {
/// Factories with the reference's plugin convention (CanonicalCalculator.cc:580-586)
CASM::clexmonte::gpu::BaseMonteCalculator *make_GpuCanonicalCalculator();
CASM::clexmonte::gpu::BaseMonteCalculator *make_GpuSemiGrandCanonicalCalculator();
CASM::clexmonte::gpu::BaseMonteCalculator *make_GpuNfoldCalculator();
}
