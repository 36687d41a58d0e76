// Python bindings (pybind11) of the C++ calculator plugin layer
// (monte_calculator.hh): the role of the reference's
// python/src/clexmonte_monte_calculator.cpp, which exposes BaseMonteCalculator
// instances made by the extern "C" factories (:34-46, :140-153), builds a
// MonteCalculator from a cloned instance (:155-173) and runs it
// (:216-269), and of MontePotential (:694-750).
//
// Module: casmcode_clexmonte_amd._clexmc_plugin (linked against libclexmc.so).
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "monte_calculator.hh"

namespace py = pybind11;
using namespace CASM::clexmonte::gpu;

PYBIND11_MODULE(_clexmc_plugin, m) {
  m.doc() = "C++ GPU Monte Carlo calculators (GpuCanonicalCalculator, GpuSemiGrandCanonicalCalculator, GpuNfoldCalculator)";
  m.attr("KB") = KB;

  py::class_<System, std::shared_ptr<System>>(m, "System")
      .def_static("from_flat_file", &System::from_flat_file, py::arg("path"))
      .def_readonly("n_sublat", &System::n_sublat)
      .def_readonly("n_occ", &System::n_occ)
      .def_readonly("n_components", &System::n_components)
      .def_readonly("n_independent", &System::n_independent)
      .def_readonly("n_corr", &System::n_corr)
      .def_readonly("eci", &System::eci)
      .def("mol_composition", &System::mol_composition, py::arg("configuration"))
      .def("param_composition", &System::param_composition, py::arg("mol_composition"))
      .def("chem_weights", &System::chem_weights, py::arg("param_chem_pot"));

  py::class_<ValueMap>(m, "ValueMap")
      .def(py::init<>())
      .def_readwrite("scalar_values", &ValueMap::scalar_values)
      .def_readwrite("vector_values", &ValueMap::vector_values);

  py::class_<Configuration>(m, "Configuration")
      .def(py::init<>())
      .def_property(
          "L", [](Configuration const &c) { return std::vector<int>{c.L[0], c.L[1], c.L[2]}; },
          [](Configuration &c, std::vector<int> const &L) {
            if (L.size() != 3) throw std::runtime_error("Configuration.L needs three values");
            for (int i = 0; i < 3; ++i) c.L[i] = L[i];
          })
      .def_readwrite("occupation", &Configuration::occupation)
      .def("n_unitcells", &Configuration::n_unitcells);

  py::class_<State>(m, "State")
      .def(py::init<>())
      .def_readwrite("configuration", &State::configuration)
      .def_readwrite("conditions", &State::conditions);

  py::class_<OccLocation>(m, "OccLocation")
      .def(py::init<>())
      .def("initialize", &OccLocation::initialize, py::arg("system"), py::arg("configuration"))
      .def("mol_size", &OccLocation::mol_size)
      .def("counts", &OccLocation::counts);

  py::class_<SamplingFixtureParams>(m, "SamplingFixtureParams")
      .def(py::init<>())
      .def_readwrite("sampler_names", &SamplingFixtureParams::sampler_names)
      .def_readwrite("n_equilibration_passes", &SamplingFixtureParams::n_equilibration_passes)
      .def_readwrite("n_passes", &SamplingFixtureParams::n_passes)
      .def_readwrite("sample_period", &SamplingFixtureParams::sample_period)
      .def_readwrite("canonical_rounds", &SamplingFixtureParams::canonical_rounds)
      .def_readwrite("n_pair_phases", &SamplingFixtureParams::n_pair_phases);

  py::class_<Results>(m, "Results")
      .def(py::init<>())
      .def_readonly("samplers", &Results::samplers)
      .def_readonly("analysis", &Results::analysis)
      .def_readonly("statistics", &Results::statistics)
      .def_readonly("n_accept", &Results::n_accept)
      .def_readonly("n_reject", &Results::n_reject)
      .def_readonly("elapsed_s", &Results::elapsed_s)
      .def_readonly("kernel_ms", &Results::kernel_ms);

  py::class_<RunManager>(m, "RunManager")
      .def(py::init<>())
      .def_readwrite("seed", &RunManager::seed)
      .def_readwrite("params", &RunManager::params)
      .def_readonly("results", &RunManager::results)
      .def_readonly("multistate_results", &RunManager::multistate_results);

  py::class_<BaseMontePotential, std::shared_ptr<BaseMontePotential>>(m, "MontePotential")
      .def("per_supercell", &BaseMontePotential::per_supercell)
      .def("per_unitcell", &BaseMontePotential::per_unitcell)
      .def("occ_delta_per_supercell", &BaseMontePotential::occ_delta_per_supercell, py::arg("linear_site_index"),
           py::arg("new_occ"));

  py::class_<BaseMonteCalculator, std::shared_ptr<BaseMonteCalculator>>(m, "BaseMonteCalculator")
      .def_readonly("calculator_name", &BaseMonteCalculator::calculator_name)
      .def_readonly("is_multistate_method", &BaseMonteCalculator::is_multistate_method)
      .def("standard_sampling_function_names", &BaseMonteCalculator::standard_sampling_function_names)
      .def("standard_analysis_function_names", &BaseMonteCalculator::standard_analysis_function_names)
      .def("make_default_sampling_fixture_params", &BaseMonteCalculator::make_default_sampling_fixture_params);

  // the extern "C" factories of the plugin convention
  m.def("make_canonical_calculator",
        []() { return std::shared_ptr<BaseMonteCalculator>(make_GpuCanonicalCalculator()); });
  m.def("make_semigrand_canonical_calculator",
        []() { return std::shared_ptr<BaseMonteCalculator>(make_GpuSemiGrandCanonicalCalculator()); });
  m.def("make_nfold_calculator", []() { return std::shared_ptr<BaseMonteCalculator>(make_GpuNfoldCalculator()); });

  py::class_<MonteCalculator, std::shared_ptr<MonteCalculator>>(m, "MonteCalculator")
      .def_property_readonly("calculator_name", &MonteCalculator::calculator_name)
      .def_property_readonly("system", &MonteCalculator::system)
      .def_property_readonly("potential", [](MonteCalculator &c) { return c.base().potential; })
      .def("set_state_and_potential", &MonteCalculator::set_state_and_potential, py::arg("state"),
           py::arg("occ_location"))
      .def(
          "run",
          [](MonteCalculator &c, State &state, OccLocation &occ_location, RunManager &run_manager) {
            c.run(state, occ_location, run_manager);
          },
          py::arg("state"), py::arg("occ_location"), py::arg("run_manager"))
      .def(
          "run_multistate",
          [](MonteCalculator &c, std::vector<State> &states, std::vector<OccLocation> &occ_locations,
             RunManager &run_manager) {
            c.run(0, states, occ_locations, run_manager);
            return states;
          },
          py::arg("states"), py::arg("occ_locations"), py::arg("run_manager"));

  // MonteCalculator(method=<BaseMonteCalculator>): the instance is cloned,
  // then reset with params and system (clexmonte_monte_calculator.cpp:155-173)
  m.def(
      "make_monte_calculator",
      [](std::map<std::string, std::string> const &params, std::shared_ptr<System> system,
         std::shared_ptr<BaseMonteCalculator> method) {
        if (!method) throw std::runtime_error("make_monte_calculator: no calculator given");
        return make_monte_calculator(params, system, method->clone());
      },
      py::arg("params"), py::arg("system"), py::arg("method"));
}
