"""The C++ calculator plugin layer from Python, through the pybind11 module
``_clexmc_plugin`` (built next to ``libclexmc.so``): the reference's
external-calculator hook, where ``MonteCalculator(method=<BaseMonteCalculator>)``
clones a calculator instance made by an ``extern "C"`` factory and runs it
(python/libcasm/clexmonte/_MonteCalculator.py:31-90,
python/src/clexmonte_monte_calculator.cpp:34-46,140-173,216-269).

Here the instance comes from ``make_canonical_calculator()`` /
``make_semigrand_canonical_calculator()`` / ``make_nfold_calculator()`` (the
library's ``make_GpuCanonicalCalculator`` / ``make_GpuSemiGrandCanonicalCalculator``
/ ``make_GpuNfoldCalculator``),
the system from the flat file written by ``system.save_flat``, and runs go
through the C++ ``GpuCanonicalCalculator::run`` / multi-state ``run``.

Container attributes (``State.conditions.scalar_values``,
``Configuration.occupation``, ...) are converted on access: assign whole
values (``state.conditions.scalar_values = {"temperature": 600.0}``), do not
modify the returned copies in place.
"""
from __future__ import annotations

import os
from typing import Dict, Optional, Sequence, Union

import numpy as np

try:
    from . import _clexmc_plugin as _p
except ImportError as e:  # the product path has no fallback
    raise ImportError("casmcode_clexmonte_amd._clexmc_plugin is not built (run __graft_entry__.build())") from e

System = _p.System
State = _p.State
Configuration = _p.Configuration
ValueMap = _p.ValueMap
OccLocation = _p.OccLocation
SamplingFixtureParams = _p.SamplingFixtureParams
RunManager = _p.RunManager
BaseMonteCalculator = _p.BaseMonteCalculator
make_canonical_calculator = _p.make_canonical_calculator
make_semigrand_canonical_calculator = _p.make_semigrand_canonical_calculator
make_nfold_calculator = _p.make_nfold_calculator

_FACTORIES = {"canonical": make_canonical_calculator, "semigrand_canonical": make_semigrand_canonical_calculator,
              "nfold": make_nfold_calculator}


def MonteCalculator(method: Union[str, "BaseMonteCalculator"], system: "System",
                    params: Optional[Dict[str, str]] = None):
    """A C++ MonteCalculator: `method` is a BaseMonteCalculator instance
    (cloned, as the reference does) or "canonical" / "semigrand_canonical" /
    "nfold"."""
    if isinstance(method, str):
        if method not in _FACTORIES:
            raise ValueError(f"Error in MonteCalculator: unknown method '{method}'")
        method = _FACTORIES[method]()
    return _p.make_monte_calculator({k: str(v) for k, v in (params or {}).items()}, system, method)


def make_state(dims: Sequence[int], occupation, conditions: Dict[str, object]) -> "State":
    """State of a diagonal supercell with scalar ("temperature") and vector
    ("param_chem_pot", "mol_composition", ...) conditions."""
    s = State()
    c = s.configuration
    c.L = [int(x) for x in dims]
    c.occupation = [int(v) for v in np.asarray(occupation).ravel()]
    scal, vec = {}, {}
    for k, v in conditions.items():
        if np.ndim(v) == 0:
            scal[k] = float(v)
        else:
            vec[k] = [float(x) for x in np.asarray(v).ravel()]
    s.conditions.scalar_values = scal
    s.conditions.vector_values = vec
    return s


def system_from_flat(flat, path: str) -> "System":
    """Write `flat` (a FlatCE) with system.save_flat and read it as a C++ System."""
    from .system import save_flat

    save_flat(flat, path)
    if not os.path.exists(path):
        raise RuntimeError(f"could not write {path}")
    return System.from_flat_file(path)
