"""Results analysis: the reference's variance / covariance analysis functions
and the autocorrelation / effective-sample-size estimators used to state
honest statistical errors.

* ``heat_capacity``  = var(potential_energy per unit cell) * n_unitcells / (kB T^2)
  (src/casm/clexmonte/monte_calculator/analysis_functions.cc:16-51,
   src/casm/clexmonte/run/io/convariance_functions.cc:29-63)
* ``mol_susc``, ``param_susc`` = cov(x_i, x_j) * n_unitcells / (kB T)
  (analysis_functions.cc:53-118, convariance_functions.cc:82-143)
* ``mol_thermochem_susc``, ``param_thermochem_susc`` = cov(E, x_j) * n / (kB T)
  (analysis_functions.cc:120-173)

Statistics use the last ``n_stats`` samples (the reference's
N_samples_for_statistics, convariance_functions.cc:48) and return NaN when
there are none (:53-54).  The variance normalisation (1/N) is this build's
choice: libcasm-monte's ``monte::variance`` is not vendored.
"""
from __future__ import annotations

from typing import Dict, Optional

import numpy as np

from .system import KB


def variance(x) -> float:
    x = np.asarray(x, dtype=float)
    if x.size == 0:
        return float("nan")
    return float(np.mean((x - x.mean()) ** 2))


def covariance(x, y) -> float:
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    if x.size == 0:
        return float("nan")
    return float(np.mean((x - x.mean()) * (y - y.mean())))


def heat_capacity(pe_per_unitcell, temperature: float, n_unitcells: int, kb: float = KB) -> float:
    return variance(pe_per_unitcell) * n_unitcells / (kb * temperature * temperature)


def susceptibility(x, y, temperature: float, n_unitcells: int, kb: float = KB) -> float:
    return covariance(x, y) * n_unitcells / (kb * temperature)


def tau_int(x, c: float = 5.0) -> float:
    """Integrated autocorrelation time (in samples) with Sokal's automatic
    window: the smallest M with M >= c * tau(M)."""
    x = np.asarray(x, dtype=float)
    n = len(x)
    if n < 4 or np.var(x) == 0:
        return 0.5
    y = x - x.mean()
    f = np.fft.rfft(y, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    acf = acf / acf[0]
    csum = np.cumsum(acf[1:])
    for M in range(1, n):
        tau = 0.5 + csum[M - 1]
        if M >= c * tau:
            return max(float(tau), 0.5)
    return max(float(0.5 + csum[-1]), 0.5)


def ess(x, c: float = 5.0) -> float:
    """Effective sample size N / (2 tau_int)."""
    return len(x) / (2.0 * tau_int(x, c))


def mean_and_error(x, c: float = 5.0):
    """Mean and its standard error corrected for autocorrelation."""
    x = np.asarray(x, dtype=float)
    n = len(x)
    if n == 0:
        return float("nan"), float("nan")
    t = tau_int(x, c)
    return float(x.mean()), float(np.sqrt(np.var(x) * 2.0 * t / n))


def analyze_chain(E_sc, n1, temperature: float, n_unitcells: int, mu: Optional[float] = None,
                  kb: float = KB) -> Dict[str, float]:
    """Per-chain observables for a binary system from the per-pass samples
    (formation energy per supercell, count of occupant 1).  The potential
    energy is E - mu * N_1 for semi-grand canonical runs (the reference's
    SemiGrandCanonicalPotential, SemiGrandCanonicalCalculator.cc:171-182)."""
    E = np.asarray(E_sc, dtype=float) / n_unitcells
    x = np.asarray(n1, dtype=float) / n_unitcells
    pe = E - (mu * x if mu is not None else 0.0)
    out = {}
    out["formation_energy"], out["formation_energy_err"] = mean_and_error(E)
    out["potential_energy"], out["potential_energy_err"] = mean_and_error(pe)
    out["x"], out["x_err"] = mean_and_error(x)
    out["heat_capacity"] = heat_capacity(pe, temperature, n_unitcells, kb)
    out["param_susc"] = susceptibility(x, x, temperature, n_unitcells, kb)
    out["param_thermochem_susc"] = susceptibility(pe, x, temperature, n_unitcells, kb)
    out["tau_int_E"] = tau_int(E)
    return out
