"""Completion checks of a sampling fixture: cutoffs, equilibration and
requested precision, evaluated on the per-pass samples of each chain.

Follows the reference's completion-check parameters
(``run_params_complete.json`` ``sampling_fixtures.<label>.completion_check``:
``cutoff.count.{min,max}``, ``begin``, ``period``, ``confidence``,
``convergence[{quantity, precision, component_name?}]``; defaults configured
in ``src/casm/clexmonte/monte_calculator/CanonicalCalculator.cc:261-270`` and
``SemiGrandCanonicalCalculator.cc:338-340``) and the ``completion_check_results``
of ``summary.json`` (``python/tests/conftest.py:205-226``).  The checks
themselves live in libcasm-monte (``RunManager``, ``BasicStatistics``,
``default_equilibration_check``), which is not vendored, so their formulas
here are this build's own and stated in full:

* a check happens at pass counts ``begin + k * period`` (k = 0, 1, ...);
* below ``cutoff.count.min`` a fixture is never complete; at or above
  ``cutoff.count.max`` it always is;
* otherwise it is complete when every requested quantity is equilibrated and
  converged.  A series of n samples is equilibrated from the first k on the
  grid {0, n/20, 2n/20, ..., n/2} whose remaining samples split into two
  halves with means within the requested precision of each other; the
  statistics use the samples after the largest such k over the requested
  quantities;
* ``calculated_precision`` = z * error of the mean, with z the two-sided
  normal quantile of ``confidence`` (1.96 at 0.95) and the error
  autocorrelation-corrected (Sokal window, ``analysis.mean_and_error``);
  a quantity is converged when its calculated precision is at most the
  requested one.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from statistics import NormalDist
from typing import Dict, Optional

import numpy as np

from .analysis import mean_and_error


@dataclass
class CompletionCheckParams:
    min_count: Optional[int] = None
    max_count: Optional[int] = None
    begin: float = 0.0
    period: float = 10.0
    confidence: float = 0.95
    # requested precision per quantity: "potential_energy", "clex.formation_energy",
    # "param_composition.a", "mol_composition.O", ...
    convergence: Dict[str, float] = field(default_factory=dict)

    @classmethod
    def from_casm(cls, d: dict) -> "CompletionCheckParams":
        cut = (d.get("cutoff") or {}).get("count") or {}
        conv = {}
        for c in d.get("convergence", []):
            q = c["quantity"]
            comp = c.get("component_name", c.get("component_index"))
            name = q if comp is None else f"{q}.{comp}"
            conv[name] = float(c["precision"])
        return cls(min_count=cut.get("min"), max_count=cut.get("max"), begin=float(d.get("begin", 0.0)),
                   period=float(d.get("period", 10.0)), confidence=float(d.get("confidence", 0.95)),
                   convergence=conv)

    def z(self) -> float:
        return NormalDist().inv_cdf(0.5 + 0.5 * self.confidence)

    def next_check(self, count: int) -> int:
        """The first pass count above `count` at which a check happens:
        ceil(begin + k * period) for the least integer k >= 0."""
        p = max(float(self.period), 1.0)
        k = 0 if count < self.begin else int(np.floor((count - self.begin) / p)) + 1
        c = int(np.ceil(self.begin + k * p))
        while c <= count:
            k += 1
            c = int(np.ceil(self.begin + k * p))
        return c


def equilibration_index(x, precision: float) -> Optional[int]:
    """First k on the grid {0, n/20, ..., n/2} after which the two halves of
    x[k:] have means within precision; None if there is none."""
    x = np.asarray(x, dtype=float)
    n = x.size
    if n < 4:
        return None
    for s in range(11):
        k = (s * n) // 20
        rest = x[k:]
        h = rest.size // 2
        if h < 2:
            break
        if abs(rest[:h].mean() - rest[h:].mean()) <= precision:
            return k
    return None


@dataclass
class CheckResult:
    complete: bool
    all_equilibrated: bool
    all_converged: bool
    n_equilibrate: int                  # samples discarded before the statistics
    quantities: Dict[str, dict]         # name -> {mean, calculated_precision, is_converged}


def check(series: Dict[str, np.ndarray], count: int, p: CompletionCheckParams) -> CheckResult:
    """Completion of one chain after `count` passes; `series` maps requested
    quantity names to their per-sample values."""
    n = min((len(v) for v in series.values()), default=0)
    eq = {}
    for name, prec in p.convergence.items():
        eq[name] = equilibration_index(series[name][:n], prec) if name in series else None
    all_eq = bool(p.convergence) and all(v is not None for v in eq.values())
    n_eq = max([v for v in eq.values() if v is not None], default=0)
    z = p.z()
    quantities = {}
    all_conv = bool(p.convergence)
    for name, prec in p.convergence.items():
        x = np.asarray(series.get(name, []), dtype=float)[n_eq:n]
        if x.size >= 2:
            m, err = mean_and_error(x)
            cp = z * err
        else:
            m, cp = float("nan"), float("inf")
        conv = bool(eq[name] is not None and cp <= prec)
        all_conv = all_conv and conv
        quantities[name] = {"mean": float(m), "calculated_precision": float(cp), "is_converged": conv}
    if p.min_count is not None and count < p.min_count:
        complete = False
    elif p.max_count is not None and count >= p.max_count:
        complete = True
    else:
        complete = bool(p.convergence) and all_eq and all_conv
    return CheckResult(complete, all_eq, all_conv, int(n_eq), quantities)
