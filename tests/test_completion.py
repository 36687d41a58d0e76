"""CPU tests of the sampling fixture's completion checks (completion.py) and
of the fixture parameters read from the reference's run params.

Pins: the check schedule (begin + k * period); the cutoffs (below min never
complete, at max always); equilibration on series with a known transient
(an exponential relaxation of amplitude 1 and decay length 50 samples,
whose equilibration point follows in closed form from the two-halves rule),
a stationary series (0) and a drifting one (never); calculated precision =
1.96 x the error of the mean at confidence 0.95, converged when at most the
requested precision; the ZrO run params' fixture (max 100 passes, potential
energy to 0.001, every pass).
"""
import json
import os

import numpy as np
import pytest

from casmcode_clexmonte_amd.analysis import mean_and_error
from casmcode_clexmonte_amd.calculator import SamplingFixtureParams, merge_summaries
from casmcode_clexmonte_amd.completion import CompletionCheckParams, check, equilibration_index

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ZRO = json.load(open(os.path.join(ROOT, "tests", "golden", "casm_clex_fixtures.json")))["Clex_ZrO_Occ"]


def test_check_schedule():
    c = CompletionCheckParams(begin=0.0, period=10.0)
    assert [c.next_check(x) for x in (0, 1, 9, 10, 11)] == [10, 10, 10, 20, 20]
    c = CompletionCheckParams(begin=5.0, period=10.0)
    assert [c.next_check(x) for x in (0, 4, 5, 6, 15)] == [5, 5, 15, 15, 25]
    c = CompletionCheckParams(begin=0.0, period=2.5)
    assert [c.next_check(x) for x in (0, 2, 3, 5)] == [3, 3, 5, 8]


def test_equilibration_index():
    rng = np.random.default_rng(0)
    n = 2000
    t = np.arange(n)
    relax = np.exp(-t / 50.0) + rng.normal(0, 1e-3, n)
    k = equilibration_index(relax, 0.01)
    # closed form of the rule: the first grid point k = 0, 100, 200, ... at which
    # the two halves (h samples each) of exp(-t/50), t >= k, differ by at most 0.01
    expect = None
    for kk in range(0, n // 2 + 1, n // 20):
        h = (n - kk) // 2
        first = np.exp(-kk / 50.0) * (1 - np.exp(-h / 50.0)) / (1 - np.exp(-1 / 50.0)) / h
        second = np.exp(-(kk + h) / 50.0) * (1 - np.exp(-h / 50.0)) / (1 - np.exp(-1 / 50.0)) / h
        if abs(first - second) <= 0.01:
            expect = kk
            break
    assert k == expect == 100, (k, expect)
    assert equilibration_index(rng.normal(0, 1e-3, n), 0.01) == 0
    assert equilibration_index(np.linspace(0, 1, n), 0.01) is None


def test_cutoffs_and_convergence():
    rng = np.random.default_rng(1)
    x = rng.normal(-0.5, 0.01, 10000)
    p = CompletionCheckParams(min_count=100, max_count=None, convergence={"potential_energy": 0.001})
    r = check({"potential_energy": x}, 50, p)
    assert not r.complete and r.all_converged          # converged, but below the minimum count
    r = check({"potential_energy": x}, 100, p)
    assert r.complete and r.all_equilibrated and r.n_equilibrate == 0
    m, err = mean_and_error(x)
    q = r.quantities["potential_energy"]
    assert q["mean"] == m and abs(q["calculated_precision"] - 1.959963984540054 * err) < 1e-15
    short = check({"potential_energy": x[:10]}, 100, p)
    assert not short.complete and not short.all_converged
    p_max = CompletionCheckParams(max_count=100, convergence={"potential_energy": 1e-9})
    assert check({"potential_energy": x}, 100, p_max).complete       # the cutoff ends the run
    assert not check({"potential_energy": x}, 99, p_max).complete


def test_zro_fixture_params():
    fx = ZRO["run_params_complete"]["sampling_fixtures"]["thermo"]
    p = SamplingFixtureParams.from_casm("thermo", fx)
    assert p.sample_period == 1 and p.n_equilibration_passes == 0
    assert "clex.formation_energy" in p.sampler_names and "potential_energy" in p.sampler_names
    c = p.completion
    assert c.max_count == 100 and c.min_count is None and c.period == 10.0 and c.begin == 0.0
    assert c.convergence == {"potential_energy": 0.001}
    assert abs(c.z() - 1.959963984540054) < 1e-12


def test_merge_summaries_layout():
    a = {"conditions": {"temperature": {"shape": [], "value": [300.0]}},
         "statistics": {"x": {"component_names": ["a"], "shape": [1], "a": {"mean": [1.0]}}}}
    b = {"conditions": {"temperature": {"shape": [], "value": [310.0]}},
         "statistics": {"x": {"component_names": ["a"], "shape": [1], "a": {"mean": [2.0]}}}}
    m = merge_summaries([a, b])
    assert m["conditions"]["temperature"] == {"shape": [], "value": [300.0, 310.0]}
    assert m["statistics"]["x"] == {"component_names": ["a"], "shape": [1], "a": {"mean": [1.0, 2.0]}}
