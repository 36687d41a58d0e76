"""The kernels' Metropolis acceptance test, accept_prob (clexmc_device.h),
compiled for the host (tests/cpp/accept_prob_check): for p from 1e-37 to 1
the acceptance probability implied by the function's own answers over all
first and second 32-bit words equals p to within FP32 rounding of p plus
2^-64, and acceptance is monotone in the first word.  (A single 24-bit
uniform compared with p accepts with probability at least 2^-24.)"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "accept_prob_check")


@pytest.mark.skipif(not os.path.exists(CHECK), reason="tests/cpp/accept_prob_check not built (run build())")
def test_acceptance_probability_matches_p_to_2_pow_minus_64():
    out = subprocess.run([CHECK], capture_output=True, text=True, check=True).stdout.split("\n")
    rows = [tuple(map(float, l.split())) for l in out if l and not l.startswith("shape")]
    shape = [l for l in out if l.startswith("shape_errors")]
    assert shape == ["shape_errors 0"]
    assert len(rows) == 149
    for p, q in rows:
        assert p - 1e-7 * p <= q <= p + 2.0 ** -64 + 1e-7 * p, (p, q)
