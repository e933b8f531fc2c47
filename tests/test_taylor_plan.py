"""Taylor degree planning from sampled exponents (CPU).

``taylor_plan_measured`` bounds the truncation error of (T_m(A/s))^s x by
s ||B^(m+1)||_1 / (m+1)! / (1 - ||B||_1 / (m+2)), B = A/s, times a safety
factor, over every sampled exponent.  The brute force below restates that
bound with ``numpy.linalg.matrix_power`` and scans (m, s) directly; the
samples include one whose norm is >= 3, for which the bound is infinite at
the smallest degrees, so a plan that stops raising the other samples' powers
at such a degree would bound them with a lower power and accept too low a
degree.
"""

import math

import numpy as np
import pytest

from superspinsim_amd.propagator import (taylor_degrees_measured,
                                         taylor_plan_measured)


def _brute_plan(exps, tol, safety=8.0, max_taylor=30, max_substeps=1024):
    best = None
    s = 1
    while s <= max_substeps:
        for m in range(1, max_taylor + 1):
            worst = 0.0
            for a in exps:
                b = a / s
                nb = np.abs(b).sum(axis=0).max()
                if nb >= m + 2:
                    worst = math.inf
                    break
                pn = np.abs(np.linalg.matrix_power(b, m + 1)).sum(axis=0).max()
                worst = max(worst,
                            pn / math.factorial(m + 1) / (1 - nb / (m + 2)))
            if s * safety * worst <= tol:
                if best is None or m * s < best[0] * best[1]:
                    best = (m, s)
                break
        s *= 2
        if best is not None and best[0] * best[1] <= s:
            break
    return best


def _samples():
    rng = np.random.default_rng(5)
    big = rng.normal(size=(6, 6))
    big *= 3.5 / np.abs(big).sum(axis=0).max()        # ||B||_1 = 3.5
    grow = 1.5 * np.eye(6) + 0.1 * rng.normal(size=(6, 6))
    small = 0.2 * rng.normal(size=(6, 6))
    return [big, grow, small]


@pytest.mark.parametrize("tol", [1e-6, 1e-10, 1e-14, 1e-17])
def test_plan_matches_brute_force(tol):
    exps = _samples()
    assert taylor_plan_measured(exps, tol=tol, max_taylor=30) == \
        _brute_plan(exps, tol)


@pytest.mark.parametrize("tol", [1e-6, 1e-10, 1e-14])
def test_one_substep_plan_matches_brute_force(tol):
    exps = _samples()
    m, s = taylor_plan_measured(exps, tol=tol, max_taylor=30,
                                max_substeps=1)
    assert (m, s) == _brute_plan(exps, tol, max_substeps=1)


def test_degrees_for_several_tolerances_in_one_pass():
    exps = _samples()
    tols = (1e-8, 1e-12, 1e-16)
    got = taylor_degrees_measured(exps, tols, max_taylor=(30, 30, 30))
    want = [taylor_plan_measured(exps, tol=t, max_taylor=30,
                                 max_substeps=1)[0] for t in tols]
    assert got == want
    assert got == sorted(got) and got[0] < got[-1]
    with pytest.raises(ValueError):
        taylor_degrees_measured(exps, (1e-16,), max_taylor=(5,))
