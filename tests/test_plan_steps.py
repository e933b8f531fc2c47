"""Which steps plan the drop-in's Taylor degree (CPU).

``generate_simulator`` plans the degree from the exact powers of the
exponents of sampled steps.  A step is sampled when its bound
sum |c| ||G||_1 on the exponent norm is among the largest, so a short
spike on a generator of large norm is planned for even where another
coefficient, on a generator of small (here zero) norm, is larger in
magnitude.  Ranking by coefficient magnitude alone misses the spike.
"""

import math

import numpy as np

from golden.cases import _qubit
from superspinsim_amd import quadratures
from superspinsim_amd.generate_simulator import plan_steps, tabulate
from superspinsim_amd.sugar import prepare_simspins

#: the spike: [39.0, 39.5) ns of a 100 ns run, 1 ns outputs, F = 100
SPIKE = (39.0e-9, 39.5e-9)
SPIKE_STEPS = set(range(3900, 3950))
FINE = 100


def spiky_fields():
    def mag_x(t):
        v = 10e-6 * math.cos(math.tau * 28e6 * t)
        if SPIKE[0] <= t < SPIKE[1]:
            v += 0.05
        return v

    def zero(t):
        return 0.0

    def light(t):
        return 0.75 + 0.25 * math.cos(math.tau * t / 100e-9)

    return [mag_x, zero, zero, light]


def spiky_qubit():
    """``(sampler, gens, vmap, kw, rho0)`` of the spiky qubit run."""
    sampler, gens, vmap, kw = prepare_simspins(
        spiky_fields(), _qubit(), [{}], {}, use_rotating=False,
        number_of_exponentials=2, number_of_fine_divisions=FINE)
    rho = np.zeros((2, 2))
    rho[0, 0] = 1.0
    return sampler, gens, vmap, kw, rho


def _table():
    sampler, gens, _, _, _ = spiky_qubit()
    nodes, _, _ = quadratures.scheme(2)
    coef = tabulate(sampler, gens.shape[0], 100, FINE, nodes, 0.0, 1e-9)
    flat = coef.reshape(-1, coef.shape[2], coef.shape[3])
    norms = np.array([np.abs(g).sum(axis=0).max() for g in gens])
    return flat, norms


def test_spike_steps_are_sampled():
    flat, norms = _table()
    # the optical generator of a qubit is zero, its coefficient the largest
    assert norms[-1] == 0.0
    assert np.max(np.abs(flat[..., -1])) > 10 * np.max(np.abs(flat[..., 1]))
    pick = plan_steps(flat, norms)
    assert len(SPIKE_STEPS & set(pick.tolist())) == 48


def test_magnitude_ranking_misses_the_spike():
    """The plausible mistake: unweighted magnitudes (and the evenly spread
    samples) sample no spike step."""
    flat, norms = _table()
    pick = plan_steps(flat, np.ones_like(norms))
    assert not SPIKE_STEPS & set(pick.tolist())


def test_short_runs_sample_every_step():
    flat, norms = _table()
    pick = plan_steps(flat[:30], norms)
    assert pick.tolist() == list(range(30))
