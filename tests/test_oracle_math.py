"""Pins of the oracle that the golden CSVs leave open.

Each check is fixed by the mathematics, so that a dropped term, a wrong
sign, a transposed operand or a swapped CF ordering fails one of them:
constant generators (exact exponential), T1 relaxation (closed form),
trace preservation, the order of CF4:2 under step halving, agreement with
an adaptive ODE solver, and rotating vs non-rotating frames as h -> 0.
"""

import math

import numpy as np
import pytest

from oracle.brute import expm_history, ode_history
from oracle.reference_simulate import flatten
from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd.generate_generators import (
    generate_atoms, make_rotating)
from superspinsim_amd.sugar import prepare_simspins


def _qutrit(extra=None):
    atom = {"S": 1, "g": 2.0, "D": math.tau * 2.87e7,
            "B0": np.array([0, 0, 2e-3]), "TS1": 2e-7, "TS2": 1e-7}
    atom.update(extra or {})
    return [[atom]], [{}], {}


def _gens(desc):
    ops, vmap = generate_atoms(*desc)
    g = ops["generators"]
    return np.array([g[k] for k in ("L0", "Gx", "Gy", "Gz", "Gr")]), vmap


@pytest.mark.parametrize("j", [1, 2, 5])
def test_constant_generator_is_exact_exponential(j):
    gens, vmap = _gens(_qutrit())
    sampler = lambda t, row: row.__setitem__(0, 1.0)  # noqa: E731
    sim = oracle_gen(sampler, gens, vmap, backend="c",
                     number_of_exponentials=j, number_of_fine_divisions=3)
    rho = np.diag([0.2, 0.5, 0.3]).astype(complex)
    rho[0, 1] = rho[1, 0] = 0.1
    t, dens = sim(rho, 0.0, 40e-9, 1e-9)
    x0 = flatten(rho, vmap, True, vmap.shape[0])
    ref = expm_history(gens[0], x0, t)
    mine = np.array([flatten(d, vmap, True, vmap.shape[0]) for d in dens])
    assert np.max(np.abs(mine - ref)) < 1e-12


def test_t1_relaxation_closed_form():
    """Qubit T1 at finite temperature: p1(t) = p_eq + (p1(0)-p_eq) e^-t/T1.

    The thermal jumps of the reference model (rates b_f/((b_i+b_f) T1))
    relax populations at exactly 1/T1 towards the Boltzmann ratio.
    """
    T1, temp, bz = 1e-7, 0.05, 0.5
    desc = ([[{"S": 1 / 2, "g": 2.0, "B0": np.array([0, 0, bz]),
               "TS1": T1, "T": temp}]], [{}], {})
    gens, vmap = _gens(desc)
    sampler = lambda t, row: row.__setitem__(0, 1.0)  # noqa: E731
    sim = oracle_gen(sampler, gens, vmap, backend="c",
                     number_of_exponentials=2, number_of_fine_divisions=5)
    rho = np.diag([0.0, 1.0]).astype(complex)
    t, dens = sim(rho, 0.0, 300e-9, 5e-9)
    from superspinsim_amd import params
    e = math.tau * 2.0 * params.fundamental.bohr_gyro * bz * np.array(
        [0.5, -0.5])
    b = np.exp(-e / (math.tau * params.fundamental.boltzmann_gyro * temp))
    p_eq = b[1] / b.sum()
    ref = p_eq + (1 - p_eq) * np.exp(-t / T1)
    assert np.max(np.abs(dens[:, 1, 1].real - ref)) < 1e-12


def test_trace_is_preserved():
    coeffs = [lambda t: 3e-4 * math.cos(math.tau * 2.87e7 * t),
              lambda t: 0.0, lambda t: 0.0, lambda t: 0.0]
    desc = _qutrit()
    sampler, gens, vmap, kw = prepare_simspins(
        coeffs, *desc, number_of_exponentials=2, number_of_fine_divisions=4,
        use_rotating=False)
    _, dens = oracle_gen(sampler, gens, vmap, backend="c", **kw)(
        np.diag([0, 1.0, 0]).astype(complex), 0, 200e-9, 2e-9)
    tr = np.trace(dens, axis1=1, axis2=2)
    assert np.max(np.abs(tr - 1)) < 1e-13


def _driven_qubit(fine, j=2, rotating=False):
    coeffs = [lambda t: 2e-3 * math.cos(math.tau * 5.6e7 * t),
              lambda t: 0.0, lambda t: 0.0, lambda t: 0.0]
    desc = ([[{"S": 1 / 2, "g": 2.0, "B0": np.array([0, 0, 2e-3]),
               "TS1": 1e-6, "T": 1.0}]], [{}], {})
    sampler, gens, vmap, kw = prepare_simspins(
        coeffs, *desc, number_of_exponentials=j,
        number_of_fine_divisions=fine, use_rotating=rotating)
    return oracle_gen(sampler, gens, vmap, backend="c", **kw)(
        np.diag([1.0, 0]).astype(complex), 0, 100e-9, 5e-9)[1]


def test_cf4_is_fourth_order():
    ref = _driven_qubit(256)
    errs = [np.max(np.abs(_driven_qubit(f) - ref)) for f in (4, 8, 16)]
    ratios = [errs[k] / errs[k + 1] for k in range(2)]
    assert all(12 < r < 20 for r in ratios), ratios


def test_agrees_with_adaptive_ode_solver():
    """Fine CF4:2 vs DOP853 on the same Lindblad ODE: within 1e-8."""
    amp, f = 2e-3, 5.6e7
    desc = ([[{"S": 1 / 2, "g": 2.0, "B0": np.array([0, 0, 2e-3]),
               "TS1": 1e-6, "T": 1.0}]], [{}], {})
    gens, vmap = _gens(desc)
    x0 = flatten(np.diag([1.0, 0]).astype(complex), vmap, True, 4)

    def coef(t):
        return np.array([1.0, amp * math.cos(math.tau * f * t), 0, 0, 0])

    ref = ode_history(gens, coef, x0, 0.0, 5e-9 * np.arange(1, 21))
    dens = _driven_qubit(64)
    mine = np.array([flatten(d, vmap, True, 4) for d in dens])
    assert np.max(np.abs(mine - ref)) < 1e-8


def test_rotating_frame_converges_to_lab_frame():
    lab = _driven_qubit(256)
    rot = _driven_qubit(256, rotating=True)
    assert np.max(np.abs(lab - rot)) < 1e-9


def test_vectorised_waveform_sampler_matches_waveform():
    """oracle.waves evaluates the same coefficient formula as Waveform."""
    from oracle.waves import vectorised_sampler
    from superspinsim_amd.sweep import Waveform
    w = Waveform(4, base=[1.0, 0.0, 0.5, 0.0])
    w.tone(1, 2e-4, 2.87e9, t_start=1e-9, t_end=3e-9, phase=0.3)
    w.add(1, 3e-9, 5e-9, 1e-4)
    w.add(3, 0.0, 2e-9, -3e-4, omega=1e8)
    rng = np.random.default_rng(0)
    ts = np.concatenate([rng.uniform(-1e-9, 6e-9, 200),
                         [0.0, 1e-9, 2e-9, 3e-9, 5e-9]])
    got = vectorised_sampler(w)(ts)
    want = np.array([w(t) for t in ts])
    assert np.array_equal(got, want)
