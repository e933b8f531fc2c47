"""Drop-in runs of the other coupling kinds on the GPU vs the oracle.

Spin-spin J (two spins in one block, driven), coherent coupling between
two blocks in L0 and in the optically controlled generator, and Bell-state
ISC of a spin pair and a singlet; each through ``simspins``' plumbing and
``generate_simulator`` against the reference algorithm (C restatement),
1e-10 relative to the largest entry.  The optically controlled coherent
coupling at constant excitation r also follows the closed form
P_a(t) = cos^2(r amp t).
"""

import math

import numpy as np
import pytest

from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd.generate_simulator import generate_simulator
from superspinsim_amd.sugar import prepare_simspins
from test_couplings import COH_AMP, bell_model, coh_model, j_model

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _fields(amp, freq, excitation):
    def mag_x(t):
        return amp * math.cos(math.tau * freq * t)

    def zero(t):
        return 0.0

    def light(t):
        return excitation

    return [mag_x, zero, zero, light]


def _compare(name, description, coefficients, rho, n_out, dt, **flags):
    spins, spin_int, groups = description
    sampler, gens, vmap, kw = prepare_simspins(coefficients, spins, spin_int,
                                               groups, **flags)
    print(f"{name}: n={gens.shape[1]} flags {flags}", flush=True)
    t_dev, d_dev = generate_simulator(sampler, gens, vmap, **kw)(
        rho, 0.0, (n_out + 0.5) * dt, dt)
    t_ref, d_ref = oracle_gen(sampler, gens, vmap, backend="c", **kw)(
        rho, 0.0, (n_out + 0.5) * dt, dt)
    assert np.array_equal(t_dev, t_ref)
    err = np.max(np.abs(d_dev - d_ref)) / np.max(np.abs(d_ref))
    print(f"  rel err {err:.3e}", flush=True)
    assert err <= TOL
    return t_dev, d_dev


@pytest.mark.parametrize("rotating", [False, True])
def test_spin_spin_coupling_vs_oracle(gpu, rotating):
    rho = np.zeros((4, 4))
    rho[0, 0] = 1.0
    _compare("J", j_model(), _fields(2e-5, 28e6, 0.0), rho, 120, 1e-9,
             use_rotating=rotating, number_of_exponentials=2,
             number_of_fine_divisions=20)


@pytest.mark.parametrize("con", [False, True])
def test_coherent_block_coupling_vs_oracle(gpu, con):
    rho = np.diag([1.0, 0.0])
    r = 0.5
    t, d = _compare(f"coh con={con}", coh_model(con), _fields(0, 0, r), rho,
                    150, 10e-9, use_rotating=False,
                    number_of_exponentials=2, number_of_fine_divisions=10)
    rate = COH_AMP * (r if con else 1.0)
    want = np.cos(rate * t) ** 2
    assert np.max(np.abs(d[:, 0, 0].real - want)) < 1e-10


def test_bell_isc_vs_oracle(gpu):
    rho = np.zeros((5, 5))
    rho[4, 4] = 1.0
    _, d = _compare("bell", bell_model(), _fields(1e-4, 28e6, 0.0), rho,
                    200, 2e-9, use_rotating=True, number_of_exponentials=2,
                    number_of_fine_divisions=20)
    assert d[-1, 4, 4].real < 0.5
