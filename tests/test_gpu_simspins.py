"""The reference's documented ODMR example through ``simspins`` (GPU).

``docs/source/examples.rst:380-433`` of the reference: the 7-level NV model
(``models.nv_7``) at a 1 mT bias, a microwave chirp from 2.82 GHz
(100 MHz over 1 ms, 100 uT) switched on at 1 us, excitation 0.01, ground
populations 1/3 each, output every 150 ns with 300 fine divisions, and
``simspins``' defaults (CF6:5, rotating frame, incoherent inter-block
couplings: n = 19).  The run is cut to its first microseconds (the chirp
keeps its 1 ms sweep rate); the coherent vectorisation (n = 49) runs too.
The drop-in result against the reference algorithm (oracle, C backend) on
the same prepared inputs: 1e-10 relative to the largest density entry, and
the fluorescence trace to 1e-10 relative to its largest value.
"""

import math

import numpy as np
import pytest

from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd import models, params
from superspinsim_amd.generate_simulator import generate_simulator
from superspinsim_amd.sugar import prepare_simspins

pytestmark = pytest.mark.gpu

TIME_START = 1e-6
SWEEP_END = 1000e-6
TIME_STEP = 150e-9
F_START = 2.82e9
F_WIDTH = 100e6


def example_fields():
    def mag_x(time):
        if time > TIME_START:
            phase = math.tau * (
                F_START * (time - TIME_START)
                + F_WIDTH * ((time - TIME_START) ** 2)
                / (SWEEP_END - TIME_START) / 2)
            return 100e-6 * math.sin(phase)
        return 0

    def zero(time):
        return 0

    def excitation(time):
        return 0.01

    return [mag_x, zero, zero, excitation]


def example_rho0():
    rho = np.zeros((7, 7), dtype=np.float64)
    rho[0, 0] = rho[1, 1] = rho[2, 2] = 1 / 3
    return rho


@pytest.mark.parametrize("incoherent,time_end", [(True, 15e-6),
                                                  (False, 6e-6)])
def test_documented_odmr_example_vs_oracle(gpu, incoherent, time_end):
    spins, spin_int, groups = models.nv_7(np.array([0, 0, 1]) * 1e-3, params)
    sampler, gens, vmap, kw = prepare_simspins(
        example_fields(), spins, spin_int, groups,
        number_of_fine_divisions=300, use_incoherent=incoherent)
    assert kw["use_rotating"] and kw["number_of_exponentials"] == 5
    n = gens.shape[-1]
    print(f"nv_7 example: n={n}, {time_end * 1e6:.0f} us", flush=True)
    t_dev, d_dev = generate_simulator(sampler, gens, vmap, **kw)(
        example_rho0(), 0, time_end, TIME_STEP)
    print("  oracle", flush=True)
    t_ref, d_ref = oracle_gen(sampler, gens, vmap, backend="c", **kw)(
        example_rho0(), 0, time_end, TIME_STEP)
    assert np.array_equal(t_dev, t_ref)
    err = np.max(np.abs(d_dev - d_ref)) / np.max(np.abs(d_ref))
    fl_dev = np.real(d_dev[:, 3, 3] + d_dev[:, 4, 4] + d_dev[:, 5, 5])
    fl_ref = np.real(d_ref[:, 3, 3] + d_ref[:, 4, 4] + d_ref[:, 5, 5])
    fl_err = np.max(np.abs(fl_dev - fl_ref)) / np.max(np.abs(fl_ref))
    print(f"  n={n}: density rel err {err:.3e}, fluorescence {fl_err:.3e} "
          f"(max {np.max(fl_ref):.4e})", flush=True)
    assert err <= 1e-10
    assert fl_err <= 1e-10
    assert np.max(fl_ref) > 1e-4
