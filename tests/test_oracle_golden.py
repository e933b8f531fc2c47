"""The oracle restatement against the reference's own golden histories.

The seven CSV files are the reference's fixtures; with the writer's flags
the oracle must reproduce them to rounding (the reference stores 19
significant digits).  Both oracle backends are checked, and on every
reader variant the two backends must agree with each other.
"""

import numpy as np
import pytest

from golden.cases import GOLDEN, READERS, case, flatten_density, load_csv
from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd.sugar import prepare_simspins


def _history(name, flags, n_out, backend):
    coeffs, spins, si, gi, rho = case(name)
    sampler, gens, vmap, kw = prepare_simspins(coeffs, spins, si, gi,
                                               **flags)
    return oracle_gen(sampler, gens, vmap, backend=backend, **kw)(
        rho, 0, n_out * 1e-9, 1e-9)


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_c_oracle_reproduces_golden(name):
    _, dens = _history(name, GOLDEN[name], 400, "c")
    mine = flatten_density(dens)
    ref = load_csv(name)[:mine.size]
    assert np.max(np.abs(mine - ref)) < 1e-14


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_numpy_oracle_reproduces_golden(name):
    _, dens = _history(name, GOLDEN[name], 40, "numpy")
    mine = flatten_density(dens)
    ref = load_csv(name)[:mine.size]
    assert np.max(np.abs(mine - ref)) < 1e-14


@pytest.mark.parametrize("name", sorted(READERS))
def test_oracle_backends_agree_on_readers(name):
    golden, flags = READERS[name]
    flags = dict(flags, number_of_fine_divisions=10)
    _, a = _history(golden, flags, 25, "c")
    _, b = _history(golden, flags, 25, "numpy")
    assert np.max(np.abs(a - b)) < 1e-13


def test_time_grid_floors_like_reference():
    """N_out = int((t1 - t0) / dt): 1 us at 1 ns gives 999 samples."""
    coeffs, spins, si, gi, rho = case("qubit-couple")
    flags = dict(GOLDEN["qubit-couple"], number_of_fine_divisions=1)
    sampler, gens, vmap, kw = prepare_simspins(coeffs, spins, si, gi,
                                               **flags)
    t, d = oracle_gen(sampler, gens, vmap, backend="c", **kw)(
        rho, 0, 1e-6, 1e-9)
    assert t.size == 999 and d.shape == (999, 2, 2)
    assert t[0] == 1e-9 and t[-1] == 1e-9 * 999
