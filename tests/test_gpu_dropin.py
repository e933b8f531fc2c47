"""Drop-in ``generate_simulator`` on the GPU vs the reference-algorithm oracle.

Parity bar: max |density difference| <= 1e-10 relative to the largest
entry of the trajectory (FP64 both sides; the reference algorithm's own
rounding is ~1e-13 over these horizons).
"""

import numpy as np
import pytest

from golden.cases import GOLDEN, READERS, case, flatten_density, load_csv
from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd.generate_simulator import generate_simulator
from superspinsim_amd.sugar import prepare_simspins
from test_plan_steps import spiky_qubit

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _run(golden, flags, n_out, backend="c"):
    coeffs, spins, si, gi, rho = case(golden)
    sampler, gens, vmap, kw = prepare_simspins(coeffs, spins, si, gi, **flags)
    print(f"case {golden} flags {flags} n_out {n_out}", flush=True)
    t_dev, d_dev = generate_simulator(sampler, gens, vmap, **kw)(
        rho, 0, n_out * 1e-9, 1e-9)
    t_ref, d_ref = oracle_gen(sampler, gens, vmap, backend=backend, **kw)(
        rho, 0, n_out * 1e-9, 1e-9)
    return t_dev, d_dev, t_ref, d_ref


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_writers_vs_oracle(gpu, name):
    flags = GOLDEN[name]
    t_dev, d_dev, t_ref, d_ref = _run(name, flags, 300)
    assert np.array_equal(t_dev, t_ref)
    scale = np.max(np.abs(d_ref))
    err = np.max(np.abs(d_dev - d_ref)) / scale
    print(f"{name}: rel err {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("name", sorted(GOLDEN))
def test_golden_csv_prefix(gpu, name):
    """The writer's own flags, non-residual qubit-couple included: within
    the 1e-10 parity bar of the golden values over the first 1 us."""
    flags = dict(GOLDEN[name])
    coeffs, spins, si, gi, rho = case(name)
    sampler, gens, vmap, kw = prepare_simspins(coeffs, spins, si, gi, **flags)
    _, dens = generate_simulator(sampler, gens, vmap, **kw)(
        rho, 0, 1000e-9, 1e-9)
    mine = flatten_density(dens)
    ref = load_csv(name)[:mine.size]
    err = np.max(np.abs(mine - ref))
    print(f"{name}: max |golden diff| {err:.3e}")
    assert err < 1e-10


#: non-residual variants beyond the golden writer: rotating frame, more
#: exponentials, damping (which this mode drops), kernel and unitary modes
NONRESIDUAL = {
    "qubit-couple-rot": ("qubit-couple", dict(use_rotating=True)),
    "qubit-couple-e2": ("qubit-couple", dict(number_of_exponentials=2)),
    "qubit-t1-mK": ("qubit-t1-mK", dict(use_rotating=False)),
    "qubit-t1-mK-rot-e5": ("qubit-t1-mK", dict(use_rotating=True,
                                               number_of_exponentials=5)),
    "qutrit-couple-ker": ("qutrit-couple", dict(use_kernel=True)),
    "qutrit-couple-unitary": ("qutrit-couple", dict(
        use_rotating=True, number_of_exponentials=5, use_unitary=True)),
    "qubit-hyperfine-rot": ("qubit-hyperfine", dict(use_rotating=True)),
}


@pytest.mark.parametrize("name", sorted(NONRESIDUAL))
def test_nonresidual_variants_vs_oracle(gpu, name):
    golden, extra = NONRESIDUAL[name]
    flags = dict(GOLDEN[golden], use_residual=False,
                 number_of_fine_divisions=20, **extra)
    t_dev, d_dev, t_ref, d_ref = _run(golden, flags, 100)
    assert np.array_equal(t_dev, t_ref)
    err = np.max(np.abs(d_dev - d_ref)) / np.max(np.abs(d_ref))
    print(f"{name}: rel err {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("name", sorted(READERS))
def test_reader_variants_vs_oracle(gpu, name):
    golden, flags = READERS[name]
    t_dev, d_dev, t_ref, d_ref = _run(golden, flags, 120)
    scale = np.max(np.abs(d_ref))
    err = np.max(np.abs(d_dev - d_ref)) / scale
    print(f"{name}: rel err {err:.3e}")
    assert err <= TOL


def test_spiky_sampler_vs_oracle(gpu):
    """A 0.5 ns, 50 mT spike on the x field inside a 100 ns run whose light
    coefficient (on the qubit's all-zero optical generator) varies more in
    magnitude: the Taylor plan must cover the spike steps (their exponents
    are 100x the others'), or their truncation error shows at ~1e-6."""
    sampler, gens, vmap, kw, rho = spiky_qubit()
    t_dev, d_dev = generate_simulator(sampler, gens, vmap,
                                      **dict(kw, verbose=True))(
        rho, 0, 100e-9, 1e-9)
    t_ref, d_ref = oracle_gen(sampler, gens, vmap, backend="c", **kw)(
        rho, 0, 100e-9, 1e-9)
    assert np.array_equal(t_dev, t_ref)
    err = np.max(np.abs(d_dev - d_ref)) / np.max(np.abs(d_ref))
    print(f"spiky sampler: rel err {err:.3e}")
    assert err <= TOL
    # the spike flips a visible part of the population
    assert np.max(np.abs(d_ref[45, 1, 1] - d_ref[35, 1, 1])) > 1e-3
