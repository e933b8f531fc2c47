"""Batched waveform sweeps on the GPU vs the reference-algorithm oracle.

Each instance's waveform is evaluated on the device; the oracle runs the
reference algorithm (C restatement) with the same waveform as a Python
sampler.  Bar: 1e-10 relative to the trace's max |value| (relative error
of a population or fluorescence trace is taken per trace, against its
largest magnitude).
"""

import numpy as np
import pytest

from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd import configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _oracle_trace(wl, wave, n_out):
    sim = oracle_gen(wave.sampler(), wl.generators, wl.vmap, backend="c",
                     number_of_exponentials=wl.number_of_exponentials,
                     number_of_fine_divisions=wl.fine)
    _, dens = sim(wl.rho0_matrix, 0.0, n_out * wl.time_step * (1 + 1e-12),
                  wl.time_step)
    flat = np.stack([np.array([
        (d[y, x].real if c == 0 else d[y, x].imag)
        for (y, x, c) in wl.vmap]) for d in dens])
    return flat @ wl.obs[0]


@pytest.mark.parametrize("factory,count,n_out", [
    (configs.odmr3, 3, 40),
    (lambda c: configs.odmr7(c), 3, 6),
    (lambda c: configs.odmr7(c, use_incoherent=True), 3, 6),
])
def test_sweep_fluorescence_vs_oracle(gpu, factory, count, n_out):
    wl = factory(count)
    # sweep points spread over the whole band
    wl.frequencies = np.linspace(2.77e9, 2.97e9, count)
    waves = wl.waveforms()
    sim = wl.simulator()
    print(f"{wl.name}: {count} instances x {n_out} outputs", flush=True)
    out = sim.run(waves, wl.rho0, 0.0, n_out * wl.time_step * (1 + 1e-12),
                  wl.time_step, obs=wl.obs, read="obs")
    dev = out[:, :, 0].cpu().numpy()
    for k, w in enumerate(waves):
        ref = _oracle_trace(wl, w, n_out)
        err = np.max(np.abs(dev[k] - ref)) / np.max(np.abs(ref))
        print(f"  f={wl.frequencies[k]:.4e}: rel err {err:.3e}")
        assert err <= TOL


def test_sweep_full_state_matches_obs(gpu):
    wl = configs.odmr7(2)
    sim = wl.simulator()
    waves = wl.waveforms()
    t_end = 5 * wl.time_step * (1 + 1e-12)
    full = sim.run(waves, wl.rho0, 0.0, t_end, wl.time_step, read="full")
    obs = sim.run(waves, wl.rho0, 0.0, t_end, wl.time_step, obs=wl.obs,
                  read="obs")
    f = (full.cpu().numpy() @ wl.obs[0])
    assert np.max(np.abs(f - obs[:, :, 0].cpu().numpy())) < 1e-15


def test_sweep_integral_is_trapezoid(gpu):
    """Integral readout vs trapezoid of the per-step trace (fine=1)."""
    wl = configs.odmr3(2)
    sim = configs.BatchSimulator(wl.generators, 2, 1)
    waves = wl.waveforms()
    h = wl.time_step / wl.fine
    n = 200
    trace = sim.run(waves, wl.rho0, 0.0, n * h * (1 + 1e-12), h,
                    obs=wl.obs, read="obs")[:, :, 0].cpu().numpy()
    integ = sim.run(waves, wl.rho0, 0.0, n * h * (1 + 1e-12), h,
                    obs=wl.obs, read="integral")[:, 0].cpu().numpy()
    f0 = wl.obs[0] @ wl.rho0
    full = np.concatenate([np.full((2, 1), f0), trace], axis=1)
    ref = np.sum(0.5 * h * (full[:, 1:] + full[:, :-1]), axis=1)
    assert np.max(np.abs(integ - ref) / np.abs(ref)) < 1e-13


@pytest.mark.parametrize("chunks", [None, 1, 5, 40])
def test_chunked_sweep_vs_sequential_and_oracle(gpu, chunks):
    """Time-chunked runs (chunk propagators chained by
    ssmx_chunk_combine) against the sequential run and the oracle."""
    wl = configs.odmr3(3)
    wl.frequencies = np.linspace(2.77e9, 2.97e9, 3)
    waves = wl.waveforms()
    sim = wl.simulator()
    n_out = 40
    t1 = n_out * wl.time_step * (1 + 1e-12)
    print(f"{wl.name}: 3 x {n_out} outputs, chunks={chunks}", flush=True)
    seq = sim.run(waves, wl.rho0, 0.0, t1, wl.time_step, obs=wl.obs,
                  read="obs")[:, :, 0].cpu().numpy()
    got = sim.run_chunked(waves, wl.rho0, 0.0, t1, wl.time_step, obs=wl.obs,
                          read="obs", chunks=chunks)[:, :, 0].cpu().numpy()
    print(f"  plan {sim.last_plan['chunks']} x {sim.last_plan['chunk_out']}"
          f", pack {sim.last_plan['chunk_pack']}", flush=True)
    assert got.shape == seq.shape == (3, n_out)
    assert np.max(np.abs(got - seq)) / np.max(np.abs(seq)) <= 1e-11
    for k, w in enumerate(waves):
        ref = _oracle_trace(wl, w, n_out)
        err = np.max(np.abs(got[k] - ref)) / np.max(np.abs(ref))
        print(f"  f={wl.frequencies[k]:.4e}: rel err {err:.3e}")
        assert err <= TOL


def test_chunked_full_states_match_sequential(gpu):
    wl = configs.odmr3(2)
    sim = wl.simulator()
    waves = wl.waveforms()
    t1 = 23 * wl.time_step * (1 + 1e-12)
    seq = sim.run(waves, wl.rho0, 0.0, t1, wl.time_step,
                  read="full").cpu().numpy()
    got = sim.run_chunked(waves, wl.rho0, 0.0, t1, wl.time_step,
                          read="full", chunks=4).cpu().numpy()
    assert got.shape == seq.shape
    assert np.max(np.abs(got - seq)) / np.max(np.abs(seq)) <= 1e-11
