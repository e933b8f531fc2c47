"""Uneven protocol batches in one persistent work-queue launch (GPU).

Short pulsed Ramsey / Hahn-echo protocols on the 7-level model (n = 49):
* the queue kernel's windowed integral equals the regular propagator's
  per-step fluorescence integrated on the host over each readout window
  (rtol 1e-12: same states, different summation order);
* two instances against the reference algorithm in C (oracle/cpu_ref.c),
  1e-10 relative;
* independence of the grid size and wavefronts per workgroup (bitwise);
* HIP-graph replays of (counter reset, launch) equal the eager launch,
  with other device work between replays;
* the autotuner's choice reproduces the result (1e-10).
"""

import numpy as np
import pytest

from superspinsim_amd import configs

pytestmark = pytest.mark.gpu

H = 20e-12


def _workload(count=24):
    return configs.pulsed7(count=count, seed=5, t_p_range=(2e-9, 20e-9),
                           tau_range=(10e-9, 100e-9), readout=50e-9)


@pytest.fixture(scope="module")
def wl(gpu):
    return _workload()


@pytest.fixture(scope="module")
def queued(wl):
    b = wl.batch().prepare(wl.waves, wl.rho0, H, wl.steps, wl.read_from,
                           wl.obs)
    out = b.run().cpu().numpy()[:, 0].copy()
    return b, out


def test_matches_per_step_readout(wl, queued):
    from superspinsim_amd.protocols import trapezoid_window
    from superspinsim_amd.sweep import BatchSimulator
    b, out = queued
    T = int(wl.steps.max())
    sim = BatchSimulator(wl.generators, 2, 1)
    series = sim.run(wl.waves, wl.rho0, 0.0, T * H * (1 + 1e-12), H,
                     obs=wl.obs, read="obs").cpu().numpy()[:, :, 0]
    assert series.shape[1] == T
    s0 = np.full((len(wl.waves), 1), float(wl.obs[0] @ wl.rho0))
    full = np.concatenate([s0, series], axis=1)
    want = trapezoid_window(full, wl.read_from, wl.steps, H)
    assert np.all(want > 0)
    assert np.max(np.abs(out - want) / np.abs(want)) < 1e-12
    assert (b.taylor, b.substeps) == (sim.last_plan["taylor"],
                                      sim.last_plan["substeps"])


def test_matches_reference_algorithm(wl, queued):
    """Two protocols vs the reference algorithm (C restatement)."""
    from oracle.reference_simulate import generate_simulator as oracle_gen
    from superspinsim_amd.protocols import trapezoid_window
    _, out = queued
    base = configs.odmr7(1)
    for i in (int(np.argmin(wl.steps)), 3):
        w = wl.waves[i]
        sim = oracle_gen(w.sampler(), wl.generators, base.vmap,
                         backend="c", number_of_exponentials=2,
                         number_of_fine_divisions=1)
        T = int(wl.steps[i])
        rho = np.zeros((7, 7))
        rho[1, 1] = 1.0
        _, dens = sim(rho, 0.0, T * H * (1 + 1e-12), H)
        pops = np.real(dens[:, 3, 3] + dens[:, 4, 4] + dens[:, 5, 5])
        full = np.concatenate([[0.0], pops])[None]
        want = trapezoid_window(full, wl.read_from[i:i + 1],
                                wl.steps[i:i + 1], H)[0]
        assert abs(out[i] - want) <= 1e-10 * abs(want), (i, out[i], want)


def test_grid_shape_does_not_change_results(wl, queued):
    _, out = queued
    for n_waves, wpb in ((1, 1), (5, 2), (7, 8)):
        b = wl.batch().prepare(wl.waves, wl.rho0, H, wl.steps, wl.read_from,
                               wl.obs, waves_per_block=wpb, n_waves=n_waves)
        got = b.run().cpu().numpy()[:, 0]
        assert np.array_equal(got, out), (n_waves, wpb)


def test_graph_replays_equal_eager(wl, queued):
    import torch
    b, out = queued
    b.capture()
    scratch = torch.ones(1 << 20, dtype=torch.float64, device="cuda")
    for k in range(3):
        scratch.mul_(1.5)               # unrelated eager work
        torch.cuda.synchronize()
        b.out.fill_(np.nan)
        b.graph.replay()
        torch.cuda.synchronize()
        got = b.out.cpu().numpy()[:, 0]
        assert np.array_equal(got, out), k
        assert int(b.counter.item()) == len(wl.waves) + b.n_waves


def test_autotuner_choice_reproduces(wl, queued, tmp_path, monkeypatch):
    from superspinsim_amd import autotune
    monkeypatch.setenv("XDG_CACHE_HOME", str(tmp_path))
    _, out = queued
    b = wl.batch().prepare(wl.waves, wl.rho0, H, wl.steps, wl.read_from,
                           wl.obs)
    choice = autotune.tune(b, trial_steps=1000, repeats=1)
    assert choice["waves_per_block"] in autotune.WAVES_PER_BLOCK
    assert len(choice["trials"]) >= len(autotune.WAVES_PER_BLOCK)
    got = b.run().cpu().numpy()[:, 0]
    assert np.max(np.abs(got - out) / np.abs(out)) < 1e-10
    again = autotune.tune(b, trial_steps=1000, repeats=1)
    assert again["cached"]


def test_launch_on_explicit_stream(wl, queued):
    """The counter reset runs on the launch's own stream: two launches in a
    row on a side stream both drain the whole queue."""
    import torch
    b, out = queued
    side = torch.cuda.Stream()
    for k in range(2):
        b.out.fill_(np.nan)
        side.wait_stream(torch.cuda.current_stream())
        b.launch(stream=side.cuda_stream)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        assert np.array_equal(b.out.cpu().numpy()[:, 0], out), k
