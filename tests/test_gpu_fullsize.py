"""Full-length runs on the GPU against the oracle and a size-independent
property.

* Oracle at full length: single sweep points of the 7-level ODMR sweep
  (999 outputs x 100 steps = 99,900 steps, n = 49) and of the 3-level sweep
  (999 x 50 = 49,950 steps, n = 9) against the reference algorithm in C
  (``oracle/cpu_ref.c``: 70 residual squarings per exponential, CF product,
  serial prefix product), the same call as the benchmark's CPU baseline,
  about 14 s per 7-level point on 16 host cores.  Bar: 1e-10 relative to
  the trace's max |value|.  The oracle itself agrees with scipy ``expm``
  applied step by step to 5.5e-13 (7-level, 2.73 GHz) and 3.6e-14
  (3-level) over these horizons.
* Oracle at full length for the pulsed protocol batch: the shortest
  Hahn-echo and the shortest Ramsey protocol of the 16384-protocol batch,
  whole (about 58k-65k steps, the last 50k the 1 us readout window),
  through the persistent queue kernel and through the reference algorithm
  with an output at every step, integrated over the readout window the same
  way; 1e-10 relative.
* Vectorisation identity: the 7-level model vectorised coherently
  (n = 49) and compressed (n = 19, coherences between incoherently coupled
  blocks dropped, which stay zero from the states used here) is the same
  physics, so the two must give the same fluorescence at the workloads'
  full lengths: 4 sweep points x 99,900 steps, and the 8 longest pulsed
  protocols (up to 742k steps, integrated readout).
"""

import numpy as np
import pytest

from oracle.reference_simulate import flatten
from oracle.reference_simulate import generate_simulator as oracle_gen
from oracle.waves import vectorised_sampler
from superspinsim_amd import configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _oracle_full_trace(wl, wave):
    sim = oracle_gen(None, wl.generators, wl.vmap, backend="c",
                     number_of_exponentials=wl.number_of_exponentials,
                     number_of_fine_divisions=wl.fine,
                     vectorised_sampler=vectorised_sampler(wave))
    _, dens = sim(wl.rho0_matrix, wl.time_start, wl.time_end, wl.time_step)
    flat = np.stack([flatten(d, wl.vmap, True, wl.n) for d in dens])
    return flat @ wl.obs[0]


@pytest.mark.parametrize("factory,freq", [
    (configs.odmr7, 2.73e9),      # ms = 0 <-> -1 line, below the band
    (configs.odmr7, 2.8713e9),    # inside the band
    (configs.odmr3, 2.87e9),
])
def test_full_length_sweep_point_vs_oracle(gpu, factory, freq):
    wl = factory(1)
    wl.frequencies = np.array([freq])
    waves = wl.waveforms()
    sim = wl.simulator()
    print(f"{wl.name} f={freq:.4e}: {wl.steps_per_instance()} steps on the "
          "GPU", flush=True)
    out = sim.run(waves, wl.rho0, wl.time_start, wl.time_end, wl.time_step,
                  obs=wl.obs, read="obs")
    assert sim.last_plan["steps"] == wl.steps_per_instance()
    dev = out[0, :, 0].cpu().numpy()
    print("  oracle (cpu_ref.c) at full length", flush=True)
    ref = _oracle_full_trace(wl, waves[0])
    assert dev.shape == ref.shape == (wl.n_out,)
    err = np.max(np.abs(dev - ref)) / np.max(np.abs(ref))
    print(f"  rel err {err:.3e} over {wl.n_out} outputs, max "
          f"{np.max(np.abs(ref)):.4e}", flush=True)
    assert err <= TOL
    assert np.max(np.abs(ref)) > 1e-3


def test_full_length_chunked_3_level_vs_oracle(gpu):
    """The 3-level sweep by time chunks (its benchmark path): one sweep
    point at full length against the oracle, and the other points of a
    16-point batch against the sequential run."""
    wl = configs.odmr3(16)
    wl.frequencies = np.linspace(2.77e9, 2.97e9, 16)
    waves = wl.waveforms()
    sim = wl.simulator()
    args = (waves, wl.rho0, wl.time_start, wl.time_end, wl.time_step)
    got = sim.run_chunked(*args, obs=wl.obs, read="obs")[:, :, 0]
    got = got.cpu().numpy()
    plan = sim.last_plan
    print(f"{wl.name}: 16 x {wl.steps_per_instance()} steps in "
          f"{plan['chunks']} chunks of {plan['chunk_out']} outputs",
          flush=True)
    seq = sim.run(*args, obs=wl.obs, read="obs")[:, :, 0].cpu().numpy()
    diff = np.max(np.abs(got - seq)) / np.max(np.abs(seq))
    print(f"  chunked vs sequential rel diff {diff:.3e}", flush=True)
    assert diff <= 1e-11
    ref = _oracle_full_trace(wl, waves[7])
    err = np.max(np.abs(got[7] - ref)) / np.max(np.abs(ref))
    print(f"  chunked vs oracle rel err {err:.3e}", flush=True)
    assert err <= TOL


def test_full_length_protocols_vs_oracle(gpu):
    from superspinsim_amd.protocols import trapezoid_window
    full = configs.pulsed7()
    echo = full.params["echo"]
    picks = [int(np.flatnonzero(echo)[np.argmin(full.steps[echo])]),
             int(np.flatnonzero(~echo)[np.argmin(full.steps[~echo])])]
    waves = [full.waves[i] for i in picks]
    batch = full.batch().prepare(waves, full.rho0, full.time_step,
                                 full.steps[picks], full.read_from[picks],
                                 full.obs)
    print(f"{full.name}: protocols {picks}, steps "
          f"{full.steps[picks].tolist()} on the GPU", flush=True)
    out = batch.run().cpu().numpy()[:, 0]
    base = configs.odmr7(1)
    for k, i in enumerate(picks):
        print(f"  protocol {i} (echo {bool(echo[i])}): oracle at every step",
              flush=True)
        sim = oracle_gen(None, full.generators, base.vmap, backend="c",
                         number_of_exponentials=2, number_of_fine_divisions=1,
                         vectorised_sampler=vectorised_sampler(full.waves[i]))
        T = int(full.steps[i])
        rho = np.zeros((7, 7))
        rho[1, 1] = 1.0
        _, dens = sim(rho, 0.0, T * full.time_step * (1 + 1e-12),
                      full.time_step)
        assert dens.shape[0] == T
        pops = np.real(dens[:, 3, 3] + dens[:, 4, 4] + dens[:, 5, 5])
        want = trapezoid_window(np.concatenate([[0.0], pops])[None],
                                full.read_from[i:i + 1],
                                full.steps[i:i + 1], full.time_step)[0]
        err = abs(out[k] - want) / abs(want)
        print(f"  integrated fluorescence {out[k]:.12e}, rel err {err:.3e}",
              flush=True)
        assert err <= TOL


def test_odmr7_full_length_vectorisations_agree(gpu):
    freqs = np.array([2.73e9, 2.78e9, 2.87e9, 2.96e9])
    traces = []
    for inc in (False, True):
        wl = configs.odmr7(4, use_incoherent=inc)
        wl.frequencies = freqs
        sim = wl.simulator()
        print(f"{wl.name}: 4 x {wl.steps_per_instance()} steps", flush=True)
        out = sim.run(wl.waveforms(), wl.rho0, wl.time_start, wl.time_end,
                      wl.time_step, obs=wl.obs, read="obs")
        assert sim.last_plan["steps"] == 99900
        traces.append(out[:, :, 0].cpu().numpy())
    a, b = traces
    err = np.max(np.abs(a - b)) / np.max(np.abs(a))
    print(f"  n49 vs n19 rel diff {err:.3e}")
    assert err <= TOL
    assert np.max(a) > 1e-3


def test_pulsed_protocols_full_length_vectorisations_agree(gpu):
    full = configs.pulsed7()
    longest = np.argsort(-full.steps, kind="stable")[:8]
    outs = []
    for inc in (False, True):
        wl = configs.pulsed7(use_incoherent=inc)
        waves = [wl.waves[i] for i in longest]
        batch = wl.batch().prepare(waves, wl.rho0, wl.time_step,
                                   wl.steps[longest], wl.read_from[longest],
                                   wl.obs)
        print(f"{wl.name} n={batch.prop.n}: 8 protocols, "
              f"{batch.total_steps()} steps", flush=True)
        outs.append(batch.run().cpu().numpy()[:, 0])
    a, b = outs
    assert np.min(full.steps[longest]) > 5e5
    err = np.max(np.abs(a - b) / np.abs(a))
    print(f"  n49 vs n19 rel diff {err:.3e}")
    assert err <= TOL
