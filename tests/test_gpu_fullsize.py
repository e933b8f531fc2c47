"""Full-length runs checked through a size-independent property (GPU).

The 7-level model vectorised coherently (n = 49) and compressed (n = 19,
coherences between incoherently coupled blocks dropped, which stay zero
from the states used here) is the same physics: at the workloads' full
lengths the two must give the same fluorescence.  The reference-algorithm
oracle cannot run these lengths (its cost per step is 70 dense squarings);
short runs of both sizes match it to 1e-10 (test_gpu_sweep.py,
test_gpu_protocols.py).

* the 7-level ODMR sweep: 4 sweep points x 99,900 steps, every output;
* the pulsed protocols: the 8 longest of the 16,384 (up to 742k steps),
  integrated readout.
"""

import numpy as np
import pytest

from superspinsim_amd import configs

pytestmark = pytest.mark.gpu

TOL = 1e-10


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
