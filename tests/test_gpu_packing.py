"""Several instances of a small model per wavefront (GPU).

``BatchSimulator.run`` packs k instances of an n <= 32 model into one
block-diagonal system of k n rows (constant block repeated, each varying
generator once per member with that member's waveform), so one wavefront
advances k instances.  A row of block j sums exactly the single-instance
row's terms plus exact zeros; only the ELL row layout of the larger system
orders them differently.  Checked:

* the 7-level model at n = 19, 2 per system, an odd instance count (the
  last system filled with a copy): per-output, integral and full-state
  readouts equal the unpacked run to 1e-13 relative, and match the
  reference algorithm (oracle) to 1e-10;
* the 3-level model (n = 9) with 7 per system and a full-state readout;
* the automatic choice on a 4096-point sweep packs 2 per system.
"""

import numpy as np
import pytest

from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd import configs

pytestmark = pytest.mark.gpu

SAME = 1e-13
TOL = 1e-10


def _run(wl, waves, n_out, pack, read):
    sim = wl.simulator()
    sim.pack = pack
    t_end = n_out * wl.time_step * (1 + 1e-12)
    out = sim.run(waves, wl.rho0, 0.0, t_end, wl.time_step,
                  obs=None if read == "full" else wl.obs, read=read)
    return out.cpu().numpy(), sim.last_plan["packed"]


def _rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("read", ["obs", "integral", "full"])
def test_packed_n19_equals_unpacked(gpu, read):
    wl = configs.odmr7(5, use_incoherent=True)
    wl.frequencies = np.linspace(2.80e9, 2.94e9, 5)
    waves = wl.waveforms()
    packed, k = _run(wl, waves, 6, 2, read)
    plain, k1 = _run(wl, waves, 6, False, read)
    assert (k, k1) == (2, 1)
    assert packed.shape == plain.shape and packed.shape[0] == 5
    d = _rel(packed, plain)
    print(f"n=19 {read}: packed vs unpacked {d:.1e}", flush=True)
    assert d <= SAME
    if read == "obs":
        for i in (0, 4):
            osim = oracle_gen(waves[i].sampler(), wl.generators, wl.vmap,
                              backend="c", number_of_exponentials=2,
                              number_of_fine_divisions=wl.fine)
            _, dens = osim(wl.rho0_matrix, 0.0,
                           6 * wl.time_step * (1 + 1e-12), wl.time_step)
            flat = np.stack([np.array([
                (dm[y, x].real if c == 0 else dm[y, x].imag)
                for (y, x, c) in wl.vmap]) for dm in dens])
            want = flat @ wl.obs[0]
            err = _rel(packed[i, :, 0], want)
            print(f"  instance {i} vs oracle {err:.1e}")
            assert err <= TOL


def test_packed_n9_seven_per_system(gpu):
    wl = configs.odmr3(9)
    wl.frequencies = np.linspace(2.85e9, 2.89e9, 9)
    waves = wl.waveforms()
    packed, k = _run(wl, waves, 20, 7, "full")
    plain, _ = _run(wl, waves, 20, False, "full")
    assert k == 7
    d = _rel(packed, plain)
    print(f"n=9 full, 7 per system: {d:.1e}", flush=True)
    assert d <= SAME


def test_automatic_packing_on_full_sweep(gpu):
    wl = configs.odmr7(4096, use_incoherent=True)
    waves = wl.waveforms()
    auto, k = _run(wl, waves, 5, True, "obs")
    plain, _ = _run(wl, waves, 5, False, "obs")
    print(f"4096 points: {k} per system, diff {_rel(auto, plain):.1e}",
          flush=True)
    assert k == 2
    assert _rel(auto, plain) <= SAME
