"""Taylor degree and the three forms of the chain (GPU).

The n <= 64 kernels evaluate exp(A) x as a Taylor series.  Degree 11 runs
a chain unrolled for exactly that degree, in a kernel of its own.  Degrees
1-16 otherwise run a chain unrolled to 16 terms that stops after term m.
Higher degrees run a loop.  At the 7-level model's 20 ps steps the
planned degree is 11 (1e-16 tolerance).  The first omitted term of degree
m is at most theta^(m+1) / (m+1)! relative per exponential, with theta
about 0.25 here: 5e-15 for m = 10, 1e-16 for m = 11, 2e-18 for m = 12.

* Degrees 13-16 (unrolled chain with exits) equal degree 17 (the loop) to
  1e-14 relative over 2,000 steps.  Their extra terms are below 1e-17, and
  a wrong 1/t or a dropped or repeated term would be off by far more.
* Degrees 11 (exact chain) and 12 equal degree 17 to 1e-12.
* Low degrees differ from degree 17 by the truncated terms, and the
  difference falls with the degree (3 > 5 > 7 > 9), so the chain really
  stops after term m; degree 9 is within 1e-13 of degree 17.
* Degree 11 matches the reference algorithm (oracle) to 1e-10.
* The pulsed-protocol queue kernel: degrees 13 and 16 equal its degree 17
  to 1e-14, degree 11 to 1e-12, degree 10 to 1e-10.
"""

import numpy as np
import pytest

from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd import configs

pytestmark = pytest.mark.gpu

N_OUT = 20
SAME = 1e-14
TOL = 1e-10


@pytest.fixture(scope="module")
def sweep_runs(gpu):
    wl = configs.odmr7(64)
    wl.frequencies = np.linspace(2.80e9, 2.94e9, 64)
    sim = wl.simulator()
    waves = wl.waveforms()
    t_end = N_OUT * wl.time_step * (1 + 1e-12)
    out = {}
    for m in (3, 5, 7, 9, 11, 12, 13, 14, 15, 16, 17):
        run = sim.prepare(waves, wl.rho0, 0.0, t_end, wl.time_step,
                          obs=wl.obs, read="obs", taylor=m)
        run.launch()
        out[m] = run.result().cpu().numpy()[:, :, 0].copy()
    planned = sim.plan(waves, wl.time_step)[2]
    return wl, waves, out, planned


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_planned_degree_is_eleven(sweep_runs):
    assert sweep_runs[3] == 11


def test_unrolled_chain_equals_loop(sweep_runs):
    _, _, out, _ = sweep_runs
    for m in (13, 14, 15, 16):
        d = _rel(out[m], out[17])
        print(f"degree {m} vs 17: {d:.1e}", flush=True)
        assert d <= SAME, m


def test_exact_chain_equals_long_series(sweep_runs):
    _, _, out, _ = sweep_runs
    for m in (11, 12):
        d = _rel(out[m], out[17])
        print(f"degree {m} vs 17: {d:.1e}", flush=True)
        assert d <= 1e-12, m


def test_low_degrees_truncate(sweep_runs):
    _, _, out, _ = sweep_runs
    errs = [_rel(out[m], out[17]) for m in (3, 5, 7, 9)]
    print("degree 3/5/7/9 vs 17:", " ".join(f"{e:.1e}" for e in errs))
    assert errs[0] > 1e-8
    assert all(a > b for a, b in zip(errs, errs[1:]))
    # the degree-9 bound theta^10 / 10! is 8e-14 per exponential: a wrong
    # constant in terms 1-9 of the chain with exits would show here
    assert errs[-1] <= 1e-13


def test_degree_eleven_matches_oracle(sweep_runs):
    wl, waves, out, _ = sweep_runs
    for i in (0, 63):
        sim = oracle_gen(waves[i].sampler(), wl.generators, wl.vmap,
                         backend="c", number_of_exponentials=2,
                         number_of_fine_divisions=wl.fine)
        _, dens = sim(wl.rho0_matrix, 0.0,
                      N_OUT * wl.time_step * (1 + 1e-12), wl.time_step)
        flat = np.stack([np.array([
            (d[y, x].real if c == 0 else d[y, x].imag)
            for (y, x, c) in wl.vmap]) for d in dens])
        assert _rel(out[11][i], flat @ wl.obs[0]) <= TOL


def test_queue_kernel_degrees(gpu):
    wl = configs.pulsed7(count=24, seed=5, t_p_range=(2e-9, 20e-9),
                         tau_range=(10e-9, 100e-9), readout=50e-9)
    res = {}
    for m in (10, 11, 13, 16, 17):
        b = wl.batch().prepare(wl.waves, wl.rho0, 20e-12, wl.steps,
                               wl.read_from, wl.obs, taylor=m)
        res[m] = b.run().cpu().numpy()[:, 0].copy()
    for m, tol in ((13, SAME), (16, SAME), (11, 1e-12), (10, 1e-10)):
        d = _rel(res[m], res[17])
        print(f"queue degree {m} vs 17: {d:.1e}", flush=True)
        assert d <= tol, m
