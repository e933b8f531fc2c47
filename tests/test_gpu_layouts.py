"""ELL row layouts of the n <= 64 kernels change speed, not results (GPU).

Layouts: column-by-column gathers ("plain"), the quad order (the other pair
of a row's quad in one 16-byte LDS read) and the pair-fetch order (every
slot pair one 16-byte read), each with and without the bank-aware slot
order.  The kernels pick their read path from the column table, so each
case below also exercises that detection.

* the 7-level sweep at n = 49 and n = 19: every layout equals the plain one
  to 1e-13 relative (only the order of the row sums changes), and the quad
  and pair layouts match the reference algorithm (oracle) to 1e-10;
* the full-state readout comes back in the caller's basis under every
  layout;
* the persistent queue kernel (pulsed protocols): every layout equals the
  plain one to 1e-13;
* the rotating frame at n = 50 (2x2 eigen-blocks kept together, the
  eigen-step factors permuted with the rows): every layout against the
  oracle's rotating path, 1e-10.
"""

import functools

import numpy as np
import pytest

from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd import configs, propagator, sweep

pytestmark = pytest.mark.gpu

LAYOUTS = [("plain", False), ("quad", False), ("quad", True),
           ("pairs", False), ("pairs", True)]
SAME = 1e-13
TOL = 1e-10


@pytest.fixture
def with_layout(monkeypatch):
    """Point the simulators' propagator factory at one bank-aware setting;
    the simulator's ``layout`` attribute picks the layout."""
    def set_bank(bank):
        monkeypatch.setattr(sweep, "Propagator",
                            functools.partial(propagator.Propagator,
                                              bank_aware=bank))
    return set_bank


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


@pytest.mark.parametrize("use_incoherent", [False, True])
def test_sweep_layouts_agree(gpu, with_layout, use_incoherent):
    wl = configs.odmr7(3, use_incoherent=use_incoherent)
    wl.frequencies = np.array([2.80e9, 2.87e9, 2.94e9])
    waves = wl.waveforms()
    n_out = 6
    t_end = n_out * wl.time_step * (1 + 1e-12)
    outs, fulls = {}, {}
    for layout, bank in LAYOUTS:
        with_layout(bank)
        sim = wl.simulator()
        sim.layout = layout
        out = sim.run(waves, wl.rho0, 0.0, t_end, wl.time_step, obs=wl.obs,
                      read="obs")
        full = sim.run(waves, wl.rho0, 0.0, t_end, wl.time_step,
                       read="full")
        prop = next(iter(sim._prop_cache.values()))
        print(f"n={wl.n} {layout} bank={bank}: {prop.ell.layout}, "
              f"{prop.ell.slots} slots", flush=True)
        assert prop.ell.layout == layout
        outs[layout, bank] = out[:, :, 0].cpu().numpy()
        fulls[layout, bank] = full.cpu().numpy()
    ref, ref_full = outs["plain", False], fulls["plain", False]
    for key in outs:
        d = np.max(np.abs(outs[key] - ref)) / np.max(np.abs(ref))
        df = np.max(np.abs(fulls[key] - ref_full)) / \
            np.max(np.abs(ref_full))
        print(f"  {key}: obs diff {d:.1e}, full diff {df:.1e}")
        assert d <= SAME and df <= SAME
    # the fetch layouts against the reference algorithm
    for key in (("quad", True), ("pairs", True)):
        for k in (0, 2):
            want = _oracle_trace(wl, waves[k], n_out)
            err = np.max(np.abs(outs[key][k] - want)) / np.max(np.abs(want))
            assert err <= TOL, (key, k, err)


def test_queue_layouts_agree(gpu, with_layout):
    wl = configs.pulsed7(count=12, seed=7, t_p_range=(2e-9, 10e-9),
                         tau_range=(10e-9, 40e-9), readout=20e-9)
    h = 20e-12
    outs = {}
    for layout, bank in LAYOUTS:
        with_layout(bank)
        b = wl.batch()
        b.sim.layout = layout
        b.prepare(wl.waves, wl.rho0, h, wl.steps, wl.read_from, wl.obs)
        assert b.prop.ell.layout == layout
        outs[layout, bank] = b.run().cpu().numpy()[:, 0].copy()
    ref = outs["plain", False]
    for key, got in outs.items():
        d = np.max(np.abs(got - ref) / np.abs(ref))
        print(f"  queue {key}: diff {d:.1e}", flush=True)
        assert d <= SAME


def test_rotating_layouts_vs_oracle(gpu, with_layout):
    """Rotating frame at n = 50 (21 complex eigen-pairs, 8 real
    eigenvalues, V = I): the quad order moves whole 2x2 blocks and their
    eigen-step factors; the pair layout keeps the order."""
    from oracle.reference_simulate import flatten
    from superspinsim_amd.sweep import BatchSimulator, Waveform
    rng = np.random.default_rng(33)
    nd, ns = 21, 8
    n = 2 * nd + ns                                  # 50 = 2 * 5 * 5
    doubles = np.stack([-rng.uniform(1e6, 1e8, nd),
                        rng.uniform(-2e10, 2e10, nd)], axis=1)
    singles = -rng.uniform(1e5, 1e7, ns)
    units = [(2 * k, 2 * k + 2) for k in range(nd)] + \
        [(2 * nd + k, 2 * nd + k + 1) for k in range(ns)]
    gens = np.zeros((2, n, n))
    for g in range(2):
        for (r0, r1) in units:
            for c in rng.choice(len(units), 3, replace=False):
                c0, c1 = units[c]
                gens[g, r0:r1, c0:c1] = rng.standard_normal((r1 - r0,
                                                             c1 - c0))
    gens *= 2e8 / np.abs(gens).sum(axis=2).max()
    eye = np.eye(n)
    F, dt, n_out = 4, 8e-11, 3                       # h = 20 ps
    freqs = [2.1e9, 2.9e9]
    waves = []
    for f in freqs:
        w = Waveform(2, base=[1.0, 0.0])
        w.tone(1, 0.7, f)
        waves.append(w)
    rho = rng.standard_normal((5, 5)) + 1j * rng.standard_normal((5, 5))
    x0 = flatten(rho, None, False, n)
    t_end = n_out * dt * (1 + 1e-12)
    refs = []
    for w in waves:
        osim = oracle_gen(w.sampler(), gens, None, use_hermitian=False,
                          use_rotating=True, vectors_real=eye,
                          inv_vectors_real=eye, doubles=doubles,
                          singles=singles, number_of_exponentials=2,
                          number_of_fine_divisions=F, backend="numpy")
        _, dens = osim(rho, 0.0, t_end, dt)
        refs.append(np.stack([flatten(d, None, False, n) for d in dens]))
    for layout, bank in LAYOUTS:
        with_layout(bank)
        sim = BatchSimulator(gens, 2, F, use_rotating=True,
                             vectors_real=eye, inv_vectors_real=eye,
                             doubles=doubles, singles=singles,
                             layout=layout)
        out = sim.run(waves, x0, 0.0, t_end, dt, read="full").cpu().numpy()
        prop = next(iter(sim._prop_cache.values()))
        assert prop.ell.layout == layout and prop.n == n
        if layout == "quad":
            # blocks moved as units
            perm = prop.perm
            assert np.all(perm[1::2] == perm[0::2] + 1)
        for k in range(len(waves)):
            err = np.max(np.abs(out[k] - refs[k])) / np.max(np.abs(refs[k]))
            print(f"  rotating {layout} bank={bank} f={freqs[k]:.1e}: "
                  f"rel err {err:.2e}", flush=True)
            assert err <= TOL
