"""Superoperators wider than one wavefront (n > 64) on the GPU.

The 21-level NV + 14N model: n = 441 (coherent vectorisation, 7 wavefronts
per instance, 2 instances per workgroup) and n = 171 (compressed, 3
wavefronts, 1 instance per workgroup, several workgroups per CU).

* Short runs against the reference-algorithm oracle (numpy backend of the
  restated CPU path): every component of every output state, and two
  readout rows at once; 1e-10 relative to the largest |component|.
  At n = 441 the odd count leaves an idle slot in the last workgroup.
* Full-length sweep points (250,000 steps): the coherent and compressed
  vectorisations of the same physics give the same fluorescence (the
  compressed one drops coherences between incoherently coupled blocks,
  which stay zero), and the state keeps unit trace.
* The integral readout equals the trapezoid of the per-step trace.
* The drop-in ``generate_simulator`` on the 21-level model (n = 171; the
  rotating frame is requested and, as in the reference, refused because
  L0 is singular) against the oracle.
"""

import math

import numpy as np
import pytest

from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd import configs
from superspinsim_amd.generate_simulator import generate_simulator
from superspinsim_amd.sugar import prepare_simspins

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _block_state(wl, seed):
    """Random density matrix, block-diagonal over the three orbital blocks
    (ground 0-8, excited 9-17, singlet 18-20)."""
    rng = np.random.default_rng(seed)
    rho = np.zeros((21, 21), dtype=complex)
    for lo, hi in ((0, 9), (9, 18), (18, 21)):
        k = hi - lo
        a = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
        rho[lo:hi, lo:hi] = a @ a.conj().T
    return rho / np.trace(rho).real


def _short_workload(count, use_incoherent, rabi_hz=20e6):
    wl = configs.odmr21(count, use_incoherent=use_incoherent)
    wl.b1 = configs.drive_amplitude(rabi_hz)
    wl.fine = 4
    wl.time_step = 1e-10                 # h = 25 ps
    wl.rho0_matrix = _block_state(wl, 5)
    wl.rho0 = wl._flat(wl.rho0_matrix)
    return wl


def _oracle_states(wl, wave, n_out):
    sim = oracle_gen(wave.sampler(), wl.generators, wl.vmap,
                     backend="numpy",
                     number_of_exponentials=wl.number_of_exponentials,
                     number_of_fine_divisions=wl.fine)
    _, dens = sim(wl.rho0_matrix, 0.0, n_out * wl.time_step * (1 + 1e-12),
                  wl.time_step)
    return np.stack([np.array([
        (d[y, x].real if c == 0 else d[y, x].imag)
        for (y, x, c) in wl.vmap]) for d in dens])


@pytest.mark.parametrize("use_incoherent,count", [(True, 7), (False, 3)])
def test_wide_states_vs_oracle(gpu, use_incoherent, count):
    wl = _short_workload(count, use_incoherent)
    n_out = 2
    waves = wl.waveforms()
    sim = wl.simulator()
    print(f"{wl.name}: {count} instances x {n_out} outputs", flush=True)
    t_end = n_out * wl.time_step * (1 + 1e-12)
    full = sim.run(waves, wl.rho0, 0.0, t_end, wl.time_step,
                   read="full").cpu().numpy()
    obs = np.stack([wl.obs[0], np.zeros(wl.n)])
    obs[1, [wl._index(k, k) for k in (3, 4, 5)]] = 1.0
    two = sim.run(waves, wl.rho0, 0.0, t_end, wl.time_step, obs=obs,
                  read="obs").cpu().numpy()
    assert sim.last_plan["n"] == wl.n > 64
    for k, w in enumerate(waves):
        ref = _oracle_states(wl, w, n_out)
        scale = np.max(np.abs(ref))
        err = np.max(np.abs(full[k] - ref)) / scale
        print(f"  f={wl.frequencies[k]:.6e}: state rel err {err:.3e}")
        assert err <= TOL
        err_obs = np.max(np.abs(two[k] - ref @ obs.T)) / scale
        assert err_obs <= TOL


@pytest.mark.parametrize("use_incoherent,n_out", [(True, 400),
                                                   (False, 20)])
def test_wide_longer_run_vs_c_oracle(gpu, use_incoherent, n_out):
    """One sweep point over 1600 steps (n = 171) and 80 steps (n = 441)
    against the reference algorithm in C (70 dense squarings of the
    171- or 441-dimensional exponentials per step): every state, 1e-10
    relative to the largest entry."""
    from oracle.waves import vectorised_sampler
    wl = _short_workload(1, use_incoherent)
    waves = wl.waveforms()
    sim = wl.simulator()
    t_end = n_out * wl.time_step * (1 + 1e-12)
    print(f"{wl.name}: {n_out * wl.fine} steps on the GPU", flush=True)
    full = sim.run(waves, wl.rho0, 0.0, t_end, wl.time_step,
                   read="full").cpu().numpy()[0]
    print("  oracle (cpu_ref.c)", flush=True)
    osim = oracle_gen(None, wl.generators, wl.vmap, backend="c",
                      number_of_exponentials=wl.number_of_exponentials,
                      number_of_fine_divisions=wl.fine,
                      vectorised_sampler=vectorised_sampler(waves[0]))
    _, dens = osim(wl.rho0_matrix, 0.0, t_end, wl.time_step)
    ref = np.stack([np.array([
        (d[y, x].real if c == 0 else d[y, x].imag)
        for (y, x, c) in wl.vmap]) for d in dens])
    assert full.shape == ref.shape == (n_out, wl.n)
    err = np.max(np.abs(full - ref)) / np.max(np.abs(ref))
    print(f"  n={wl.n}: state rel err {err:.3e}", flush=True)
    assert err <= TOL


def test_wide_vectorisations_agree_full_length(gpu):
    """250,000 steps at four sweep points, n = 441 vs n = 171."""
    freqs = np.array([2.8656e9, 2.8678e9, 2.8700e9, 2.8722e9])
    traces = []
    for inc in (False, True):
        wl = configs.odmr21(4, use_incoherent=inc)
        wl.frequencies = freqs
        sim = wl.simulator()
        waves = wl.waveforms()
        print(f"{wl.name}: 4 x {wl.n_out} outputs x {wl.fine}", flush=True)
        out = sim.run(waves, wl.rho0, 0.0, wl.time_end, wl.time_step,
                      obs=wl.obs, read="obs")
        traces.append(out[:, :, 0].cpu().numpy())
        assert sim.last_plan["steps"] == 250000
    a, b = traces
    err = np.max(np.abs(a - b)) / np.max(np.abs(a))
    print(f"  n441 vs n171 rel diff {err:.3e}")
    assert err <= TOL
    assert np.all(np.isfinite(a)) and np.max(a) > 1e-3


def test_wide_trace_preserved(gpu):
    wl = configs.odmr21(2, use_incoherent=True)
    sim = wl.simulator()
    waves = wl.waveforms()
    n_out = 50
    full = sim.run(waves, wl.rho0, 0.0, n_out * wl.time_step * (1 + 1e-12),
                   wl.time_step, read="full").cpu().numpy()
    diag = [wl._index(k, k) for k in range(21)]
    tr = full[:, :, diag].sum(axis=2)
    assert np.max(np.abs(tr - 1.0)) < 1e-12


def test_wide_integral_is_trapezoid(gpu):
    wl = configs.odmr21(3, use_incoherent=True)
    sim = configs.BatchSimulator(wl.generators, 2, 1)
    waves = wl.waveforms()
    h = wl.time_step / wl.fine
    n = 300
    trace = sim.run(waves, wl.rho0, 0.0, n * h * (1 + 1e-12), h,
                    obs=wl.obs, read="obs")[:, :, 0].cpu().numpy()
    integ = sim.run(waves, wl.rho0, 0.0, n * h * (1 + 1e-12), h,
                    obs=wl.obs, read="integral")[:, 0].cpu().numpy()
    f0 = wl.obs[0] @ wl.rho0
    full = np.concatenate([np.full((3, 1), f0), trace], axis=1)
    ref = np.sum(0.5 * h * (full[:, 1:] + full[:, :-1]), axis=1)
    assert np.max(np.abs(integ - ref) / np.abs(ref)) < 1e-13


def test_wide_dropin_generate_simulator_vs_oracle(gpu):
    b1 = configs.drive_amplitude(20e6)
    coeffs = [lambda t: b1 * math.cos(math.tau * 2.8678e9 * t),
              lambda t: 0.0, lambda t: 0.0, lambda t: 5 / 65]
    spins, si, gi = configs._nv21_description()
    sampler, gens, vmap, kw = prepare_simspins(
        coeffs, spins, si, gi, number_of_exponentials=2,
        number_of_fine_divisions=4, use_rotating=True, use_incoherent=True)
    assert gens[0].shape == (171, 171) and not kw["use_rotating"]
    rho = np.zeros((21, 21))
    for k in (3, 4, 5):
        rho[k, k] = 1 / 3
    rho[3, 4] = rho[4, 3] = 0.1
    print("nv21 drop-in n=171, 3 outputs", flush=True)
    t_dev, d_dev = generate_simulator(sampler, gens, vmap, **kw)(
        rho, 0, 3e-10, 1e-10)
    t_ref, d_ref = oracle_gen(sampler, gens, vmap, backend="numpy", **kw)(
        rho, 0, 3e-10, 1e-10)
    assert np.array_equal(t_dev, t_ref)
    err = np.max(np.abs(d_dev - d_ref)) / np.max(np.abs(d_ref))
    print(f"  rel err {err:.3e}")
    assert err <= TOL


def test_wide_rotating_frame_vs_oracle(gpu):
    """Rotating frame at n = 98 (node-wise generators, eigen-step) against
    the oracle's rotating path on the same tables.

    L0 is given by its real block eigen-data (45 complex pairs, 8 real
    eigenvalues) in the standard basis (V = I), so the interaction-picture
    generators keep the block pattern of two random sparse generators; one
    has a constant coefficient, the other a microwave tone."""
    from superspinsim_amd.sweep import BatchSimulator, Waveform
    from oracle.reference_simulate import flatten
    rng = np.random.default_rng(21)
    nd, ns = 45, 8
    n = 2 * nd + ns                                  # 98 = 2 * 7 * 7
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
    rho = rng.standard_normal((7, 7)) + 1j * rng.standard_normal((7, 7))
    x0 = flatten(rho, None, False, n)
    sim = BatchSimulator(gens, 2, F, use_rotating=True, vectors_real=eye,
                         inv_vectors_real=eye, doubles=doubles,
                         singles=singles)
    t_end = n_out * dt * (1 + 1e-12)
    out = sim.run(waves, x0, 0.0, t_end, dt, read="full").cpu().numpy()
    assert sim.last_plan["n"] == n > 64
    for k, w in enumerate(waves):
        osim = oracle_gen(w.sampler(), gens, None, use_hermitian=False,
                          use_rotating=True, vectors_real=eye,
                          inv_vectors_real=eye, doubles=doubles,
                          singles=singles, number_of_exponentials=2,
                          number_of_fine_divisions=F, backend="numpy")
        _, dens = osim(rho, 0.0, t_end, dt)
        ref = np.stack([flatten(d, None, False, n) for d in dens])
        err = np.max(np.abs(out[k] - ref)) / np.max(np.abs(ref))
        print(f"  f={freqs[k]:.2e}: rotating n=98 rel err {err:.3e}")
        assert err <= TOL
