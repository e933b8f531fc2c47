"""Adjoint gradients for superoperators wider than one wavefront (GPU).

The 21-level NV + 14N model (n = 171 compressed, n = 441 coherent) with
gradients of the integrated fluorescence with respect to the 7-level
parameters and the 14N quadrupole / hyperfine constants (12 in all).  A
20 MHz drive and a random block-diagonal initial state (coherences in every
orbital block) make every parameter but T2* enter at first order.  T2*
enters no generator of this model: the reference's label combination drops
the dephasing of an atom with both an electron and a nuclear spin
(gg.py:2226-2314), so its gradient is exactly 0.

* the gradient oracle (oracle/adjoint.py: dense scipy expm and
  expm_frechet, pinned by tests/test_oracle_adjoint.py) on short runs
  whose checkpoint segments have uneven lengths: F, every d F / d theta
  and d F / d rho0, rtol 1e-8;
* d F / d rho0: F is linear in rho0, so lambda . rho0 = F;
* the checkpoint interval does not change the gradients (bitwise: the
  recomputation reproduces the forward states exactly);
* central differences of the HIP forward at the 1e-6 relative step,
  within their error bound.
"""

import numpy as np
import pytest

from superspinsim_amd import configs

pytestmark = pytest.mark.gpu

FREQS = [2.8678e9, 2.8700e9, 2.8722e9]


def _block_rho0(wl, seed=11):
    rng = np.random.default_rng(seed)
    rho = np.zeros((21, 21), dtype=complex)
    for lo, hi in ((0, 9), (9, 18), (18, 21)):
        k = hi - lo
        a = rng.standard_normal((k, k)) + 1j * rng.standard_normal((k, k))
        rho[lo:hi, lo:hi] = a @ a.conj().T
    rho /= np.trace(rho).real
    return wl._flat(rho)


def _problem(use_incoherent, time_end, time_step, fine, checkpoint_every,
             freqs=FREQS):
    wl = configs.odmr21_gradient(
        count=len(freqs), time_end=time_end, time_step=time_step, fine=fine,
        checkpoint_every=checkpoint_every, use_incoherent=use_incoherent,
        frequencies=freqs, rabi_hz=20e6)
    base = configs.odmr21(1, use_incoherent=use_incoherent)
    wl.problem.rho0 = _block_rho0(base)
    return wl


def _run(wl):
    pb = wl.problem
    out, state = pb.forward(wl.theta, wl.amplitudes)
    n_inst = state["n_inst"]
    res = pb.backward(state, np.ones((n_inst, 1)), lambda_out=True)
    print("n=%d Taylor %d adjoint %d slots %d/%d/%d"
          % (state["prop"].n, state["taylor"], state["taylor_adj"],
             state["slots_f"], state["slots_t"], state["slots"]),
          flush=True)
    return (out.cpu().numpy()[:, 0],
            {k: v.cpu().numpy() for k, v in res.items()}, state)


def _oracle_factors(problem, state, inst):
    """Per-step factors (T, J, M) of one instance at the reference node
    times, from its kernel waveform."""
    from superspinsim_amd.generate_simulator import tabulate
    prop = state["prop"]
    kw = state["kernel_waves"][inst]
    M = prop.K
    h = problem.dt / problem.F
    T = problem.n_out * problem.F
    rows = tabulate(lambda t, row: row.__setitem__(slice(None), kw(t)), M,
                    problem.n_out, problem.F, prop.nodes, problem.t0,
                    problem.dt)
    rows = rows.reshape(T, prop.S, M)
    return h * np.einsum("js,tsm->tjm", prop.weights, rows)


def _check_vs_oracle(wl, F, res, state, count):
    from oracle import adjoint as oadj
    pb = wl.problem
    prop = state["prop"]
    h = pb.dt / pb.F
    for i in range(count):
        factors = _oracle_factors(pb, state, i)
        Fo, go, _, lam = oadj.gradient(prop._dense, factors, pb.rho0,
                                       pb.obs, [1.0], h,
                                       dmats=state["dconst"], theta_gen=0)
        assert abs(F[i] - Fo[0]) <= 1e-10 * abs(Fo[0])
        floor = 1e-12 * abs(Fo[0]) / configs.THETA21_SCALE
        err = np.abs(res["theta"][i] - go)
        for k, name in enumerate(configs.THETA21_NAMES):
            print("inst %d %-10s kernel %+.12e oracle %+.12e rel %.1e"
                  % (i, name, res["theta"][i, k], go[k],
                     err[k] / max(abs(go[k]), 1e-300)))
        assert np.all(err <= 1e-8 * np.abs(go) + floor)
        lerr = np.max(np.abs(res["lambda"][i] - lam)) / np.max(np.abs(lam))
        print("inst %d lambda rel err %.1e" % (i, lerr))
        assert lerr <= 1e-10


def test_wide_gradients_match_oracle_n171():
    """20 steps, checkpoints every 7 (segments 7, 7, 6)."""
    wl = _problem(True, 4e-10, 1e-10, 5, 7)
    F, res, state = _run(wl)
    assert state["prop"].n == 171
    _check_vs_oracle(wl, F, res, state, 3)


def test_wide_gradients_match_oracle_n441():
    """10 steps, checkpoints every 4 (segments 4, 4, 2), one sweep
    point."""
    wl = _problem(False, 2e-10, 1e-10, 5, 4, freqs=[2.8700e9])
    F, res, state = _run(wl)
    assert state["prop"].n == 441
    _check_vs_oracle(wl, F, res, state, 1)


@pytest.fixture(scope="module")
def longer(gpu):
    """n = 171, 2000 steps (40 ns), checkpoints every 700."""
    wl = _problem(True, 4e-8, 2e-9, 100, 700)
    return wl, _run(wl)


def test_wide_lambda_is_dF_drho0(longer):
    wl, (F, res, _) = longer
    lin = res["lambda"] @ wl.problem.rho0
    assert np.max(np.abs(lin - F) / np.abs(F)) < 1e-10


def test_wide_checkpoint_interval_does_not_change_gradients(longer):
    wl, (F, res, _) = longer
    other = _problem(True, 4e-8, 2e-9, 100, 2000)
    F2, res2, _ = _run(other)
    assert np.array_equal(F, F2)
    assert np.array_equal(res["theta"], res2["theta"])
    assert np.array_equal(res["lambda"], res2["lambda"])


#: every parameter of the 21-level gradient
FD_PARAMS = configs.THETA21_NAMES


def test_wide_theta_gradient_central_fd(longer):
    wl, (F, res, _) = longer
    pb = wl.problem
    theta = wl.theta
    steps = pb._fd_steps(theta)

    def fwd(t):
        return pb.readout(t, wl.amplitudes).cpu().numpy()[:, 0]

    for name in FD_PARAMS:
        k = configs.THETA21_NAMES.index(name)
        tp = theta.copy()
        tm = theta.copy()
        tp[k] += steps[k]
        tm[k] -= steps[k]
        fd = (fwd(tp) - fwd(tm)) / (tp[k] - tm[k])
        adj = res["theta"][:, k]
        # truncation (relative 1e-6 steps) plus rounding of F over 2000
        # steps, divided by the step
        bound = 1e-6 * np.abs(fd) + 1e-11 * np.abs(F) / (2 * steps[k])
        print("%-10s adjoint %+.10e  fd %+.10e  err/bound %.2f"
              % (name, adj[0], fd[0], np.max(np.abs(adj - fd) / bound)))
        assert np.all(np.abs(adj - fd) <= bound), name
        if name == "T2star":
            # no generator depends on it: both sides are exactly zero
            assert np.all(adj == 0.0) and np.all(fd == 0.0)
