"""Adjoint gradients of integrated fluorescence (GPU).

The HIP adjoint (``ssmx_adjoint``) is checked against finite differences
of the HIP forward, which itself matches the oracle to 1e-10
(test_gpu_sweep.py), on a short 7-level NV sweep (0.1 us, 5000 steps of
20 ps) with a piecewise-constant microwave amplitude, driven near the
ms = 0 <-> -1 line (2.73 GHz at Bz = 5 mT) so that the readout resolves the
parameters:

* the gradient oracle (oracle/adjoint.py: scipy expm and expm_frechet,
  pinned by tests/test_oracle_adjoint.py) on the same instances, 1000
  steps: per-instance d F / d theta for all 9 parameters and d F / d amp
  for every segment record, rtol 1e-8;
* central differences of the HIP forward with step 1e-6 relative, within
  their own error bound (truncation plus the readout's rounding, about
  1e-12 relative over 5000 steps, divided by the step);
* d F / d rho0 (the adjoint state at step 0): F is linear in rho0, so
  lambda . rho0 must equal F;
* independence of the checkpoint interval (the recomputation reproduces the
  forward states exactly);
* the torch.autograd.Function on a sum and a squared-error loss.
"""

import numpy as np
import pytest

from superspinsim_amd import configs
from superspinsim_amd.sweep import WAVE_DTYPE

pytestmark = pytest.mark.gpu

FREQS = [2.728e9, 2.730e9, 2.733e9]
COUNT = len(FREQS)
SEGMENTS = 8
T_END = 0.1e-6          # 5000 propagator steps at h = 20 ps


def _workload(checkpoint_every=700):
    return configs.odmr7_gradient(count=COUNT, segments=SEGMENTS,
                                  time_end=T_END,
                                  checkpoint_every=checkpoint_every,
                                  frequencies=FREQS)


@pytest.fixture(scope="module")
def workload(gpu):
    return _workload()


@pytest.fixture(scope="module")
def grads(workload):
    pb = workload.problem
    out, state = pb.forward(workload.theta, workload.amplitudes)
    res = pb.backward(state, np.ones((COUNT, 1)), lambda_out=True)
    print("Taylor degree %d, adjoint degree %d, slots %d / %d"
          % (state["taylor"], state["taylor_adj"], state["prop"].ell.slots,
             state["slots_t"]))
    return out.cpu().numpy()[:, 0], {k: v.cpu().numpy() for k, v in
                                     res.items()}, state


def _central(fun, x, k, step):
    xp = x.copy()
    xm = x.copy()
    xp[k] += step
    xm[k] -= step
    return (fun(xp) - fun(xm)) / (xp[k] - xm[k])


def _richardson(fun, x, k, step):
    return (4 * _central(fun, x, k, step / 2) - _central(fun, x, k, step)) / 3


def test_forward_matches_batch_readout(workload, grads):
    """The checkpointing forward gives the plain integral readout."""
    F, _, _ = grads
    plain = workload.problem.readout(workload.theta,
                                     workload.amplitudes).cpu().numpy()
    assert np.array_equal(F, plain[:, 0])
    assert np.all(F > 0)


def test_lambda_is_dF_drho0(workload, grads):
    F, res, _ = grads
    lin = res["lambda"] @ workload.problem.rho0
    assert np.max(np.abs(lin - F) / np.abs(F)) < 1e-10


def _oracle_inputs(problem, state, inst):
    """Per-step factors (T, J, M) and record weights (T, J, R) of one
    instance, from its kernel waveform at the reference node times."""
    from superspinsim_amd.generate_simulator import tabulate
    prop = state["prop"]
    kw = state["kernel_waves"][inst]
    M = prop.K
    h = problem.dt / problem.F
    T = problem.n_out * problem.F
    rows = tabulate(lambda t, row: row.__setitem__(slice(None), kw(t)), M,
                    problem.n_out, problem.F, prop.nodes, problem.t0,
                    problem.dt)                       # (N, F, S, M)
    rows = rows.reshape(T, prop.S, M)
    factors = h * np.einsum("js,tsm->tjm", prop.weights, rows)
    gg = state["grad_gen"]
    order = sorted(range(len(kw.segments[gg])),
                   key=lambda k: kw.segments[gg][k])
    segs = [kw.segments[gg][k] for k in order]
    times = tabulate(lambda t, row: row.__setitem__(0, t), 1, problem.n_out,
                     problem.F, prop.nodes, problem.t0,
                     problem.dt).reshape(T, prop.S)
    amp = np.zeros((T, prop.J, len(segs)))
    for r, (ts, te, _, om, ph) in enumerate(segs):
        inside = (times >= ts) & (times < te)
        dc = np.where(inside, np.cos(om * times + ph), 0.0)   # (T, S)
        amp[:, :, r] = h * dc @ prop.weights.T
    return factors, amp


def test_gradients_match_oracle():
    """Kernel vs oracle/adjoint.py per instance (1000 steps)."""
    from oracle import adjoint as oadj
    wl = configs.odmr7_gradient(count=COUNT, segments=4, time_end=0.02e-6,
                                checkpoint_every=300, frequencies=FREQS)
    pb = wl.problem
    out, state = pb.forward(wl.theta, wl.amplitudes)
    res = pb.backward(state, np.ones((COUNT, 1)))
    g_theta = res["theta"].cpu().numpy()
    g_rec = res["records"].cpu().numpy()
    F = out.cpu().numpy()[:, 0]
    prop = state["prop"]
    waves = np.frombuffer(state["d_waves"].cpu().numpy().tobytes(),
                          dtype=WAVE_DTYPE)
    gg = state["grad_gen"]
    h = pb.dt / pb.F
    for i in range(COUNT):
        factors, amp = _oracle_inputs(pb, state, i)
        Fo, go, gro, _ = oadj.gradient(
            prop._dense, factors, pb.rho0, pb.obs, [1.0], h,
            dmats=state["dconst"], theta_gen=0, amp_weights=amp,
            grad_gen=gg)
        assert abs(F[i] - Fo[0]) <= 1e-10 * abs(Fo[0])
        floor = 1e-12 * abs(Fo[0]) / configs.THETA_SCALE
        err = np.abs(g_theta[i] - go)
        for k, name in enumerate(configs.THETA_NAMES):
            print("inst %d %-10s kernel %+.12e oracle %+.12e rel %.1e"
                  % (i, name, g_theta[i, k], go[k],
                     err[k] / max(abs(go[k]), 1e-300)))
        assert np.all(err <= 1e-8 * np.abs(go) + floor)
        b = waves["seg_begin"][i, gg]
        c = waves["seg_count"][i, gg]
        kr = g_rec[b:b + c]
        print("inst %d records kernel %s oracle %s" % (i, kr, gro))
        assert np.all(np.abs(kr - gro) <= 1e-8 * np.max(np.abs(gro)))


#: every parameter (Bx and By are 0 at theta0, where F is even in them)
SPEC_PARAMS = configs.THETA_NAMES


def test_theta_gradient_central_fd_spec_step(workload, grads):
    """Central FD with h = 1e-6 relative, within its error bound."""
    F, res, _ = grads
    pb = workload.problem
    theta = workload.theta
    steps = pb._fd_steps(theta)

    def fwd(t):
        return pb.readout(t, workload.amplitudes).cpu().numpy()[:, 0]

    for name in SPEC_PARAMS:
        k = configs.THETA_NAMES.index(name)
        fd = _central(fwd, theta, k, steps[k])
        adj = res["theta"][:, k]
        # truncation (h^2 terms, 1e-6) plus rounding of F over 5000 steps
        bound = 1e-6 * np.abs(fd) + 1e-11 * np.abs(F) / (2 * steps[k])
        print("%-10s adjoint %+.10e  fd %+.10e  err/bound %.2f"
              % (name, adj[0], fd[0], np.max(np.abs(adj - fd) / bound)))
        assert np.all(np.abs(adj - fd) <= bound), name


def test_amplitude_gradient_central_fd_spec_step(workload, grads):
    F, res, _ = grads
    pb = workload.problem
    amps = workload.amplitudes

    def fwd(a):
        return pb.readout(workload.theta, a).cpu().numpy()[:, 0].sum()

    adj = res["amplitudes"]
    for k in range(SEGMENTS):
        step = 1e-6 * abs(amps[k])
        fd = _central(fwd, amps, k, step)
        bound = 1e-6 * abs(fd) + 1e-11 * np.abs(F).sum() / (2 * step)
        print("a[%d] adjoint %+.10e  fd %+.10e  err/bound %.2f"
              % (k, adj[k], fd, abs(adj[k] - fd) / bound))
        assert abs(adj[k] - fd) <= bound, k


def test_checkpoint_interval_does_not_change_gradients(grads):
    _, res, _ = grads
    other = _workload(checkpoint_every=5000)
    out, state = other.problem.forward(other.theta, other.amplitudes)
    r2 = other.problem.backward(state, np.ones((COUNT, 1)))
    assert np.array_equal(res["theta"], r2["theta"].cpu().numpy())
    assert np.array_equal(res["amplitudes"], r2["amplitudes"].cpu().numpy())


def test_autograd_function(workload, grads):
    import torch
    from superspinsim_amd.adjoint import IntegratedReadout
    F, res, state = grads
    pb = workload.problem

    def leaves():
        return (torch.tensor(workload.theta, dtype=torch.float64,
                             requires_grad=True),
                torch.tensor(workload.amplitudes, dtype=torch.float64,
                             requires_grad=True))

    # linear loss with power-of-two weights: the seeds scale the seed-1
    # run exactly (only the order of the instance sum may differ)
    c = np.array([0.5, 2.0, 0.25])
    theta, amps = leaves()
    out = IntegratedReadout.apply(theta, amps, pb)
    (out[:, 0] * torch.from_numpy(c).cuda()).sum().backward()
    want = (c[:, None] * res["theta"]).sum(axis=0)
    size = (c[:, None] * np.abs(res["theta"])).sum(axis=0)
    assert np.all(np.abs(theta.grad.numpy() - want) <= 1e-14 * size)
    # amplitudes are shared by the instances: seed-weighted record sums
    waves = np.frombuffer(state["d_waves"].cpu().numpy().tobytes(),
                          dtype=WAVE_DTYPE)
    gg = state["grad_gen"]
    want_a = np.zeros(SEGMENTS)
    size_a = np.zeros(SEGMENTS)
    for i in range(COUNT):
        b = waves["seg_begin"][i, gg]
        for r in range(b, b + waves["seg_count"][i, gg]):
            k = state["tag_index"][r]
            v = c[i] * state["tag_scale"][r] * res["records"][r]
            want_a[k] += v
            size_a[k] += abs(v)
    assert np.all(np.abs(amps.grad.numpy() - want_a) <= 1e-14 * size_a)

    # squared error: seeds 2 (F - target); components resolved to 1e-8
    # (those not cancelling inside an instance) agree to 1e-6
    theta, amps = leaves()
    target = torch.from_numpy(0.98 * F).cuda()
    out = IntegratedReadout.apply(theta, amps, pb)
    ((out[:, 0] - target) ** 2).sum().backward()
    w = 2 * (F - 0.98 * F)
    want = (w[:, None] * res["theta"]).sum(axis=0)
    size = (np.abs(w[:, None]) * np.abs(res["theta"])).sum(axis=0)
    got = theta.grad.numpy()
    assert np.all(np.abs(got - want) <= 1e-6 * size)


@pytest.mark.parametrize("adjoint_tolerance", [1e-13, 1e-12])
def test_stored_inputs_vs_recomputation(gpu, adjoint_tolerance):
    """The backward from stored exponential inputs (no recomputation)
    against the checkpointed one (adjoint degree 10): bit for bit at the
    same degree; at the default tolerance (degree 9) within 1e-8 of each
    parameter's largest gradient."""
    res = {}
    for store in (False, True):
        wl = configs.odmr7_gradient(
            count=COUNT, segments=SEGMENTS, time_end=T_END,
            checkpoint_every=700, frequencies=FREQS,
            adjoint_tolerance=adjoint_tolerance, store_inputs=store)
        pb = wl.problem
        out, state = pb.forward(wl.theta, wl.amplitudes)
        assert state["stored"] == store
        r = pb.backward(state, np.ones((COUNT, 1)), lambda_out=True)
        res[store] = ({k: v.cpu().numpy() for k, v in r.items()},
                      out.cpu().numpy(), int(state["taylor_adj"]))
        print(f"store={store}: adjoint degree {state['taylor_adj']}",
              flush=True)
    (a, fa, ma), (b, fb, mb) = res[False], res[True]
    assert np.array_equal(fa, fb)
    if mb >= 10:
        for k in ("theta", "records", "amplitudes", "lambda"):
            assert np.array_equal(a[k], b[k]), k
    else:
        # relative to each gradient, with a floor on the parameter's natural
        # scale F / theta_scale for components that cancel (dF/dBy is zero
        # by symmetry with the field along z and the drive along x)
        F = fa[:, 0]
        err = np.abs(a["theta"] - b["theta"])
        floor = 1e-9 * np.abs(F)[:, None] / configs.THETA_SCALE[None, :]
        for k, name in enumerate(configs.THETA_NAMES):
            print(f"  degree {mb}: {name:10s} |grad| {np.abs(a['theta'][:, k]).max():.3e}"
                  f" err {err[:, k].max():.2e} floor {floor[:, k].max():.2e}")
        assert np.all(err <= 1e-8 * np.abs(a["theta"]) + floor)
        size = np.abs(a["amplitudes"]).max()
        assert np.max(np.abs(a["amplitudes"] - b["amplitudes"])) <= \
            1e-8 * size
        assert np.max(np.abs(a["lambda"] - b["lambda"])) <= \
            1e-12 * np.max(np.abs(a["lambda"]))


@pytest.mark.parametrize("store", [False, True])
def test_handoff_overrun_poisons_every_gradient(gpu, monkeypatch, store):
    """A handoff wait that gives up (bound forced to 0 polls from the host;
    every kernel still exits normally) must turn every output of the
    affected instances into NaN, amplitude records included."""
    wl = configs.odmr7_gradient(count=COUNT, segments=SEGMENTS,
                                time_end=0.02e-6, checkpoint_every=300,
                                frequencies=FREQS, store_inputs=store)
    pb = wl.problem
    out, state = pb.forward(wl.theta, wl.amplitudes)
    monkeypatch.setenv("SSMX_ADJOINT_MAX_POLLS", "0")
    r = pb.backward(state, np.ones((COUNT, 1)), lambda_out=True)
    theta = r["theta"].cpu().numpy()
    lam = r["lambda"].cpu().numpy()
    rec = r["records"].cpu().numpy()
    monkeypatch.delenv("SSMX_ADJOINT_MAX_POLLS")
    assert np.all(np.isnan(theta)) and np.all(np.isnan(lam))
    touched = np.asarray(state["tag_index"]) >= 0
    assert np.all(np.isnan(rec[touched]))
    # the bound restored: finite again
    r = pb.backward(state, np.ones((COUNT, 1)))
    assert np.all(np.isfinite(r["theta"].cpu().numpy()))


def test_backward_refuses_forward_only_state(workload):
    """A state set up without derivatives (as readout builds it) has no
    adjoint tables: backward says so instead of launching."""
    pb = workload.problem
    state = pb.setup(workload.theta, workload.amplitudes, derivatives=False)
    assert state["d_col"] is None and not state["derivatives"]
    with pytest.raises(ValueError, match="without derivatives"):
        pb.backward(state, np.ones((COUNT, 1)))
