"""Pins of the gradient oracle (oracle/adjoint.py), CPU only.

On small random Lindblad-like systems (stable, non-normal) the oracle's
reverse-mode gradient must match central differences of its own readout
(step 1e-6, rtol 1e-7) for a direction in a generator matrix, for segment
amplitude weights and for rho0; F is linear in rho0, so lambda . rho0 = F.
The oracle's exponentials and Frechet derivatives are scipy's
(expm, expm_frechet), not the device's Taylor evaluation.
"""

import numpy as np
import pytest

from oracle import adjoint as oadj


def _system(seed, n=6, M=2, T=25, J=2, R=3):
    rng = np.random.default_rng(seed)
    mats = rng.normal(size=(M, n, n)) * 0.4
    mats[0] -= 1.5 * np.eye(n)              # decaying drift
    factors = rng.uniform(0.02, 0.08, size=(T, J, M))
    dmats = rng.normal(size=(2, n, n))
    amp = rng.normal(size=(T, J, R)) * 0.03
    rho0 = rng.normal(size=n)
    obs = rng.normal(size=(1, n))
    return mats, factors, dmats, amp, rho0, obs


@pytest.mark.parametrize("seed", [0, 1])
def test_theta_gradient_matches_central_differences(seed):
    mats, factors, dmats, amp, rho0, obs = _system(seed)
    h = 0.05
    F, g, grec, lam = oadj.gradient(mats, factors, rho0, obs, [1.0], h,
                                    dmats=dmats, theta_gen=0,
                                    amp_weights=amp, grad_gen=1)
    eps = 1e-6
    for t in range(dmats.shape[0]):
        mp = mats.copy()
        mm = mats.copy()
        mp[0] += eps * dmats[t]
        mm[0] -= eps * dmats[t]
        fd = (oadj.readout(mp, factors, rho0, obs, h)[0]
              - oadj.readout(mm, factors, rho0, obs, h)[0]) / (2 * eps)
        assert abs(g[t] - fd) <= 1e-7 * abs(fd)
    for r in range(amp.shape[2]):
        fp = factors.copy()
        fm = factors.copy()
        fp[:, :, 1] += eps * amp[:, :, r]
        fm[:, :, 1] -= eps * amp[:, :, r]
        fd = (oadj.readout(mats, fp, rho0, obs, h)[0]
              - oadj.readout(mats, fm, rho0, obs, h)[0]) / (2 * eps)
        assert abs(grec[r] - fd) <= 1e-7 * abs(fd)
    assert abs(lam @ rho0 - F[0]) <= 1e-12 * abs(F[0])


def test_rho0_gradient_matches_central_differences():
    mats, factors, _, _, rho0, obs = _system(3)
    h = 0.05
    _, _, _, lam = oadj.gradient(mats, factors, rho0, obs, [2.0], h)
    eps = 1e-6
    for i in range(rho0.size):
        rp = rho0.copy()
        rm = rho0.copy()
        rp[i] += eps
        rm[i] -= eps
        fd = 2.0 * (oadj.readout(mats, factors, rp, obs, h)[0]
                    - oadj.readout(mats, factors, rm, obs, h)[0]) / (2 * eps)
        assert abs(lam[i] - fd) <= 1e-8 * max(abs(fd), 1e-3)


def test_a_dropped_transpose_is_caught():
    """A plausible mistake (Frechet derivative at A instead of A^T) fails."""
    mats, factors, dmats, _, rho0, obs = _system(0)
    h = 0.05
    _, g, _, _ = oadj.gradient(mats, factors, rho0, obs, [1.0], h,
                               dmats=dmats)
    from scipy.linalg import expm, expm_frechet
    T, J, _ = factors.shape
    x = rho0.copy()
    inputs = []
    for k in range(T):
        ys = [None] * J
        for j in range(J - 1, -1, -1):
            ys[j] = x.copy()
            x = expm(np.tensordot(factors[k, j], mats, axes=(0, 0))) @ x
        inputs.append(ys)
    lam = 0.5 * h * obs[0]
    wrong = 0.0
    for k in range(T - 1, -1, -1):
        for j in range(J):
            A = np.tensordot(factors[k, j], mats, axes=(0, 0))
            E, abar = expm_frechet(A, np.outer(lam, inputs[k][j]))
            lam = E.T @ lam
            wrong += factors[k, j, 0] * np.sum(abar * dmats[0])
        lam = lam + (h if k > 0 else 0.5 * h) * obs[0]
    assert abs(wrong - g[0]) > 1e-4 * abs(g[0])
