"""Oracle for gradients of the integrated readout (test infrastructure).

Dense restatement of the discrete adjoint of the commutator-free Magnus
propagation (generate_simulator.py:1954-2015 for the step; the reference has
no gradients):

    x_{k+1} = E_{k,0} ... E_{k,J-1} x_k,  E_kj = expm(A_kj),
    A_kj = sum_m f[k, j, m] G_m,
    F = sum_k w_k obs . x_k  (trapezoid: w_0 = w_T = h/2, else h).

The exponentials are scipy ``expm``; the exponent gradient of each
exponential with input y and output cotangent mu is the adjoint Frechet
derivative ``Abar = L(A^T, mu y^T)`` from scipy ``expm_frechet``
(Al-Mohy and Higham 2009), using <mu, L(A, E) y> = <L(A^T, mu y^T), E>.
This is independent of the device's Taylor/Beta-coefficient evaluation.

Pinned by tests/test_oracle_adjoint.py against central differences of
``readout`` on small random systems and by lambda . rho0 = F.
"""

import numpy as np
from scipy.linalg import expm, expm_frechet


def readout(mats, factors, rho0, obs, h):
    """F (n_obs,) for factors (T, J, M), mats (M, n, n)."""
    x = np.asarray(rho0, dtype=np.float64).copy()
    T, J, _ = factors.shape
    tot = 0.5 * h * (obs @ x)
    for k in range(T):
        for j in range(J - 1, -1, -1):
            x = expm(np.tensordot(factors[k, j], mats, axes=(0, 0))) @ x
        tot = tot + (h if k < T - 1 else 0.5 * h) * (obs @ x)
    return tot


def gradient(mats, factors, rho0, obs, seed, h, dmats=None, theta_gen=0,
             amp_weights=None, grad_gen=1):
    """Reverse-mode gradient of seed . F.

    ``dmats`` (n_theta, n, n): derivatives of matrix ``theta_gen``;
    ``amp_weights`` (T, J, R): d f[k, j, grad_gen] / d amp_r per segment
    record r.  Returns (F, grad_theta (n_theta,), grad_records (R,),
    lambda0 (n,)).
    """
    mats = np.asarray(mats, dtype=np.float64)
    T, J, _ = factors.shape
    x = np.asarray(rho0, dtype=np.float64).copy()
    inputs = []
    F = 0.5 * h * (obs @ x)
    for k in range(T):
        ys = [None] * J
        for j in range(J - 1, -1, -1):
            ys[j] = x.copy()
            x = expm(np.tensordot(factors[k, j], mats, axes=(0, 0))) @ x
        inputs.append(ys)
        F = F + (h if k < T - 1 else 0.5 * h) * (obs @ x)
    s = np.asarray(seed, dtype=np.float64) @ obs
    lam = 0.5 * h * s
    n_theta = 0 if dmats is None else dmats.shape[0]
    g_theta = np.zeros(n_theta)
    R = 0 if amp_weights is None else amp_weights.shape[2]
    g_rec = np.zeros(R)
    for k in range(T - 1, -1, -1):
        for j in range(J):
            A = np.tensordot(factors[k, j], mats, axes=(0, 0))
            y = inputs[k][j]
            E, abar = expm_frechet(A.T, np.outer(lam, y))
            lam = E @ lam                 # expm(A^T) = expm(A)^T
            for t in range(n_theta):
                g_theta[t] += factors[k, j, theta_gen] * np.sum(abar *
                                                               dmats[t])
            if R:
                inner = np.sum(abar * mats[grad_gen])
                g_rec += amp_weights[k, j] * inner
        lam = lam + (h if k > 0 else 0.5 * h) * s
    return F, g_theta, g_rec, lam
