"""Brute-force references from the mathematics (test oracle only).

* ``expm_history`` - constant generator: rho(t) = expm(L t) rho0 (scipy).
* ``ode_history`` - the vectorised Lindblad ODE dx/dt = (sum_g c_g(t) G_g) x
  integrated with scipy ``solve_ivp`` (DOP853), standing in for QuTiP's
  ``mesolve`` (not installed here).
"""

import numpy as np
from scipy.integrate import solve_ivp
from scipy.linalg import expm


def expm_history(generator, x0, times):
    return np.array([expm(generator * t) @ x0 for t in times])


def ode_history(generators, coefficients, x0, t0, times, rtol=1e-13,
                atol=1e-15):
    """``coefficients(t) -> (G,)``; returns x at ``times`` (N, n)."""
    gens = np.asarray(generators)

    def rhs(t, x):
        c = coefficients(t)
        return np.tensordot(c, gens, axes=(0, 0)) @ x

    sol = solve_ivp(rhs, (t0, times[-1]), x0, method="DOP853",
                    t_eval=times, rtol=rtol, atol=atol)
    if not sol.success:
        raise RuntimeError(sol.message)
    return sol.y.T
