"""``mesolve`` on the GPU (sugar.py:21-161).

A driven, damped qubit: H = H0 + f(t) H1 and c_ops = [L0, [L1, g]], whose
time-dependent jump operator enters through sqrt(g(t)) on its dissipator
(sugar.py:123-137).

* Against the oracle: the reference algorithm on the generators and the
  sampler ``mesolve`` builds, 1e-10 relative to the largest entry.
* Against the Lindblad equation: rho' = -i[H0 + f H1, rho] + D[L0] rho +
  sqrt(g) D[L1] rho, written here with numpy Kronecker products on the
  complex N^2 vectorisation and integrated by scipy DOP853 (rtol 1e-13),
  the stand-in for QuTiP's mesolve; bar 1e-8.  The non-rotating run uses
  ``intended_coefficients=True``; with the reference's slot wiring (the
  default) it is compared with the oracle only, and differs from the
  Lindblad equation by design.
"""

import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd.generate_simulator import generate_simulator
from superspinsim_amd.sugar import (
    _make_superoperators, _sort_operators, mesolve, mesolve_coefficients)

pytestmark = pytest.mark.gpu

W0 = math.tau * 10e6
OMEGA = math.tau * 1e6
SZ = np.diag([0.5, -0.5]).astype(complex)
SX = np.array([[0, 0.5], [0.5, 0]], dtype=complex)
SM = np.array([[0, 0], [1, 0]], dtype=complex)
H0 = W0 * SZ
H1 = OMEGA * 2 * SX
L0 = math.sqrt(2e5) * SM
L1 = math.sqrt(1e5) * 2 * SZ


def f(t):
    return math.cos(W0 * t)


def g(t):
    return (1.0 + 0.5 * math.sin(math.tau * 3e6 * t)) ** 2


RHO0 = np.array([[0.8, 0.1 - 0.2j], [0.1 + 0.2j, 0.2]])
T_END, DT = 1e-6, 10e-9


def _lindblad_ode(times):
    eye = np.eye(2)

    def comm(h):
        return -1j * (np.kron(eye, h) - np.kron(h.T, eye))

    def diss(lop):
        ld = lop.conj().T @ lop
        return (np.kron(lop.conj(), lop)
                - 0.5 * (np.kron(eye, ld) + np.kron(ld.T, eye)))

    a0 = comm(H0) + diss(L0)
    a1 = comm(H1)
    a2 = diss(L1)

    def rhs(t, y):
        v = y[:4] + 1j * y[4:]
        dv = (a0 + f(t) * a1 + math.sqrt(g(t)) * a2) @ v
        return np.concatenate([dv.real, dv.imag])

    v0 = RHO0.reshape(-1, order="F")
    sol = solve_ivp(rhs, (0.0, times[-1]), np.concatenate([v0.real, v0.imag]),
                    method="DOP853", t_eval=times, rtol=1e-13, atol=1e-15)
    assert sol.success
    v = sol.y[:4] + 1j * sol.y[4:]
    return np.stack([v[:, k].reshape(2, 2, order="F")
                     for k in range(v.shape[1])])


def _run(use_rotating, intended, fine):
    return mesolve([H0, [H1, f]], RHO0, 0.0, T_END, DT,
                   c_ops=[L0, [L1, g]], use_rotating=use_rotating,
                   number_of_fine_divisions=fine,
                   intended_coefficients=intended)


def _oracle(use_rotating, intended, fine):
    hq, hd, hc = _sort_operators([H0, [H1, f]], "H")
    lq, ld, lc = _sort_operators([L0, [L1, g]], "c_ops")
    gens, vmap, rot, used = _make_superoperators([sum(hq)], hd, lq, ld,
                                                 None, use_rotating)
    assert used == use_rotating
    sampler = mesolve_coefficients(hc, lc, used, intended)
    return oracle_gen(sampler, gens, vmap, backend="c",
                      number_of_exponentials=5,
                      number_of_fine_divisions=fine, use_rotating=used,
                      **rot)(RHO0, 0.0, T_END, DT)


@pytest.mark.parametrize("use_rotating,intended", [
    (True, False), (False, False), (False, True)])
def test_mesolve_vs_oracle(gpu, use_rotating, intended):
    print(f"mesolve rotating={use_rotating} intended={intended}", flush=True)
    t, rho = _run(use_rotating, intended, 20)
    t_ref, rho_ref = _oracle(use_rotating, intended, 20)
    assert np.array_equal(t, t_ref)
    err = np.max(np.abs(rho - rho_ref)) / np.max(np.abs(rho_ref))
    print(f"  rel err vs oracle {err:.3e}", flush=True)
    assert err <= 1e-10


@pytest.mark.parametrize("use_rotating", [True, False])
def test_mesolve_vs_lindblad_equation(gpu, use_rotating):
    print(f"mesolve rotating={use_rotating} vs DOP853", flush=True)
    t, rho = _run(use_rotating, True, 50)
    want = _lindblad_ode(t)
    err = np.max(np.abs(rho - want))
    print(f"  max |diff| vs DOP853 {err:.3e}", flush=True)
    assert err <= 1e-8
    # the dissipators act: purity falls and the state is not the free one
    assert np.real(np.trace(rho[-1] @ rho[-1])) < 0.99


def test_reference_slot_wiring_is_not_the_lindblad_equation(gpu):
    t, rho = _run(False, False, 20)
    want = _lindblad_ode(t)
    assert np.max(np.abs(rho - want)) > 1e-3
