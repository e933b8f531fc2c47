"""Coupling kinds beyond the NV model, pinned by physics (CPU).

* spin-spin J between two spins of one block (``generate_generators.py``
  ``_add_spin_spin_coupling`` :694-751): without dissipation L0 is
  -i[H, .] with H = Zeeman + J_perp (S1x S2x + S1y S2y) + J S1z S2z, so its
  spectrum is -i(E_a - E_b) of that H built with numpy Kronecker products;
* coherent coupling between blocks (``"coh"``, ``_couple_coherent_blocks``
  :1961-1993): amp (|b><a| + |a><b|) joins the two blocks coherently, in
  L0 (``Vc``) or, with ``"con"``, in the optically controlled generator Gr
  (``Vr``); two singlets then Rabi-oscillate, P_a(t) = cos^2(amp t);
* Bell-state ISC of a spin pair and a singlet (hBN C-like defect,
  ``_couple_bell_excited``/``_couple_bell_ground`` :1796-1960): the
  population block of L0 is the classical rate matrix singlet -> aligned
  (phi) and anti-aligned (psi) product states and back.

Each check also runs against a plausible mistake (J and J_perp swapped, a
factor 2 in the coupling, psi and phi rates swapped) that must fail it.
"""

import math

import numpy as np
import pytest
from scipy.linalg import expm

from superspinsim_amd import params
from superspinsim_amd.generate_generators import generate_atoms

TAU = math.tau
MU_B = params.fundamental.bohr_gyro
B0 = np.array([0.3e-3, 0.0, 1e-3])
J_ISO, J_PERP = TAU * 5e6, TAU * 2e6


def spin_half():
    sx = np.array([[0, 0.5], [0.5, 0]], dtype=complex)
    sy = np.array([[0, -0.5j], [0.5j, 0]])
    sz = np.diag([0.5, -0.5]).astype(complex)
    return sx, sy, sz


def j_model():
    spins = [[{"S": 0.5, "g": 2.0, "B0": B0},
              {"S": 0.5, "g": 2.1, "B0": B0}]]
    return spins, [{(0, 1): {"J": J_ISO, "J_perp": J_PERP}}], {}


def j_hamiltonian(swap=False):
    sx, sy, sz = spin_half()
    eye = np.eye(2)
    j_par, j_perp = (J_PERP, J_ISO) if swap else (J_ISO, J_PERP)
    h = np.zeros((4, 4), dtype=complex)
    for g, side in ((2.0, 0), (2.1, 1)):
        for b, s in zip(B0, (sx, sy, sz)):
            op = np.kron(s, eye) if side == 0 else np.kron(eye, s)
            h += g * TAU * MU_B * b * op
    h += j_perp * (np.kron(sx, sx) + np.kron(sy, sy)) + j_par * np.kron(sz, sz)
    return h


def _spectrum_im(matrix):
    ev = np.linalg.eigvals(matrix)
    return ev.real, np.sort(ev.imag)


def _differences(h):
    e = np.linalg.eigvalsh(h)
    return np.sort((-(e[:, None] - e[None, :])).ravel())


@pytest.mark.parametrize("swap", [False, True])
def test_spin_spin_coupling_spectrum(swap):
    ops, vmap = generate_atoms(*j_model())
    assert vmap.shape[0] == 16
    re, im = _spectrum_im(ops["generators"]["L0"])
    want = _differences(j_hamiltonian(swap))
    assert np.max(np.abs(re)) <= 1e-6 * np.max(np.abs(im))
    err = np.max(np.abs(im - want)) / np.max(np.abs(want))
    if swap:
        assert err > 1e-6
    else:
        assert err <= 1e-12


COH_AMP = TAU * 1e6


def coh_model(con):
    return ([[{"S": 0}], [{"S": 0}]], [{}, {}],
            {((0, 0), (1, 0)): {"coh": [{"ini": 0, "fin": 0, "amp": COH_AMP,
                                         "con": con}]}})


@pytest.mark.parametrize("con", [False, True])
def test_coherent_block_coupling_rabi(con):
    ops, vmap = generate_atoms(*coh_model(con))
    gens = ops["generators"]
    # the coupling joins the blocks: the (0, 1) coherence is vectorised
    assert vmap.tolist() == [[0, 0, 0], [1, 1, 0], [0, 1, 0], [0, 1, 1]]
    live, idle = ("Gr", "L0") if con else ("L0", "Gr")
    assert np.max(np.abs(gens[idle])) == 0.0
    re, im = _spectrum_im(gens[live])
    assert np.max(np.abs(re)) <= 1e-9 * COH_AMP
    assert np.allclose(im, [-2 * COH_AMP, 0, 0, 2 * COH_AMP], rtol=1e-12,
                       atol=1e-6)
    rho0 = np.array([1.0, 0.0, 0.0, 0.0])
    for t in (0.1e-6, 0.37e-6, 1.3e-6):
        p0 = (expm(gens[live] * t) @ rho0)[0]
        assert math.isclose(p0, math.cos(COH_AMP * t) ** 2, abs_tol=1e-12)
        # a plausible mistake: a factor 2 in the coupling strength
        assert abs(p0 - math.cos(2 * COH_AMP * t) ** 2) > 1e-3


RATES = {"psi_gets_e": 3e6, "phi_gets_e": 1e6, "g_gets_psi": 2e6,
         "g_gets_phi": 0.5e6}


def bell_model():
    b = np.array([0.0, 0.0, 1e-3])
    spins = [[{"S": 0.5, "g": 2.0, "B0": b}, {"S": 0.5, "g": 2.0, "B0": b}],
             [{"S": 0}]]
    groups = {((1, 0), (0, 0)): {"psi_gets_e": RATES["psi_gets_e"],
                                 "phi_gets_e": RATES["phi_gets_e"]},
              ((0, 0), (1, 0)): {"g_gets_psi": RATES["g_gets_psi"],
                                 "g_gets_phi": RATES["g_gets_phi"]}}
    return spins, [{}, {}], groups


def bell_rate_matrix(swap=False):
    """Levels: pair (m1, m2) = ++, +-, -+, -- (0-3), singlet (4)."""
    psi_e, phi_e = RATES["psi_gets_e"], RATES["phi_gets_e"]
    g_psi, g_phi = RATES["g_gets_psi"], RATES["g_gets_phi"]
    if swap:
        psi_e, phi_e = phi_e, psi_e
    rates = {(4, 1): psi_e, (4, 2): psi_e, (4, 0): phi_e, (4, 3): phi_e,
             (1, 4): g_psi, (2, 4): g_psi, (0, 4): g_phi, (3, 4): g_phi}
    w = np.zeros((5, 5))
    for (i, f), r in rates.items():
        w[f, i] += r
        w[i, i] -= r
    return w


@pytest.mark.parametrize("swap", [False, True])
def test_bell_isc_population_block(swap):
    ops, vmap = generate_atoms(*bell_model())
    gens = ops["generators"]
    where = {(int(y), int(x), int(c)): k for k, (y, x, c) in enumerate(vmap)}
    pop = [where[(i, i, 0)] for i in range(5)]
    # level labels: the (++, --) coherence turns at 2 gamma B, (+-, -+) at 0
    gamma_b = 2.0 * TAU * MU_B * 1e-3
    l0 = gens["L0"]
    rot = l0[where[(0, 3, 1)], where[(0, 3, 0)]]
    assert math.isclose(abs(rot), 2 * gamma_b, rel_tol=1e-12)
    assert l0[where[(1, 2, 1)], where[(1, 2, 0)]] == 0.0
    got = l0[np.ix_(pop, pop)]
    ok = np.allclose(got, bell_rate_matrix(swap), rtol=1e-12, atol=1e-9)
    assert ok != swap
