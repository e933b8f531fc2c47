"""Invariants of the host-side model construction (generate_generators).

The golden CSVs pin qubit/qutrit/hyperfine models end to end; these checks
cover the NV models and the basis conventions directly.
"""

import math

import numpy as np
import pytest

from superspinsim_amd import models
from superspinsim_amd.generate_generators import (
    _generate_dissipator, _generate_valid_indices, _generate_von_neumann,
    find_kernel, generate_atoms, make_rotating)


def _gens(desc, **kw):
    ops, vmap = generate_atoms(*desc, **kw)
    g = ops["generators"]
    return [g[k] for k in ("L0", "Gx", "Gy", "Gz", "Gr")], vmap


def _diag_rows(vmap):
    return np.array([k for k, (y, x, c) in enumerate(vmap) if y == x])


@pytest.mark.parametrize("incoherent,n", [(False, 49), (True, 19)])
def test_nv7_sizes_and_trace_preservation(incoherent, n):
    gens, vmap = _gens(models.nv_7(np.array([0, 0, 5e-3])),
                       use_incoherent=incoherent)
    assert vmap.shape == (n, 3)
    diag = _diag_rows(vmap)
    assert diag.tolist() == list(range(7))
    for g in gens:
        assert g.shape == (n, n)
        # the trace functional annihilates every generator
        assert np.max(np.abs(g[diag].sum(axis=0))) <= 1e-6 * max(
            1.0, np.abs(g).max()) * 1e-3


def test_nv21_sizes():
    gens, vmap = _gens(models.nv_21(np.array([0, 0, 5e-3])),
                       use_incoherent=False)
    assert gens[0].shape == (441, 441)
    gens, vmap = _gens(models.nv_21(np.array([0, 0, 5e-3])))
    assert gens[0].shape == (171, 171)


def test_valid_indices_order_and_default_mask():
    vi = _generate_valid_indices(np.ones((3, 3)))
    assert vi[:3].tolist() == [[0, 0, 0], [1, 1, 0], [2, 2, 0]]
    assert vi[3:].tolist() == [[0, 1, 0], [0, 1, 1], [0, 2, 0], [0, 2, 1],
                               [1, 2, 0], [1, 2, 1]]
    assert _generate_valid_indices(None).shape == (19, 3)


def _orthonormal_scale(vmap):
    return np.array([1.0 if y == x else math.sqrt(2) for y, x, _ in vmap])


def test_von_neumann_is_antisymmetric_in_orthonormal_basis():
    rng = np.random.default_rng(1)
    h = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    h = h + h.conj().T
    vmap = _generate_valid_indices(np.ones((3, 3)))
    L = _generate_von_neumann(h, vmap)
    d = _orthonormal_scale(vmap)
    Lo = d[:, None] * L / d[None, :]
    assert np.max(np.abs(Lo + Lo.T)) < 1e-12


def test_von_neumann_matches_commutator():
    """L x equals the coordinates of -i [H, rho] for a random rho."""
    rng = np.random.default_rng(2)
    h = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    h = h + h.conj().T
    r = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    r = r + r.conj().T
    vmap = _generate_valid_indices(np.ones((3, 3)))
    L = _generate_von_neumann(h, vmap)
    coords = lambda m: np.array([m[y, x].real if c == 0 else m[y, x].imag  # noqa
                                 for y, x, c in vmap])
    assert np.allclose(L @ coords(r), coords(-1j * (h @ r - r @ h)),
                       atol=1e-12)


def test_single_jump_dissipator():
    """L = |0><1|: population 1 -> 0 at rate 1, coherence decays at 1/2."""
    jump = np.zeros((2, 2), complex)
    jump[0, 1] = 1
    vmap = _generate_valid_indices(np.ones((2, 2)))
    D = _generate_dissipator(jump, vmap)
    # rows/cols: rho00, rho11, Re rho01, Im rho01
    expect = np.array([[0, 1, 0, 0], [0, -1, 0, 0],
                       [0, 0, -0.5, 0], [0, 0, 0, -0.5]])
    assert np.allclose(D, expect)


def test_nv7_quiescent_steady_state_is_physical():
    gens, vmap = _gens(models.nv_7(np.array([0, 0, 5e-3])))
    L = gens[0] + 0.1 * gens[4]
    w, v = np.linalg.eig(L)
    k = np.argmin(np.abs(w))
    ss = np.real(v[:, k])
    ss = ss / ss[:7].sum()
    assert abs(w[k]) < 1e-3
    assert np.all(ss[:7] > -1e-12)


def test_rotating_frame_reconstructs_generators():
    gens, vmap = _gens(models.nv_7(np.array([0, 0, 5e-3])))
    rot, V, Vi, doubles, singles = make_rotating(gens[0], gens[1:])
    assert np.allclose(V @ Vi, np.eye(V.shape[0]), atol=1e-10)
    for g, r in zip(gens[1:], rot):
        assert np.allclose(V @ r @ Vi, g, atol=1e-6 * np.abs(g).max())
    # block-diagonal form of L0 in the eigenbasis
    blk = Vi @ gens[0] @ V
    nd = doubles.shape[0]
    for y in range(nd):
        a, b = doubles[y]
        assert np.allclose(blk[2 * y:2 * y + 2, 2 * y:2 * y + 2],
                           [[a, b], [-b, a]], atol=1e-6 * abs(gens[0]).max())


def test_find_kernel_projects_out_identity():
    desc = ([[{"S": 1 / 2, "g": 2.0, "B0": np.array([0, 0, 1e-3])}]],
            [{}], {})
    gens, vmap = _gens(desc)
    full, image, proj = find_kernel(gens)
    assert image.shape == (4, 3)
    ident = np.array([1.0, 1.0, 0.0, 0.0]) / math.sqrt(2)
    assert np.allclose(image.T @ ident, 0, atol=1e-12)
    assert np.allclose(image.T @ image, np.eye(3), atol=1e-12)
