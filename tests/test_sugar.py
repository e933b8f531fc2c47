"""API plumbing of simspins / mesolve / generate_simulator (no device)."""



import numpy as np
import pytest

from superspinsim_amd.generate_generators import (
    _generate_dissipator, _generate_von_neumann, _generate_valid_indices,
    generate_atoms)
from superspinsim_amd.generate_simulator import (
    UnsupportedSchemeError, generate_simulator)
from superspinsim_amd.sugar import (
    _make_superoperators, _sort_operators, prepare_simspins)


class FakeQobj:
    def __init__(self, m):
        self._m = np.asarray(m, dtype=complex)

    def full(self):
        return self._m


SZ = np.diag([0.5, -0.5]).astype(complex)
SX = np.array([[0, 0.5], [0.5, 0]], dtype=complex)


def test_sort_operators_forms():
    f = lambda t: 1.0  # noqa: E731
    assert _sort_operators(SZ, "H")[0][0] is SZ
    assert np.array_equal(_sort_operators(FakeQobj(SZ), "H")[0][0], SZ)
    q, d, c = _sort_operators([SX, f], "H")
    assert not q and d[0] is SX and c[0] is f
    q, d, c = _sort_operators([SZ, [SX, f], FakeQobj(SX)], "H")
    assert len(q) == 2 and len(d) == 1
    with pytest.raises(TypeError):
        _sort_operators("nope", "H")
    assert _sort_operators(None, "c") == [[], [], []]


def test_mesolve_superoperators_match_direct_construction():
    vmap = _generate_valid_indices(np.ones((2, 2)))
    jump = np.array([[0, 1], [0, 0]], dtype=complex)
    gens, vi, rot, used = _make_superoperators([SZ], [SX], [jump], [],
                                               None, False)
    assert not used and np.array_equal(vi, vmap)
    L0 = _generate_von_neumann(SZ, vmap) + _generate_dissipator(jump, vmap)
    assert np.allclose(gens[0], L0)
    assert np.allclose(gens[1], _generate_von_neumann(SX, vmap))


def test_prepare_simspins_generator_rows():
    coeffs = [lambda t: 0.0] * 4
    spins = [[{"S": 1 / 2, "g": 2.0, "B0": np.array([0, 0, 1e-3])}]]
    sampler, gens, vmap, kw = prepare_simspins(
        coeffs, spins, [{}], {}, use_rotating=False,
        number_of_exponentials=2)
    assert gens.shape == (5, 4, 4)
    row = np.zeros(5)
    sampler(0.0, row)
    assert row.tolist() == [1.0, 0.0, 0.0, 0.0, 0.0]
    sampler, gens, vmap, kw = prepare_simspins(coeffs, spins, [{}], {})
    assert kw["use_rotating"] and gens.shape == (4, 4, 4)
    assert "vectors_real" in kw


def test_zero_l0_disables_rotating_frame():
    coeffs = [lambda t: 0.0] * 4
    spins = [[{"S": 1 / 2, "g": 2.0}]]
    _, gens, _, kw = prepare_simspins(coeffs, spins, [{}], {})
    assert not kw["use_rotating"] and gens.shape[0] == 5


def test_factory_errors_mirror_reference():
    gens = np.zeros((5, 4, 4))
    vmap = _generate_valid_indices(np.ones((2, 2)))
    s = lambda t, r: None  # noqa: E731
    with pytest.raises(NameError):
        generate_simulator(s, gens, vmap, number_of_exponentials=4)
    with pytest.raises(UnsupportedSchemeError):
        generate_simulator(s, gens, vmap, number_of_exponentials=7)
    with pytest.raises(TypeError):
        generate_simulator(s, gens, vmap, use_cayley=True)
    with pytest.raises(NotImplementedError):
        generate_simulator(s, gens, vmap, datatype=np.float32)


def test_nv7_generators_are_reference_sized():
    from superspinsim_amd import models
    ops, vmap = generate_atoms(*models.nv_7(np.array([0, 0, 1e-3])))
    assert vmap.shape == (19, 3)
    assert set(ops["generators"]) == {"L0", "Gx", "Gy", "Gz", "Gr"}
    assert np.abs(ops["generators"]["Gr"]).max() > 0



def test_mesolve_sampler_slots_follow_the_reference():
    """Rotating frame: slot k is coefficient k.  Outside it the reference
    writes the same slots, so its quiescent generator (slot 0) takes the
    first coefficient; intended_coefficients=True shifts them by one."""
    from superspinsim_amd.sugar import mesolve_coefficients
    f = lambda t: 2.0 + t  # noqa: E731
    g = lambda t: 9.0  # noqa: E731
    for rotating, intended, want in (
            (True, False, [2.5, 3.0, 0.0]), (True, True, [2.5, 3.0, 0.0]),
            (False, False, [2.5, 3.0, 0.0]), (False, True, [1.0, 2.5, 3.0])):
        row = np.zeros(3)
        mesolve_coefficients([f], [g], rotating, intended)(0.5, row)
        assert row.tolist() == want, (rotating, intended)
