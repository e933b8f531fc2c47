"""Pins of the NV generator blocks, written from the physics they model.

Every configuration the benchmarks run builds its superoperators with
``generate_atoms`` from the NV descriptions (``models.nv_7`` after
``superspinsim/models.py:7-87``, ``models.nv_21``).  The GPU comparisons
feed those same matrices to both sides, so here they are checked against
quantities constructed independently of the builder:

* population block: with the static field along the NV axis the energy
  eigenstates are the |m> states, so the populations obey a classical
  rate equation dp/dt = W p.  W is written from the transition list of the
  reference's coupling functions (``generate_generators.py``: optical
  ``_couple_optical`` :1608-1711 with the rate halving for |m| = S at
  :1677-1704, ISC ``_couple_isc_excited``/``_couple_isc_ground``
  :1712-1795, thermalisation :459-558 with rate b_f / ((b_i + b_f) T1)
  and Boltzmann factors b = exp(-E / (2 pi k_B/h T))).  L0 must hold the
  decay, ISC and T1 rates, Gr the optical pumping rates, Gx/Gy/Gz none;
* coherent spectrum: without dissipation the generators are -i[H, .], so
  their eigenvalues are -i(E_a - E_b) for the eigenvalues E of H, built
  here with numpy spin matrices: D Sz^2 + g mu_B/h 2 pi B.S per triplet
  (g_perp on x and y of the ground triplet, :250-283), and for the 14N
  blocks P Iz^2 + A_par Sz Iz + A_perp (Sx Ix + Sy Iy) - g_N mu_N/h 2 pi
  B.I (nuclear quadrupole :355-372, hyperfine :559-618, nuclear Zeeman
  sign :410);
* coherence decay: with diagonal jump operators the coherence rho_yx of
  the ground triplet decays at (m_y - m_x)^2 / (2 T2) (dephasing jump
  S / (lambda_max sqrt(T2)) B0.G, :300-315) plus half the T1 rates out of
  both levels, self-jumps i -> i included (:524-555).

Each check is also run against a plausible mistake (swapped ISC rates, a
dropped halving, a T1 ratio inverted, a dropped dephasing normalisation,
a nuclear Zeeman sign flip, A_par and A_perp swapped) and must fail it.
"""

import copy
import math

import numpy as np
import pytest

from superspinsim_amd import configs, models, params
from superspinsim_amd.generate_generators import generate_atoms

TAU = math.tau
MU_B = params.fundamental.bohr_gyro
MU_N = params.fundamental.nuclear_gyro
K_B = params.fundamental.boltzmann_gyro


# --- independent spin algebra ---------------------------------------------

def spin_ops(s):
    """(Sx, Sy, Sz) for spin s in the basis m = s, s-1, ..., -s."""
    m = np.arange(s, -s - 1, -1)
    sp = np.zeros((m.size, m.size), dtype=complex)
    for k in range(1, m.size):
        mk = m[k]
        sp[k - 1, k] = math.sqrt(s * (s + 1) - mk * (mk + 1))
    sx = (sp + sp.conj().T) / 2
    sy = (sp - sp.conj().T) / 2j
    return sx, sy, np.diag(m).astype(complex)


def build(description, use_incoherent=True):
    spins, spin_int, groups = description
    ops, vmap = generate_atoms(copy.deepcopy(spins), spin_int,
                               copy.deepcopy(groups),
                               use_incoherent=use_incoherent)
    return ops["generators"], vmap


def population_index(vmap, levels):
    where = {(int(y), int(x)): k for k, (y, x, c) in enumerate(vmap)
             if c == 0}
    return np.array([where[(i, i)] for i in range(levels)])


def rate_matrix(rates, levels):
    """W with W[f, i] = rate(i -> f) and columns summing to zero."""
    w = np.zeros((levels, levels))
    for (i, f), r in rates.items():
        if i != f:
            w[f, i] += r
            w[i, i] -= r
    return w


def thermal_rates(energies, t1, temperature, levels):
    """T1 jumps between eigenstates |k> (the |m> states for B along z)."""
    b = np.exp(-np.asarray(energies) / (TAU * K_B * temperature))
    out = {}
    for i, e_i in zip(levels, b):
        for f, e_f in zip(levels, b):
            if i != f:
                out[(i, f)] = e_f / ((e_i + e_f) * t1)
    return out


def triplet_energies(d, g, bz):
    """Diagonal of D Sz^2 + g mu_B 2 pi Bz Sz for m = +1, 0, -1."""
    return np.array([d + g * TAU * MU_B * bz, 0.0, d - g * TAU * MU_B * bz])


# --- 7-level population block ---------------------------------------------

def nv7_expected(description, halve=True, swap_isc=False,
                 invert_boltzmann=False):
    """(W_L0, W_Gr) of the 7-level description: levels ground +1, 0, -1
    (0-2), excited +1, 0, -1 (3-5), singlet (6)."""
    spins, _, groups = description
    ground, excited = spins[0][0], spins[1][0]
    opt = groups[((0, 0), (1, 0))]
    isc_e = dict(groups[((1, 0), (2, 0))])
    isc_g = groups[((2, 0), (0, 0))]
    if swap_isc:
        isc_e["s_gets_0"], isc_e["s_gets_1"] = (isc_e["s_gets_1"],
                                                isc_e["s_gets_0"])
    m_of = {0: 1, 1: 0, 2: -1}
    decay, pump = {}, {}
    rel, rel_n = opt["rel"], opt.get("rel_n")
    for a in range(3):
        decay[(3 + a, a)] = rel
        pump[(a, 3 + a)] = rel
        if rel_n is None:
            continue
        for b in range(3):
            if abs(m_of[a] - m_of[b]) != 1:
                continue
            # pump ground a -> excited b: halved when the source |m| = 1
            pump[(a, 3 + b)] = rel_n / 2 if (halve and abs(m_of[a]) == 1) \
                else rel_n
            # decay excited b -> ground a: halved when the source |m| = 1
            decay[(3 + b, a)] = rel_n / 2 if (halve and abs(m_of[b]) == 1) \
                else rel_n
    for b in range(3):
        decay[(3 + b, 6)] = isc_e["s_gets_0"] if m_of[b] == 0 \
            else isc_e["s_gets_1"]
        decay[(6, b)] = isc_g["0_gets_s"] if m_of[b] == 0 \
            else isc_g["1_gets_s"]
    for atom, offset in ((ground, 0), (excited, 3)):
        e = triplet_energies(atom["D"], atom["g"], atom["B0"][2])
        if invert_boltzmann:
            e = -e
        for key, r in thermal_rates(e, atom["TS1"], atom["T"],
                                    [offset, offset + 1, offset + 2]).items():
            decay[key] = decay.get(key, 0.0) + r
    return rate_matrix(decay, 7), rate_matrix(pump, 7)


DESCRIPTIONS = {
    "odmr7": lambda: configs._nv7_description(),
    "nv_7 default params": lambda: models.nv_7(np.array([0.0, 0.0, 5e-3])),
}


def _population_blocks(description):
    gens, vmap = build(description)
    pop = population_index(vmap, 7)
    return {k: gens[k][np.ix_(pop, pop)] for k in gens}


def _same_rates(got, want):
    """Entry by entry to 1e-12 relative (rates run from 50 /s to 1e8 /s);
    entries that should be zero must be below 1e-9 /s."""
    return np.allclose(got, want, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("name", sorted(DESCRIPTIONS))
def test_nv7_population_block_is_the_rate_matrix(name):
    desc = DESCRIPTIONS[name]()
    blocks = _population_blocks(desc)
    w0, wr = nv7_expected(desc)
    assert _same_rates(blocks["L0"], w0)
    assert _same_rates(blocks["Gr"], wr)
    for g in ("Gx", "Gy", "Gz"):
        assert np.max(np.abs(blocks[g])) == 0.0


@pytest.mark.parametrize("mistake", [
    dict(swap_isc=True), dict(halve=False), dict(invert_boltzmann=True)])
def test_nv7_population_check_catches_mistakes(mistake):
    desc = DESCRIPTIONS["nv_7 default params"]()
    blocks = _population_blocks(desc)
    w0, wr = nv7_expected(desc, **mistake)
    assert not (_same_rates(blocks["L0"], w0)
                and _same_rates(blocks["Gr"], wr)), mistake


def test_thermal_rates_obey_detailed_balance_and_sum_to_inverse_t1():
    desc = DESCRIPTIONS["nv_7 default params"]()
    blocks = _population_blocks(desc)
    ground = desc[0][0][0]
    w = blocks["L0"][:3, :3]
    e = triplet_energies(ground["D"], ground["g"], ground["B0"][2])
    kt = TAU * K_B * ground["T"]
    for i in range(3):
        for f in range(3):
            if i == f:
                continue
            assert math.isclose(w[f, i] + w[i, f], 1 / ground["TS1"],
                                rel_tol=1e-12)
            assert math.isclose(w[f, i] / w[i, f],
                                math.exp(-(e[f] - e[i]) / kt), rel_tol=1e-12)


# --- coherent spectra ------------------------------------------------------

def _coherent_only(description):
    """The description with every dissipative and incoherent term removed."""
    spins, spin_int, groups = copy.deepcopy(description)
    for block in spins:
        for atom in block:
            for key in ("TS1", "TS2", "TI1", "TI2"):
                atom.pop(key, None)
    return spins, spin_int, {}


def _block_spectrum(hamiltonians):
    """Sorted -i(E_a - E_b) over a, b within each block (imaginary parts)."""
    out = []
    for h in hamiltonians:
        e = np.linalg.eigvalsh(h)
        out.extend((-(e[:, None] - e[None, :])).ravel())
    return np.sort(np.array(out))


def _generator_spectrum(matrix):
    ev = np.linalg.eigvals(matrix)
    return ev.real, np.sort(ev.imag)


def _triplet_h(atom, field, g_perp=None):
    sx, sy, sz = spin_ops(1)
    g_par = atom["g"]
    g_t = g_perp if g_perp is not None else atom.get("g_perp", g_par)
    return (atom["D"] * sz @ sz
            + TAU * MU_B * (g_t * (field[0] * sx + field[1] * sy)
                            + g_par * field[2] * sz))


FIELD = np.array([1.3e-3, -0.7e-3, 4.1e-3])


@pytest.mark.parametrize("g_perp_mistake", [False, True])
def test_nv7_coherent_spectrum(g_perp_mistake):
    desc = models.nv_7(FIELD)
    gens, vmap = build(_coherent_only(desc))
    ground, excited = desc[0][0][0], desc[0][1][0]
    g_perp = ground["g"] if g_perp_mistake else None
    want = _block_spectrum([_triplet_h(ground, FIELD, g_perp),
                            _triplet_h(excited, FIELD), np.zeros((1, 1))])
    re, im = _generator_spectrum(gens["L0"])
    assert np.max(np.abs(re)) <= 1e-6 * np.max(np.abs(im))
    err = np.max(np.abs(im - want)) / np.max(np.abs(want))
    if g_perp_mistake:
        assert err > 1e-9
    else:
        assert err <= 1e-12


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_nv7_field_generators_spectrum(axis):
    """G_axis = -i[g mu_B 2 pi S_axis, .] per triplet (per tesla)."""
    desc = models.nv_7(FIELD)
    gens, _ = build(_coherent_only(desc))
    unit = np.zeros(3)
    unit[axis] = 1.0
    ground, excited = desc[0][0][0], desc[0][1][0]
    hs = [_triplet_h(dict(ground, D=0.0), unit),
          _triplet_h(dict(excited, D=0.0), unit), np.zeros((1, 1))]
    _, im = _generator_spectrum(gens["G" + "xyz"[axis]])
    want = _block_spectrum(hs)
    assert np.max(np.abs(im - want)) <= 1e-12 * np.max(np.abs(want))


# --- 14N: quadrupole, hyperfine, nuclear Zeeman ----------------------------

def _n14_h(atom, field, zeeman_sign=-1.0, swap_a=False):
    sx, sy, sz = spin_ops(atom["S"]) if atom["S"] > 0 else \
        (np.zeros((1, 1)),) * 3
    ix, iy, iz = spin_ops(atom["I"])
    es, ei = np.eye(sx.shape[0]), np.eye(3)
    h = np.zeros((sx.shape[0] * 3,) * 2, dtype=complex)
    if atom["S"] > 0:
        g_t = atom.get("g_perp", atom["g"])
        h += np.kron(atom["D"] * sz @ sz
                     + TAU * MU_B * (g_t * (field[0] * sx + field[1] * sy)
                                     + atom["g"] * field[2] * sz), ei)
        a_par, a_perp = atom["A"], atom["A_perp"]
        if swap_a:
            a_par, a_perp = a_perp, a_par
        h += a_par * np.kron(sz, iz) + a_perp * (np.kron(sx, ix)
                                                 + np.kron(sy, iy))
    h += np.kron(es, atom["P"] * iz @ iz)
    gamma_n = atom["gN"] * TAU * MU_N
    h += zeeman_sign * gamma_n * np.kron(
        es, field[0] * ix + field[1] * iy + field[2] * iz)
    return h


@pytest.mark.parametrize("mistake", [None, "zeeman_sign", "swap_a"])
def test_nv21_coherent_spectrum(mistake):
    desc = configs._nv21_description(bz=0.0)
    for block in desc[0]:
        block[0]["B0"] = FIELD
    gens, _ = build(_coherent_only(desc))
    kw = {"zeeman_sign": 1.0} if mistake == "zeeman_sign" else \
        {"swap_a": True} if mistake == "swap_a" else {}
    want = _block_spectrum([_n14_h(block[0], FIELD, **kw)
                            for block in desc[0]])
    re, im = _generator_spectrum(gens["L0"])
    assert np.max(np.abs(re)) <= 1e-6 * np.max(np.abs(im))
    err = np.max(np.abs(im - want)) / np.max(np.abs(want))
    if mistake:
        assert err > 1e-9, mistake
    else:
        assert err <= 1e-12


# --- coherence decay: T2 dephasing and T1 ---------------------------------

def _coherence_decay(atom, y, x, dephasing_norm=True):
    e = triplet_energies(atom["D"], atom["g"], atom["B0"][2])
    b = np.exp(-e / (TAU * K_B * atom["T"]))
    out_rate = [sum(b[f] / ((b[i] + b[f]) * atom["TS1"]) for f in range(3))
                for i in range(3)]
    m = [1, 0, -1]
    strength = 1.0 if dephasing_norm else \
        (atom["g"] * TAU * MU_B * atom["B0"][2]) ** 2
    return (strength * (m[y] - m[x]) ** 2 / (2 * atom["TS2"])
            + 0.5 * (out_rate[y] + out_rate[x]))


@pytest.mark.parametrize("mistake", [False, True])
def test_ground_coherence_decay_is_t2_plus_t1(mistake):
    desc = models.nv_7(np.array([0.0, 0.0, 5e-3]))
    gens, vmap = build(desc)
    ground = desc[0][0][0]
    index = {(int(a), int(b), int(c)): k for k, (a, b, c) in enumerate(vmap)}
    for y, x in ((0, 1), (0, 2), (1, 2)):
        re, im = index[(y, x, 0)], index[(y, x, 1)]
        want = _coherence_decay(ground, y, x, dephasing_norm=not mistake)
        for k in (re, im):
            got = -gens["L0"][k, k]
            if mistake:
                assert abs(got - want) > 1e-3 * want
            else:
                assert math.isclose(got, want, rel_tol=1e-12), (y, x)
