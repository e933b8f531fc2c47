"""Host-side logic of the HIP path (no device needed).

ELL packing, basis permutation, Taylor planning, rotating tables, node
times, flattening and waveform records must be exact, because the kernel
consumes them verbatim.
"""

import math

import numpy as np
import pytest
from scipy.linalg import expm

from oracle.reference_simulate import (
    flatten, node_coefficients, rotating_precompute, unflatten)
from superspinsim_amd import configs, quadratures
from superspinsim_amd.generate_generators import make_rotating
from superspinsim_amd.generate_simulator import _flatten, _unflatten, tabulate
from superspinsim_amd.propagator import (
    EllStack, coefficient_samples, eig_lane_factors, pair_permutation,
    choose_layout, fetch_pairs, lds_instructions, pair_fetch_order,
    pair_fetches, quad_gathers, quad_order,
    rotating_tables, taylor_plan, taylor_plan_measured, taylor_remainder,
    width_order)
from superspinsim_amd.sweep import WAVE_DTYPE, Waveform, waveform_records


def _dense_from_ell(ell):
    M = ell.val.shape[0]
    out = np.zeros((M, ell.n, ell.n))
    for sl in range(ell.slots):
        for r in range(ell.n):
            c = r if sl == 0 else ell.col[sl, r]
            if c < ell.n:
                out[:, r, c] += ell.val[:, sl, r]
    return out


def test_ell_roundtrip_and_slot_conventions():
    wl = configs.odmr7(1)
    mats = wl.generators[[0, 1, 4]]
    perm = pair_permutation(np.any(mats != 0, axis=0))
    assert sorted(perm) == list(range(mats.shape[1]))
    pm = mats[:, perm][:, :, perm]
    ell = EllStack(pm)
    assert np.array_equal(_dense_from_ell(ell), pm)
    assert np.all(ell.col[1, :ell.n] == np.arange(ell.n) ^ 1)
    # slot 0 carries the used-slot count of every row
    assert np.all(ell.col[0] >= 2) and ell.col[0].max() == ell.slots


def test_pair_permutation_pairs_coherences():
    """On the NV basis every Re/Im pair of a coherence becomes (2k, 2k+1)."""
    wl = configs.odmr7(1)
    mats = wl.generators[[0, 1, 4]]
    perm = pair_permutation(np.any(mats != 0, axis=0))
    vm = wl.vmap[perm]
    paired = sum(1 for k in range(0, len(perm) - 1, 2)
                 if tuple(vm[k][:2]) == tuple(vm[k + 1][:2])
                 and vm[k][2] != vm[k + 1][2])
    assert paired == 21  # all off-diagonal (y, x) of the 7-level model


def _pairs_bruteforce(pat):
    """Per row: distinct aligned pairs (c // 2) other than its own holding
    an off-diagonal column."""
    n = pat.shape[0]
    return np.array([len({c // 2 for c in range(n) if pat[r, c] and c != r}
                         - {r // 2}) for r in range(n)])


def _odmr7_pattern(use_incoherent):
    wl = configs.odmr7(4, use_incoherent=use_incoherent)
    sim = wl.simulator()
    active, _, _ = sim.program(wl.waveforms())
    return np.asarray(wl.generators)[active]


@pytest.mark.parametrize("use_incoherent", [False, True])
def test_pair_fetch_order_on_nv7(use_incoherent):
    """The 7-level sweep's pattern (n = 49 / 19): gathering column by column
    takes 5 LDS reads per product; the pair-fetch order leaves at most 3
    aligned pairs per row (one 16-byte read each)."""
    mats = _odmr7_pattern(use_incoherent)
    pat = np.any(mats != 0, axis=0)
    n = pat.shape[0]
    perm = pair_permutation(pat)
    pp = pat[perm][:, perm]
    assert EllStack(mats[:, perm][:, :, perm]).gather_slots == 5
    order = pair_fetch_order(pp)
    assert sorted(order) == list(range(n))
    qp = pp[order][:, order]
    assert pair_fetches(qp) == 3
    assert _pairs_bruteforce(qp).max() == 3
    assert [len(f) for f in fetch_pairs(qp)] == list(_pairs_bruteforce(qp))
    ell = EllStack(mats[:, perm[order]][:, :, perm[order]], layout="pairs")
    assert ell.slots == 8 and ell.gather_slots == 3
    # deterministic: same pattern, same order
    assert np.array_equal(pair_fetch_order(pp), order)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 9, 30, 64])
def test_pair_fetch_ell_layout(n):
    """Pair-fetch layout: slot 1 is r ^ 1 and every later slot pair is an
    aligned column pair (2p, 2p + 1) on every padded row; the table
    reproduces the matrices exactly; the order never fetches more pairs
    than the order it started from."""
    rng = np.random.default_rng(n)
    pat = (rng.random((n, n)) < 0.15) | np.eye(n, dtype=bool)
    mats = np.where(pat[None], rng.standard_normal((2, n, n)), 0.0)
    order = pair_fetch_order(pat)
    assert sorted(order) == list(range(n))
    pm = mats[:, order][:, :, order]
    qp = pat[order][:, order]
    ell = EllStack(pm, layout="pairs")
    G = pair_fetches(qp)
    assert ell.layout == "pairs" and ell.slots == 2 + 2 * G
    assert G == (_pairs_bruteforce(qp).max() if n else 0)
    lanes = np.arange(64)
    assert np.array_equal(ell.col[1], lanes ^ 1)
    for k in range(G):
        assert np.all(ell.col[2 + 2 * k] % 2 == 0)
        assert np.array_equal(ell.col[3 + 2 * k], ell.col[2 + 2 * k] + 1)
    assert np.array_equal(_dense_from_ell(ell), pm)
    assert np.all(ell.val[:, :, n:] == 0)
    assert G <= pair_fetches(pat)
    with pytest.raises(ValueError):
        EllStack(pm, partner=False, layout="pairs")


def _outside_quad_bruteforce(pat):
    """Per row: off-diagonal columns whose lane is not r ^ 1, r ^ 2, r ^ 3."""
    n = pat.shape[0]
    return np.array([sum(1 for c in range(n) if pat[r, c] and c != r
                         and c not in (r ^ 1, r ^ 2, r ^ 3))
                     for r in range(n)])


@pytest.mark.parametrize("use_incoherent", [False, True])
def test_quad_order_leaves_three_gathers_on_nv7(use_incoherent):
    """The 7-level sweep's pattern (n = 49 / 19): the partner-only layout
    gathers 5 columns per row through LDS; the quad order leaves 3 (rows
    may land anywhere), 4 when (2k, 2k + 1) pairs must stay together."""
    mats = _odmr7_pattern(use_incoherent)
    pat = np.any(mats != 0, axis=0)
    n = pat.shape[0]
    perm = pair_permutation(pat)
    pp = pat[perm][:, perm]
    assert EllStack(mats[:, perm][:, :, perm]).gather_slots == 5
    for pair_units, expect in ((False, 3), (True, 4)):
        order = quad_order(pp, pair_units=pair_units)
        assert sorted(order) == list(range(n))
        qp = pp[order][:, order]
        assert quad_gathers(qp) == expect
        assert _outside_quad_bruteforce(qp).max() == expect
        if pair_units:
            assert np.all(order[0:2 * (n // 2):2] % 2 == 0)
            assert np.all(order[1:2 * (n // 2):2] ==
                          order[0:2 * (n // 2):2] + 1)
            if n % 2:
                assert order[-1] == n - 1
    # deterministic: same pattern, same order
    assert np.array_equal(quad_order(pp), quad_order(pp))


@pytest.mark.parametrize("n", [1, 3, 4, 9, 30, 64])
def test_quad_ell_layout(n):
    """Quad layout: slots 1..3 are r ^ 1 and the other pair of the quad
    (even row, odd row) on every padded row, the other columns follow, and
    the table reproduces the matrices exactly."""
    rng = np.random.default_rng(n)
    pat = (rng.random((n, n)) < 0.15) | np.eye(n, dtype=bool)
    mats = np.where(pat[None], rng.standard_normal((2, n, n)), 0.0)
    order = quad_order(pat)
    assert sorted(order) == list(range(n))
    pm = mats[:, order][:, :, order]
    qp = pat[order][:, order]
    ell = EllStack(pm, layout="quad")
    assert ell.layout == "quad" and ell.slots == 4 + quad_gathers(qp)
    lanes = np.arange(64)
    other = (lanes & ~1) ^ 2
    for k, c in enumerate((lanes ^ 1, other, other | 1), start=1):
        assert np.array_equal(ell.col[k], c)
    # the three columns are the other members of the lane's quad
    assert np.all(np.sort(ell.col[1:4], axis=0) ==
                  np.sort(np.stack([lanes ^ 1, lanes ^ 2, lanes ^ 3]),
                          axis=0))
    assert np.array_equal(_dense_from_ell(ell), pm)
    assert np.all(ell.val[:, :, n:] == 0)
    assert quad_gathers(qp) == _outside_quad_bruteforce(qp).max()
    if n > 4:
        # never worse than the order it started from
        assert quad_gathers(qp) <= quad_gathers(pat)
    with pytest.raises(ValueError):
        EllStack(pm, partner=False, layout="quad")


def _bank_passes(addrs, group, banks):
    """LDS cycles of one read: per group of lanes, the most distinct
    addresses that share a bank."""
    total = 0
    for g0 in range(0, addrs.size, group):
        per = {}
        for a in set(addrs[g0:g0 + group].tolist()):
            per[a % banks] = per.get(a % banks, 0) + 1
        total += max(per.values())
    return total


@pytest.mark.parametrize("layout", ["quad", "pairs"])
@pytest.mark.parametrize("n", [19, 49, 64])
def test_bank_aware_slot_order(layout, n):
    """Bank-aware slot order: same layout rules and the same matrices, and
    no more bank passes slot by slot summed (8-lane groups for the 16-byte
    reads, 16-lane groups for single gathers)."""
    rng = np.random.default_rng(n)
    pat = (rng.random((n, n)) < 0.1) | np.eye(n, dtype=bool)
    mats = np.where(pat[None], rng.standard_normal((2, n, n)), 0.0)
    order, _ = choose_layout(pat, layout)
    pm = mats[:, order][:, :, order]
    plain_ord = EllStack(pm, layout=layout)
    aware = EllStack(pm, layout=layout, bank_aware=True)
    assert aware.slots == plain_ord.slots
    assert np.array_equal(_dense_from_ell(aware), pm)
    assert np.all(aware.val[:, :, n:] == 0)
    # every non-zero lies below the row's used-slot count
    used = aware.col[0, :n]
    for r in range(n):
        for sl in range(int(used[r]), aware.slots):
            assert np.all(aware.val[:, sl, r] == 0)
    lanes = np.arange(64)
    assert np.array_equal(aware.col[1], lanes ^ 1)
    npair = 1 if layout == "quad" else (aware.slots - 2) // 2
    for k in range(npair):
        assert np.all(aware.col[2 + 2 * k] % 2 == 0)
        assert np.array_equal(aware.col[3 + 2 * k], aware.col[2 + 2 * k] + 1)

    def passes(ell):
        c = ell.col.astype(np.int64)
        tot = sum(_bank_passes(c[2 + 2 * k] >> 1, 8, 8)
                  for k in range(npair))
        return tot + sum(_bank_passes(c[sl], 16, 16)
                         for sl in range(2 + 2 * npair, ell.slots))
    assert passes(aware) <= passes(plain_ord)


@pytest.mark.parametrize("use_incoherent,expect,instr",
                         [(False, "quad", 4), (True, "quad", 4)])
def test_auto_layout_on_nv7(use_incoherent, expect, instr):
    """The 7-level sweep (n = 49 and 19): quads, below the plain layout's 6
    LDS instructions per product."""
    mats = _odmr7_pattern(use_incoherent)
    pat = np.any(mats != 0, axis=0)
    perm = pair_permutation(pat)
    pp = pat[perm][:, perm]
    assert lds_instructions(pp, "plain") == 6
    order, layout = choose_layout(pp)
    assert layout == expect
    assert lds_instructions(pp[order][:, order], layout) == instr
    # the rotating frame keeps its (2k, 2k + 1) blocks together
    order, layout = choose_layout(pp, pair_units=True)
    n = pp.shape[0]
    assert np.all(order[1:2 * (n // 2):2] == order[0:2 * (n // 2):2] + 1)


@pytest.mark.parametrize("n", [9, 10, 171])
def test_width_order_keeps_pairs_and_sorts_units(n):
    rng = np.random.default_rng(n)
    pat = rng.random((n, n)) < 0.08
    pat |= np.eye(n, dtype=bool)
    order = width_order(pat)
    assert sorted(order) == list(range(n))
    # aligned pairs move as units; an odd last row stays last
    assert np.all(order[1:2 * (n // 2):2] == order[0:2 * (n // 2):2] + 1)
    assert np.all(order[0:2 * (n // 2):2] % 2 == 0)
    if n % 2:
        assert order[-1] == n - 1
    # widths (used ELL slots) of the reordered matrix: nonincreasing units
    pm = pat[order][:, order]
    ell = EllStack(pm[None].astype(float) + 0.0)
    w = ell.col[0, :n]
    units = np.maximum(w[0:2 * (n // 2):2], w[1:2 * (n // 2):2])
    assert np.all(np.diff(units) <= 0)
    # the reordering is a similarity: same matrix up to relabelling
    assert np.array_equal(pm, pat[np.ix_(order, order)])


def test_wide_propagator_orders_pairs_by_width():
    """The 21-level model's slot counts per 64-row wavefront after the
    width ordering (what the wide kernel reads)."""
    wl = configs.odmr21(1)
    mats = np.stack([wl.generators[0] + wl.excitation * wl.generators[4],
                     wl.generators[1]])
    pat = np.any(mats != 0, axis=0)
    perm = pair_permutation(pat)
    perm = perm[width_order(pat[perm][:, perm])]
    ell = EllStack(mats[:, perm][:, :, perm])
    per_wave = [int(ell.col[0, w * 64:(w + 1) * 64].max())
                for w in range(ell.npad // 64)]
    assert per_wave[0] == ell.slots == 14
    assert per_wave == sorted(per_wave, reverse=True)
    assert sum(per_wave) <= 64


def test_wide_adjoint_tables():
    """Extended forward, transposed and derivative tables for n > 64."""
    from superspinsim_amd.adjoint import wide_adjoint_ell
    rng = np.random.default_rng(7)
    n, M, nt = 130, 2, 3
    mats = np.where(rng.random((M, n, n)) < 0.04,
                    rng.standard_normal((M, n, n)), 0.0)
    mats[:, np.arange(n), np.arange(n)] = 1.0
    pat = np.any(mats != 0, axis=0)
    perm = pair_permutation(pat)
    perm = perm[width_order(pat[perm][:, perm])]
    mp = mats[:, perm][:, :, perm]
    dmats = np.where(rng.random((nt, n, n)) < 0.03,
                     rng.standard_normal((nt, n, n)), 0.0)
    ell = EllStack(mp)
    col_e, val_e, col_t, dval_e, slots_f = wide_adjoint_ell(ell, mp, dmats)
    NP = ell.npad
    assert slots_f == ell.slots and col_e.shape[1] == NP
    used = col_e[0] & 0xFF
    assert np.array_equal(col_e[0] >> 8, ell.col[0])
    # the matrices' columns first, in the propagator's order and values
    assert np.array_equal(col_e[1:ell.slots][:, :n][ell.col[1:, :n] != 
                          np.arange(n)], ell.col[1:, :n][ell.col[1:, :n] !=
                          np.arange(n)])
    assert np.array_equal(val_e[:, :ell.slots], ell.val)
    assert not np.any(val_e[:, ell.slots:])
    # D is exactly representable on the extended pattern
    D = np.zeros((nt, n, n))
    for r in range(n):
        for sl in range(int(used[r])):
            c = r if sl == 0 else int(col_e[sl, r])
            if c < n:
                D[:, r, c] += dval_e[:, sl, r]
    assert np.array_equal(D, dmats)
    # transposed pattern covers A^T; every entry has a staged source that
    # the kernel's scan of row c finds (slot 0 diagonal, 1 partner, else
    # the slot of column r among the matrices' slots)
    for r in range(n):
        cols = {r, r ^ 1} | {int(c) for c in col_t[2:col_t[0, r], r]}
        assert set(np.nonzero(pat[perm][:, perm][:, r])[0]) <= cols
        for sl in range(2, int(col_t[0, r])):
            c = int(col_t[sl, r])
            hits = [k for k in range(2, slots_f) if ell.col[k, c] == r]
            assert len(hits) == 1


def test_transposed_adjoint_tables():
    """The adjoint chain's own pattern holds A^T and D^T exactly once per
    entry, keeps slot 1 the partner, and needs fewer modelled LDS passes
    than the shared forward/transposed pattern."""
    from superspinsim_amd.adjoint import (shared_adjoint_ell,
                                          transposed_adjoint_ell)
    rng = np.random.default_rng(7)
    n = 49
    pat = rng.random((n, n)) < 0.08
    pat |= np.eye(n, dtype=bool)
    mats = np.stack([pat * rng.normal(size=(n, n)),
                     (rng.random((n, n)) < 0.04) * rng.normal(size=(n, n))])
    dmats = np.stack([(rng.random((n, n)) < 0.03) * rng.normal(size=(n, n))
                      for _ in range(3)])
    col_t, val_t, dval_t = transposed_adjoint_ell(mats, dmats)
    lanes = np.arange(64)
    assert np.array_equal(col_t[1], lanes ^ 1)
    At = np.zeros((2, 64, 64))
    Dt = np.zeros((3, 64, 64))
    for r in range(64):
        for sl in range(col_t.shape[0]):
            c = r if sl == 0 else int(col_t[sl, r])
            At[:, r, c] += val_t[:, sl, r]
            Dt[:, r, c] += dval_t[:, sl, r]
    assert np.array_equal(At[:, :n, :n], mats.transpose(0, 2, 1))
    assert np.array_equal(Dt[:, :n, :n], dmats.transpose(0, 2, 1))
    assert not At[:, n:].any() and not At[:, :, n:].any()

    def passes(col):
        tot = 0
        for sl in range(2, col.shape[0]):
            for g in range(0, 64, 16):
                per = {}
                for a in {int(col[sl, r]) for r in range(g, g + 16)}:
                    per[a % 16] = per.get(a % 16, 0) + 1
                tot += max(per.values())
        return tot

    fwd = np.tile(lanes, (7, 1)).astype(np.int32)
    fwd[1] = lanes ^ 1
    for r in range(n):
        cols = [c for c in np.nonzero(pat[r])[0] if c not in (r, r ^ 1)][:5]
        fwd[2:2 + len(cols), r] = cols
        fwd[0, r] = 2 + len(cols)
    shared = shared_adjoint_ell(fwd, mats, dmats)[0]
    assert passes(col_t) <= passes(shared)


@pytest.mark.parametrize("norm", [0.01, 0.19, 0.8, 3.0, 40.0])
def test_taylor_plan_meets_tolerance(norm):
    m, s = taylor_plan(norm, tol=1e-16)
    assert s * taylor_remainder(norm / s, m) <= 1e-16
    if m > 1:
        assert s * taylor_remainder(norm / s, m - 1) > 1e-16 or s > 1


def test_measured_plan_bounds_actual_truncation():
    rng = np.random.default_rng(3)
    A = 0.3 * rng.standard_normal((12, 12)) / math.sqrt(12)
    m, s = taylor_plan_measured([A], tol=1e-14)
    x = rng.standard_normal(12)
    y = x.copy()
    B = A / s
    for _ in range(s):
        term = y.copy()
        acc = y.copy()
        for k in range(1, m + 1):
            term = B @ term / k
            acc += term
        y = acc
    exact = expm(A) @ x
    assert np.max(np.abs(y - exact)) / np.max(np.abs(exact)) < 1e-13
    m2, _ = taylor_plan_measured([A], tol=1e-8)
    assert m2 < m


def test_coefficient_samples_cover_corners():
    lo = np.array([[1.0, -2.0, 0.5]])
    hi = np.array([[1.0, 2.0, 0.5]])
    smp = coefficient_samples(lo, hi, count=4)
    assert smp.shape[1:] == (1, 3)
    assert np.any(smp[:, 0, 1] == -2.0) and np.any(smp[:, 0, 1] == 2.0)
    assert np.all(smp[:, 0, 0] == 1.0) and np.all(smp[:, 0, 2] == 0.5)


def test_rotating_tables_match_oracle_precompute():
    wl = configs.odmr7(1, use_incoherent=True)
    g = wl.generators
    rot, V, Vi, doubles, singles = make_rotating(g[0], list(g[1:]))
    rot = np.array(rot)
    nodes, _, _ = quadratures.scheme(2)
    for residual in (True, False):
        a, da, sa = rotating_tables(rot, doubles, singles, nodes, 2e-9, 100,
                                    residual)
        b, db, sb = rotating_precompute(rot, doubles, singles, nodes, 2e-9,
                                        100, residual)
        assert np.max(np.abs(a - b)) <= 1e-15 * np.abs(b).max()
        assert np.array_equal(da, db) and np.array_equal(sa, sb)


def test_eig_lane_factors_layout():
    d = np.array([[0.1, 0.2], [0.3, 0.4]])
    s = np.array([0.5])
    f = eig_lane_factors(d, s, 5)
    own, part = f[:64], f[64:]
    assert own[:5].tolist() == [0.1, 0.1, 0.3, 0.3, 0.5]
    assert part[:5].tolist() == [-0.2, 0.2, -0.4, 0.4, 0.0]


def test_node_times_match_oracle_bit_for_bit():
    nodes, _, _ = quadratures.scheme(5)
    rows = []
    sampler = lambda t, row: row.__setitem__(0, t)  # noqa: E731
    mine = tabulate(sampler, 1, 7, 13, nodes, 1.5e-7, 3e-9)[..., 0]
    ref = node_coefficients(sampler, 1, 7, nodes.size, 13, nodes, 1.5e-7,
                            3e-9)[..., 0]
    assert np.array_equal(mine, np.transpose(ref, (1, 0, 2)))
    del rows


def test_flatten_roundtrip_matches_oracle():
    rng = np.random.default_rng(4)
    r = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    r = r + r.conj().T
    vmap = np.array([[y, x, c] for y in range(4) for x in range(y, 4)
                     for c in ((0,) if x == y else (0, 1))])
    order = np.argsort([(0 if y == x else 1, y, x, c) for y, x, c in vmap],
                       axis=0)[:, 0]
    vmap = vmap[order]
    for herm, n_d in ((True, vmap.shape[0]), (False, 32)):
        a = _flatten(r, vmap, herm, n_d)
        b = flatten(r, vmap, herm, n_d)
        assert np.array_equal(a, b)
        back = _unflatten(a[None], vmap, herm, 4)[0]
        assert np.allclose(back, r)
        assert np.array_equal(back, unflatten(b[None], vmap, herm, 4)[0])


def test_waveform_records_layout():
    w = Waveform(3, base=[1.0, 0.0, 0.5])
    w.tone(1, 2e-4, 2.87e9, t_start=0.0, t_end=1e-6)
    w.add(1, 2e-6, 3e-6, 1e-4)
    recs, segs = waveform_records([w, w], 3)
    assert recs.dtype == WAVE_DTYPE and recs.itemsize == 128
    assert recs["seg_count"][0].tolist()[:3] == [0, 2, 0]
    assert recs["seg_begin"][1, 1] == 2
    assert segs.shape == (4, 5)
    assert segs[0, 3] == math.tau * 2.87e9
    assert w(0.5e-6)[1] == 2e-4 * math.cos(math.tau * 2.87e9 * 0.5e-6)
    assert w(1.5e-6)[1] == 0.0 and w(2.5e-6)[1] == 1e-4


def test_overlapping_segments_rejected():
    w = Waveform(1)
    w.add(0, 0.0, 2.0, 1.0)
    w.add(0, 1.0, 3.0, 1.0)
    with pytest.raises(ValueError):
        waveform_records([w], 1)


def test_pulsed_protocol_workload():
    from superspinsim_amd import configs
    wl = configs.pulsed7(count=512)
    assert len(wl.waves) == 512
    assert np.all(wl.read_from < wl.steps)
    # readout window is 1 us = 50,000 steps of 20 ps (rounding of the ends)
    assert np.all(np.abs((wl.steps - wl.read_from) - 50000) <= 1)
    assert wl.steps.min() >= 0.1e-6 / 20e-12
    assert wl.steps.max() <= (4e-6 + 10e-6 + 1e-6) / 20e-12 + 1
    for w, ro in zip(wl.waves[:5], wl.read_from[:5]):
        # laser off before the readout window, on inside it
        assert w(ro * 20e-12 - 1e-9)[4] == 0.0
        assert w(ro * 20e-12 + 1e-9)[4] > 0.0


def test_trapezoid_window():
    from superspinsim_amd.protocols import trapezoid_window
    s = np.arange(11, dtype=float)[None] ** 2
    got = trapezoid_window(s, [2], [7], 0.5)
    want = 0.5 * 0.5 * sum(s[0, k] + s[0, k + 1] for k in range(2, 7))
    assert got[0] == want
    assert trapezoid_window(s, [5], [5], 0.5)[0] == 0.0


def test_waveform_shift_reproduces_the_later_clock():
    """shifted(T)(t) = w(t + T) to rounding: tones keep their phase, windows
    move."""
    import math
    from superspinsim_amd.sweep import Waveform
    w = Waveform(3, base=[1.0, 0.0, 0.2])
    w.tone(1, 3e-4, 2.87e9, phase=0.4)
    w.add(2, 1e-6, 2e-6, 0.5, omega=1e7)
    T = 7.3e-7
    sw = w.shifted(T)
    for t in np.linspace(0, 2e-6, 97):
        a, b = sw(t), w(t + T)
        assert np.allclose(a, b, rtol=0, atol=1e-14 * 0.5), t
    assert sw.segments[2][0][0] == 1e-6 - T
    assert abs(sw.segments[1][0][4]) <= math.pi


def test_chunk_plan_covers_every_output():
    from superspinsim_amd.sweep import BatchSimulator
    sim = BatchSimulator(np.zeros((2, 9, 9)), 2, 50)
    for n_inst, N in ((256, 999), (3, 40), (1, 7), (5, 1)):
        P, L, k = sim.chunk_plan(n_inst, N, resident=4096)
        assert (P - 1) * L < N <= P * L
        assert k * 9 <= 64
