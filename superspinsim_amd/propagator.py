"""Device-side propagation: ELL packing, Taylor planning, kernel launch.

``Propagator`` owns the device copies of a generator stack (HBM, via torch
tensors on the HIP device) and launches ``ssmx_propagate`` for any number of
instances.  It is the layer both the drop-in ``generate_simulator`` and the
batched sweep API (``superspinsim_amd.sweep``) call.
"""

import ctypes
import math

import numpy as np

from superspinsim_amd import _lib

_f64 = np.float64


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError(
            "superspinsim_amd needs a HIP device (torch.cuda.is_available() "
            "is False); there is no CPU fallback")
    return torch


def device_tensor(array, dtype=None):
    torch = _torch()
    arr = np.ascontiguousarray(array)
    t = torch.from_numpy(arr)
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


# ---------------------------------------------------------------------------
# ELL packing
# ---------------------------------------------------------------------------

def padded_rows(n):
    """Row stride of the device tables: 64 * ceil(n / 64)."""
    return 64 * ((int(n) + 63) // 64)


def pair_permutation(pattern):
    """Basis order that puts coupled index pairs at (2k, 2k + 1).

    ``pattern`` is the (n, n) boolean union of the generators' non-zeros.
    Pairs (a, b) coupled both ways are matched greedily (most couplings
    first); matched pairs come first, unmatched indices after.  Returns
    ``perm`` with ``new[k] = old[perm[k]]``.
    """
    n = pattern.shape[0]
    weight = pattern.astype(int) + pattern.T.astype(int)
    cand = [(weight[a, b], a, b) for a in range(n) for b in range(a + 1, n)
            if pattern[a, b] and pattern[b, a]]
    cand.sort(key=lambda t: (-t[0], t[1], t[2]))
    used = np.zeros(n, dtype=bool)
    order = []
    for _, a, b in cand:
        if not used[a] and not used[b]:
            used[a] = used[b] = True
            order += [a, b]
    order += [k for k in range(n) if not used[k]]
    return np.array(order, dtype=np.int64)


def width_order(pattern, partner=True):
    """Order of aligned row pairs (2k, 2k + 1), widest first.

    ``pattern`` is the (n, n) union pattern in the paired basis.  The width
    of a row is its used ELL slots (diagonal, partner, other non-zeros);
    pairs move as units, so a pair stays at (2k, 2k + 1), and an odd last
    row stays last.  Returns ``order`` with ``new[k] = old[order[k]]``.
    """
    n = pattern.shape[0]
    fixed = 2 if partner else 1
    width = np.empty(n, dtype=np.int64)
    for r in range(n):
        others = np.count_nonzero(pattern[r])
        others -= int(pattern[r, r])
        if partner and (r ^ 1) < n:
            others -= int(pattern[r, r ^ 1])
        width[r] = fixed + others
    units = n // 2
    unit_width = np.maximum(width[0:2 * units:2], width[1:2 * units:2])
    rank = np.argsort(-unit_width, kind="stable")
    order = np.stack([2 * rank, 2 * rank + 1], axis=1).ravel()
    if n % 2:
        order = np.append(order, n - 1)
    return order.astype(np.int64)


def _off_diagonal(pattern):
    need = np.asarray(pattern, dtype=bool).copy()
    np.fill_diagonal(need, False)
    return need


def fetch_pairs(pattern):
    """Per row, the aligned pairs (2p, 2p + 1) other than its own that hold
    one of its off-diagonal columns, ascending.

    In the pair-fetch layout each such pair costs one 16-byte LDS read per
    product (both columns at once); the row's own pair is its partner
    r ^ 1, read by DPP.
    """
    need = _off_diagonal(pattern)
    n = need.shape[0]
    return [sorted({int(c) >> 1 for c in np.nonzero(need[r])[0]}
                   - {r >> 1}) for r in range(n)]


def pair_fetches(pattern):
    """Largest number of fetched pairs over the rows (``fetch_pairs``)."""
    pairs = fetch_pairs(pattern)
    return max((len(p) for p in pairs), default=0)


def quad_columns(r):
    """Columns of slots 1, 2, 3 of row r in the quad layout: r ^ 1, then
    the even and the odd row of the other pair of r's quad."""
    other = (r & ~1) ^ 2
    return (r ^ 1, other, other | 1)


def quad_gathers(pattern):
    """Largest number of columns of a row outside its quad (rows 4q..4q+3).

    These are the LDS gathers per product left in the quad layout, beside
    the write; the quad's other lanes come by DPP.
    """
    need = _off_diagonal(pattern)
    n = need.shape[0]
    grp = np.arange(n) // 4
    outside = need & (grp[:, None] != grp[None, :])
    return int(outside.sum(axis=1).max()) if n else 0


def quad_order(pattern, pair_units=False, restarts=8, seed=0):
    """Basis order that puts coupled rows into common quads (4q .. 4q + 3).

    Within a quad every row reads the three other lanes by DPP, so only
    the columns outside its quad cost an LDS gather.  The order minimises, lexicographically, the largest count of such
    columns over the rows (the wavefront executes that many gathers per
    product), the number of rows at that count, and their sum, by steepest
    descent over swaps between quads, from the current order and from a few
    seeded random starts.  ``pair_units`` keeps the rows (2k, 2k + 1) of the
    current basis together as one unit (the rotating frame's 2x2
    eigen-blocks), so pairs stay aligned.  With n not a multiple of 4 the
    last quad is the partial one (an odd last row stays last).  Returns
    ``order`` with ``new[k] = old[order[k]]``.
    """
    need = _off_diagonal(pattern).astype(np.int64)
    n = need.shape[0]
    if n <= 4:
        return np.arange(n, dtype=np.int64)
    size = 2 if pair_units else 1
    per_quad = 4 // size
    U = (n + size - 1) // size
    Q = (U + per_quad - 1) // per_quad
    # unit u holds rows size*u .. ; a padding row n and a padding unit U
    # (no couplings) fill partial quads
    unit_rows = np.full((U + 1, size), n, dtype=np.int64)
    for u in range(U):
        rows = np.arange(size * u, min(size * u + size, n))
        unit_rows[u, :rows.size] = rows
    need_p = np.zeros((n + 1, n + 1), dtype=np.int64)
    need_p[:n, :n] = need
    Nu = need_p[:, unit_rows].sum(axis=2)            # (n + 1, U + 1)
    base = need_p.sum(axis=1)
    fixed_tail = 1 if (pair_units and n % 2) else 0
    movable = U - fixed_tail
    rng = np.random.default_rng(seed)
    ii, jj = np.triu_indices(movable, 1)

    def evaluate(slot_unit):
        quads = slot_unit.reshape(Q, per_quad)
        S = Nu[:, quads].sum(axis=2)                  # (n + 1, Q)
        q_of_unit = np.zeros(U + 1, dtype=np.int64)
        q_of_unit[quads.ravel()] = np.repeat(np.arange(Q), per_quad)
        unit_of_row = np.full(n + 1, U, dtype=np.int64)
        for u in range(U):
            unit_of_row[unit_rows[u][unit_rows[u] < n]] = u
        q_of_row = q_of_unit[unit_of_row]
        rem = base - S[np.arange(n + 1), q_of_row]
        rem[n] = 0
        return quads, S, q_of_row, unit_of_row, rem

    def key_of(rem):
        m = rem.max()
        return (int(m), int((rem == m).sum()), int(rem.sum()))

    best = None
    for start in range(restarts + 1):
        order = np.arange(movable) if start == 0 else \
            rng.permutation(movable)
        slot_unit = np.full(Q * per_quad, U, dtype=np.int64)
        slot_unit[:movable] = order
        if fixed_tail:
            slot_unit[movable] = U - 1
        while True:
            quads, S, q_of_row, unit_of_row, rem = evaluate(slot_unit)
            cur = key_of(rem[:n])
            qa, qb = ii // per_quad, jj // per_quad
            ok = qa != qb
            pi, pj, qa, qb = ii[ok], jj[ok], qa[ok], qb[ok]
            ua, ub = slot_unit[pi], slot_unit[pj]
            if pi.size == 0:
                break
            rowsA = unit_rows[quads[qa]].reshape(pi.size, -1)   # (P, 4)
            rowsB = unit_rows[quads[qb]].reshape(pi.size, -1)
            uA = unit_of_row[rowsA]
            uB = unit_of_row[rowsB]
            Sa_A = S[rowsA, qa[:, None]]
            Sa_B = S[rowsA, qb[:, None]]
            Sb_A = S[rowsB, qa[:, None]]
            Sb_B = S[rowsB, qb[:, None]]
            na_a, na_b = Nu[rowsA, ua[:, None]], Nu[rowsA, ub[:, None]]
            nb_a, nb_b = Nu[rowsB, ua[:, None]], Nu[rowsB, ub[:, None]]
            covA = np.where(uA == ua[:, None], Sa_B - na_b + na_a,
                            Sa_A - na_a + na_b)
            covB = np.where(uB == ub[:, None], Sb_A - nb_a + nb_b,
                            Sb_B - nb_b + nb_a)
            R = np.broadcast_to(rem, (pi.size, n + 1)).copy()
            rows = np.arange(pi.size)[:, None]
            R[rows, rowsA] = base[rowsA] - covA
            R[rows, rowsB] = base[rowsB] - covB
            R[:, n] = 0
            mx = R.max(axis=1)
            cnt = (R == mx[:, None]).sum(axis=1)
            tot = R.sum(axis=1)
            k = int(np.lexsort((tot, cnt, mx))[0])
            if (int(mx[k]), int(cnt[k]), int(tot[k])) >= cur:
                break
            slot_unit[pi[k]], slot_unit[pj[k]] = \
                slot_unit[pj[k]], slot_unit[pi[k]]
        if best is None or cur < best[0]:
            best = (cur, slot_unit.copy())
    order = [int(r) for u in best[1] if u < U for r in unit_rows[u]
             if r < n]
    return np.array(order, dtype=np.int64)


_PAIR_FETCH_CACHE = {}


def pair_fetch_order(pattern, attempts=4, seed=0):
    """Basis order that minimises the pairs each row fetches.

    Rows are placed at positions 0..n-1; positions (2p, 2p + 1) form pair p.
    A row reads its partner by DPP and every other pair holding one of its
    columns with one 16-byte LDS read, so the wavefront executes
    1 + max_r(pairs of row r) LDS instructions per product (with the write
    of its own component).  Search: steepest descent over swaps of two rows
    in different pairs on (max, rows at max, sum), then simulated annealing
    of the excess over a target one below the best maximum (up to
    ``attempts`` seeded runs per target), lowering the target until it
    fails or reaches the bound max_r ceil((d_r - 1) / 2) for d_r
    off-diagonal columns (the partner is free, a pair holds two).
    Deterministic for a given seed; cached per pattern.  Returns ``order``
    with ``new[k] = old[order[k]]``.
    """
    need = _off_diagonal(pattern).astype(np.int64)
    n = need.shape[0]
    if n <= 2:
        return np.arange(n, dtype=np.int64)
    key = (n, np.packbits(need.astype(bool)).tobytes(), attempts, seed)
    if key in _PAIR_FETCH_CACHE:
        return _PAIR_FETCH_CACHE[key].copy()
    npairs = (n + 1) // 2
    rows = np.arange(n)
    pair_ind = np.eye(npairs, dtype=np.int64)

    def state(pos):
        pr = pos >> 1
        Mc = need @ pair_ind[pr]                  # columns of row r in pair p
        return Mc, pr, (Mc > 0).sum(axis=1) - (Mc[rows, pr] > 0)

    # steepest descent on (max, rows at max, sum) over all swaps
    ii, jj = np.triu_indices(n, 1)
    pos = np.arange(n)
    while True:
        Mc, pr, cost = state(pos)
        m = int(cost.max())
        cur = (m, int((cost == m).sum()), int(cost.sum()))
        ok = pr[ii] != pr[jj]
        I, J = ii[ok], jj[ok]
        pi, pj = pr[I], pr[J]
        Mi = Mc[:, pi].T - need[:, I].T + need[:, J].T         # (P, n)
        Mj = Mc[:, pj].T - need[:, J].T + need[:, I].T
        D = (Mc > 0).sum(axis=1)[None, :] - (Mc[:, pi].T > 0) \
            - (Mc[:, pj].T > 0) + (Mi > 0) + (Mj > 0)
        prn = np.broadcast_to(pr, (I.size, n)).copy()
        k = np.arange(I.size)
        prn[k, I] = pj
        prn[k, J] = pi
        Mown = np.where(prn == pi[:, None], Mi,
                        np.where(prn == pj[:, None], Mj,
                                 Mc[rows, pr][None, :]))
        C = D - (Mown > 0)
        mx = C.max(axis=1)
        cnt = (C == mx[:, None]).sum(axis=1)
        tot = C.sum(axis=1)
        b = int(np.lexsort((tot, cnt, mx))[0])
        if (int(mx[b]), int(cnt[b]), int(tot[b])) >= cur:
            break
        pos[I[b]], pos[J[b]] = pos[J[b]], pos[I[b]]

    def anneal(start, target, rs, steps, temp0=0.3, batch=32):
        # Metropolis over random swaps, `batch` candidates evaluated at once
        # against the current order; the first accepted one is applied and
        # the candidates examined count as steps
        pos = start.copy()
        Mc, pr, cost = state(pos)
        Mc = Mc.copy()
        dist = (Mc > 0).sum(axis=1)               # distinct pairs per row
        own = Mc[rows, pr] > 0
        E = float(np.maximum(cost - target, 0).sum()) + 0.01 * cost.sum()
        k = np.arange(batch)
        st = 0
        while st < steps:
            temp = temp0 * (1.0 - st / steps) + 1e-3
            I = rs.integers(n, size=batch)
            J = rs.integers(n, size=batch)
            u = rs.random(batch)
            pi, pj = pr[I], pr[J]
            # only pairs pi and pj change: their column counts, and the
            # own pair of the rows now in them
            ci = Mc[:, pi].T + need[:, J].T - need[:, I].T     # (batch, n)
            cj = Mc[:, pj].T + need[:, I].T - need[:, J].T
            dist2 = dist[None, :] - (Mc[:, pi].T > 0) \
                - (Mc[:, pj].T > 0) + (ci > 0) + (cj > 0)
            pr2 = np.broadcast_to(pr, (batch, n)).copy()
            pr2[k, I] = pj
            pr2[k, J] = pi
            own2 = np.where(pr2 == pi[:, None], ci > 0,
                            np.where(pr2 == pj[:, None], cj > 0,
                                     own[None, :]))
            c2 = dist2 - own2
            E2 = np.maximum(c2 - target, 0).sum(axis=1) + 0.01 * c2.sum(axis=1)
            ok = (pi != pj) & ((E2 <= E) |
                               (u < np.exp(np.minimum(E - E2, 0.0) / temp)))
            hits = np.flatnonzero(ok)
            if hits.size == 0:
                st += batch
                continue
            b = int(hits[0])
            st += b + 1
            Mc[:, pi[b]], Mc[:, pj[b]] = ci[b], cj[b]
            pr, dist, own = pr2[b], dist2[b], own2[b]
            cost, E = c2[b], float(E2[b])
            a, c = int(I[b]), int(J[b])
            pos[a], pos[c] = pos[c], pos[a]
            if cost.max() <= target:
                return pos
        return None

    degree = need.sum(axis=1)
    bound = int(np.max(degree // 2))      # ceil((d - 1) / 2)
    best = pos
    T = int(state(best)[2].max())
    rs = np.random.default_rng(seed)
    while T > bound:
        found = None
        for _ in range(attempts):
            found = anneal(best, T - 1, rs, 1500 * n)
            if found is not None:
                break
        if found is None:
            break
        best = found
        T -= 1
    order = np.empty(n, dtype=np.int64)
    order[best] = np.arange(n)
    _PAIR_FETCH_CACHE[key] = order.copy()
    return order


def _bank_slots(entries, G, pad, group, banks):
    """Order each row's fetch addresses over G slots so that, slot by slot,
    the lanes of one LDS cycle (``group`` consecutive lanes) read few
    distinct addresses that share a bank (``address % banks``).

    ``entries[r]`` lists row r's addresses (any order, at most G); slots
    left over take an address another lane of the group reads in that slot
    (a broadcast), else ``pad(r)``.  Local search by swapping two slots of
    one row, group by group, while the summed passes drop.  Returns an
    (R, G) integer array.
    """
    R = len(entries)
    out = np.zeros((R, G), dtype=np.int64)

    def slot_passes(cur, k):
        per = {}
        for a in {sl[k] for sl in cur if sl[k] is not None}:
            per[a % banks] = per.get(a % banks, 0) + 1
        return max(per.values(), default=0)

    for g0 in range(0, R, group):
        rows = list(range(g0, min(g0 + group, R)))
        cur = [list(entries[r]) + [None] * (G - len(entries[r]))
               for r in rows]
        cost = [slot_passes(cur, k) for k in range(G)]
        improved = True
        while improved:
            improved = False
            for i in range(len(rows)):
                for k1 in range(G):
                    for k2 in range(k1 + 1, G):
                        a, b = cur[i][k1], cur[i][k2]
                        if a == b:
                            continue
                        cur[i][k1], cur[i][k2] = b, a
                        c1, c2 = slot_passes(cur, k1), slot_passes(cur, k2)
                        if c1 + c2 < cost[k1] + cost[k2]:
                            cost[k1], cost[k2] = c1, c2
                            improved = True
                        else:
                            cur[i][k1], cur[i][k2] = a, b
        for i, r in enumerate(rows):
            for k in range(G):
                a = cur[i][k]
                if a is None:
                    used = [cur[q][k] for q in range(len(rows))
                            if cur[q][k] is not None]
                    a = used[0] if used else pad(r)
                out[r, k] = a
    return out


class EllStack:
    """Generators packed in one ELL pattern (the union of their non-zeros).

    ``mats`` has shape (M, n, n).  Slot 0 of row r is always column r (the
    diagonal); with ``partner`` slot 1 is always column r ^ 1.  ``layout``
    places the other non-zeros of the union:

    * ``"plain"``: slots 2.. in column order, gathered one by one;
    * ``"quad"`` (n <= 64): slots 2, 3 are the other pair of r's quad
      (``quad_columns``), the columns outside the quad follow;
    * ``"pairs"`` (n <= 64): slots 2 + 2k, 3 + 2k are the two columns
      2p, 2p + 1 of the row's k-th fetched pair (``fetch_pairs``).

    In the last two every row of the padded table follows the layout, and
    the kernels read such slot pairs whole (the quad's by DPP, the pairs
    by one 16-byte LDS read each).  With
    ``bank_aware`` the fetched addresses are ordered over the slots so that
    lanes read by one LDS cycle meet fewer bank conflicts.  ``col`` is int32
    (slots, NP), ``val`` float64 (M, slots, NP) with NP = 64 ceil(n / 64);
    unused slots point at the row's own pair with value 0, rows >= n are
    all zero.
    """

    def __init__(self, mats, partner=True, layout="plain", bank_aware=False):
        mats = np.asarray(mats, dtype=_f64)
        if mats.ndim != 3 or mats.shape[1] != mats.shape[2]:
            raise ValueError(f"generator stack must be (M, n, n), got "
                             f"{mats.shape}")
        M, n, _ = mats.shape
        if n > _lib.MAX_N_WIDE:
            raise ValueError(f"superoperator dimension {n} > "
                             f"{_lib.MAX_N_WIDE}")
        if layout not in ("plain", "quad", "pairs"):
            raise ValueError(f"unknown ELL layout {layout!r}")
        NP = padded_rows(n)
        if layout != "plain" and (not partner or NP != 64):
            raise ValueError(f"the {layout} layout needs partner=True and "
                             "n <= 64")
        union = np.any(mats != 0, axis=0)
        lanes = np.arange(NP, dtype=np.int64)
        fixed = 2 if partner else 1
        if layout == "pairs":
            pairs = fetch_pairs(union) + [[] for _ in range(NP - n)]
            G = max((len(p) for p in pairs), default=0)
            if bank_aware:
                fp = _bank_slots(pairs, G, lambda r: r >> 1, 8, 8)
            else:
                fp = np.array([p + [r >> 1] * (G - len(p))
                               for r, p in enumerate(pairs)],
                              dtype=np.int64).reshape(NP, G)
            slot_cols = np.empty((NP, 2 * G), dtype=np.int64)
            slot_cols[:, 0::2] = 2 * fp
            slot_cols[:, 1::2] = 2 * fp + 1
            used = [fixed + 2 * len(p) for p in pairs[:n]]
        else:
            head = []
            singles = []
            for r in range(NP):
                special = {r, r ^ 1} if partner else {r}
                if layout == "quad":
                    q = quad_columns(r)
                    special |= set(q)
                    head.append([q[1], q[2]])
                cols = [] if r >= n else \
                    [int(c) for c in np.nonzero(union[r])[0]
                     if c not in special]
                singles.append(cols)
            S = max((len(c) for c in singles[:n]), default=0)
            if bank_aware and layout == "quad":
                sc = _bank_slots(singles, S, lambda r: r, 16, 16)
            else:
                sc = np.array([c + [r] * (S - len(c))
                               for r, c in enumerate(singles)],
                              dtype=np.int64).reshape(NP, S)
            slot_cols = np.concatenate([np.array(head, dtype=np.int64)
                                        .reshape(NP, -1), sc], axis=1) \
                if layout == "quad" else sc
            used = [fixed + (2 if layout == "quad" else 0) + len(c)
                    for c in singles[:n]]
        slots = fixed + slot_cols.shape[1]
        col = np.empty((slots, NP), dtype=np.int32)
        col[0] = lanes
        if partner:
            col[1] = lanes ^ 1
        col[fixed:] = slot_cols.T
        val = np.zeros((M, slots, NP), dtype=_f64)
        for r in range(n):
            val[:, 0, r] = mats[:, r, r]
            if partner and (r ^ 1) < n:
                val[:, 1, r] = mats[:, r, r ^ 1]
            seen = {r, r ^ 1} if partner else {r}
            for sl in range(fixed, slots):
                c = int(col[sl, r])
                # each non-zero once: padding repeats read value 0
                if c < n and c not in seen and union[r, c]:
                    val[:, sl, r] = mats[:, r, c]
                    seen.add(c)
        # slot 0 is the diagonal by definition; its column entry carries the
        # number of used slots of each row: the row's non-zeros lie in the
        # slots below it (a bank-aware order may interleave padding, so the
        # fetch layouts then count every slot)
        if bank_aware and layout != "plain":
            used = [slots] * n
        col[0, :] = fixed
        col[0, :n] = used
        self.n = n
        self.npad = NP
        self.slots = slots
        self.partner = bool(partner)
        self.layout = layout
        self.col = col
        self.val = val
        self.nnz = int(union.sum())
        #: LDS reads per product beside the write (one per fetched pair)
        self.gather_slots = {"plain": slots - fixed,
                             "quad": slots - fixed - 1,
                             "pairs": (slots - fixed) // 2}[layout]
        self.norms = np.abs(mats).sum(axis=1).max(axis=1)  # 1-norms


# ---------------------------------------------------------------------------
# Taylor degree and substeps
# ---------------------------------------------------------------------------

def taylor_remainder(x, m):
    """Bound on sum_{k>m} x^k / k! for x >= 0."""
    if x == 0:
        return 0.0
    if x >= m + 2:
        return math.inf
    return x ** (m + 1) / math.factorial(m + 1) / (1 - x / (m + 2))


def coefficient_samples(cmin, cmax, count=48, seed=0):
    """Node-coefficient samples (count, S, K) covering the box [cmin, cmax].

    Corners of the varying coordinates (up to 2^8), the centre and uniform
    random points; constant coordinates (cmin == cmax) are fixed.
    """
    cmin = np.asarray(cmin, dtype=_f64)
    cmax = np.asarray(cmax, dtype=_f64)
    vary = np.nonzero((cmax - cmin).ravel() != 0)[0]
    rng = np.random.default_rng(seed)
    out = [0.5 * (cmin + cmax)]
    if vary.size <= 8:
        for bits in range(2 ** vary.size):
            c = cmin.copy().ravel()
            for k, v in enumerate(vary):
                if (bits >> k) & 1:
                    c[v] = cmax.ravel()[v]
            out.append(c.reshape(cmin.shape))
    for _ in range(count):
        out.append(cmin + (cmax - cmin) * rng.random(cmin.shape))
    return np.array(out)


def taylor_plan(norm, tol=1e-17, max_taylor=_lib.MAX_TAYLOR):
    """Cheapest (m, s) with s * remainder(norm / s, m) <= tol.

    ``norm`` bounds the 1-norm of every exponent A; the action
    exp(A) x = (exp(A/s))^s x is evaluated with degree-m Taylor polynomials.
    Cost is m * s sparse products; ties prefer fewer substeps.
    """
    best = None
    s = 1
    while s <= 1024:
        x = norm / s
        for m in range(1, max_taylor + 1):
            if s * taylor_remainder(x, m) <= tol:
                if best is None or m * s < best[0] * best[1]:
                    best = (m, s)
                break
        s *= 2
    if best is None:
        raise ValueError(f"no Taylor plan reaches tol={tol} for "
                         f"norm={norm}")
    return best


def taylor_degrees_measured(exponents, tols, safety=8.0, max_taylor=None):
    """Least degree m with one substep meeting each tolerance of ``tols``.

    One pass of the powers of the sampled exponents gives the degree
    ``taylor_plan_measured(exponents, tol, safety, max_taylor=mt,
    max_substeps=1)`` returns for every ``tol`` of ``tols`` (and ``mt`` of
    ``max_taylor``, a sequence, or ``_lib.MAX_TAYLOR`` for all).
    """
    tols = [float(t) for t in tols]
    if max_taylor is None:
        max_taylor = [_lib.MAX_TAYLOR] * len(tols)
    max_taylor = [int(m) for m in max_taylor]
    exps = [np.asarray(a, dtype=_f64) for a in exponents]
    norms = [float(np.abs(b).sum(axis=0).max()) for b in exps]
    powers = [b.copy() for b in exps]
    degrees = [None] * len(tols)
    fact = 1.0
    for m in range(1, max(max_taylor) + 1):
        fact *= (m + 1)
        worst = 0.0
        for k, b in enumerate(exps):
            powers[k] = powers[k] @ b          # B^(m+1), every sample
            if norms[k] >= m + 2:
                worst = math.inf
            elif worst < math.inf:
                pn = float(np.abs(powers[k]).sum(axis=0).max())
                worst = max(worst, pn / fact / (1 - norms[k] / (m + 2)))
        for i, tol in enumerate(tols):
            if degrees[i] is None and m <= max_taylor[i] and \
                    safety * worst <= tol:
                degrees[i] = m
        if all(d is not None for d in degrees):
            break
    for i, d in enumerate(degrees):
        if d is None:
            raise ValueError(f"no Taylor plan reaches tol={tols[i]}")
    return degrees


def taylor_plan_measured(exponents, tol=1e-17, safety=8.0,
                         max_taylor=_lib.MAX_TAYLOR, max_substeps=1024):
    """(m, s) from exact powers of representative exponents A.

    For each candidate substep count s and degree m the truncation error of
    (T_m(A/s))^s x is bounded by
        s * ||B^(m+1)||_1 / (m+1)! / (1 - ||B||_1 / (m+2)),  B = A/s,
    which is rigorous for the sampled A (||B^(m+1+i)|| <= ||B^(m+1)||
    ||B||^i); ``safety`` covers coefficients between the samples.  Returns
    the cheapest (m * s) plan meeting ``tol`` with s <= ``max_substeps``.
    """
    exps = [np.asarray(a, dtype=_f64) for a in exponents]
    best = None
    s = 1
    while s <= max_substeps:
        scaled = [a / s for a in exps]
        norms = [float(np.abs(b).sum(axis=0).max()) for b in scaled]
        powers = [b.copy() for b in scaled]
        fact = 1.0
        for m in range(1, max_taylor + 1):
            fact *= (m + 1)
            worst = 0.0
            for k, b in enumerate(scaled):
                powers[k] = powers[k] @ b          # B^(m+1), every sample
                if norms[k] >= m + 2:
                    worst = math.inf
                elif worst < math.inf:
                    pn = float(np.abs(powers[k]).sum(axis=0).max())
                    worst = max(worst, pn / fact / (1 - norms[k] / (m + 2)))
            if s * safety * worst <= tol:
                if best is None or m * s < best[0] * best[1]:
                    best = (m, s)
                break
        s *= 2
        if best is not None and best[0] * best[1] <= s:
            break
    if best is None:
        raise ValueError(f"no Taylor plan reaches tol={tol}")
    return best


# ---------------------------------------------------------------------------
# Rotating-frame precompute (restates generate_simulator.py:1567-1677)
# ---------------------------------------------------------------------------

def rotating_tables(generators, doubles, singles, nodes, time_step, fine,
                    residual=True):
    """Interaction-picture generators per node and exp(L0 h) step factors.

    For node offset tau_s = (c_s dt) / F the generators are conjugated
    ``exp(-L0 tau) G exp(L0 tau)`` in the real block eigenbasis of L0
    (2x2 blocks for eigenvalue pairs a +- ib, 1x1 for real eigenvalues).
    Returns ``(gens_rot (S, G, n, n), d_step (nd, 2), s_step (ns,))`` where
    the step factors describe exp(L0 h) - I when ``residual`` (the
    reference's expm1 forms, :1617-1624) and exp(L0 h) otherwise.
    """
    generators = np.asarray(generators, dtype=_f64)
    nd = doubles.shape[0]
    ns = singles.size
    tz = nodes * time_step / fine
    h = time_step / fine
    a = doubles[:, 0] if nd else np.zeros(0)
    b = doubles[:, 1] if nd else np.zeros(0)
    out = np.empty((nodes.size,) + generators.shape, dtype=_f64)
    for s, tau in enumerate(tz):
        amp = np.exp(-a * tau)
        att = np.exp(a * tau)
        cos = np.cos(b * tau)
        sin = np.sin(b * tau)
        db0, db1 = amp * cos, amp * sin
        df0, df1 = att * cos, -att * sin
        g = generators.copy()
        if nd:
            r0 = g[:, 0:2 * nd:2, :].copy()
            r1 = g[:, 1:2 * nd:2, :].copy()
            g[:, 0:2 * nd:2, :] = db0[None, :, None] * r0 \
                - db1[None, :, None] * r1
            g[:, 1:2 * nd:2, :] = db1[None, :, None] * r0 \
                + db0[None, :, None] * r1
        if ns:
            g[:, 2 * nd:2 * nd + ns, :] *= np.exp(-singles * tau)[None, :,
                                                                   None]
        if nd:
            c0 = g[:, :, 0:2 * nd:2].copy()
            c1 = g[:, :, 1:2 * nd:2].copy()
            tr = df0[None, None, :] * c0 + df1[None, None, :] * c1
            ti = df1[None, None, :] * c0 - df0[None, None, :] * c1
            g[:, :, 0:2 * nd:2] = tr
            g[:, :, 1:2 * nd:2] = -ti
        if ns:
            g[:, :, 2 * nd:2 * nd + ns] *= np.exp(singles * tau)[None, None,
                                                                  :]
        out[s] = g
    if residual:
        s_step = np.expm1(singles * h)
        sin2 = -2 * (np.sin(b * h / 2) ** 2)
        sin1 = np.sin(b * h)
        expmr = np.expm1(a * h)
        d_step = np.stack([(1 + expmr) * sin2 + expmr, -(1 + expmr) * sin1],
                          axis=1) if nd else np.zeros((0, 2))
    else:
        s_step = np.exp(singles * h)
        att = np.exp(a * h)
        d_step = np.stack([att * np.cos(b * h), -att * np.sin(b * h)],
                          axis=1) if nd else np.zeros((0, 2))
    return out, d_step, s_step


def eig_lane_factors(d_step, s_step, n, residual=True):
    """Per-row (own, partner) factors so that x <- x + own x + part x[r^1].

    With ``residual`` the step factors are exp(L0 h) - I (expm1 forms);
    otherwise they are exp(L0 h) and 1 is subtracted here.
    """
    own = np.zeros(padded_rows(n), dtype=_f64)
    part = np.zeros(padded_rows(n), dtype=_f64)
    nd = d_step.shape[0]
    shift = 0.0 if residual else 1.0
    for y in range(nd):
        d0, d1 = d_step[y]
        own[2 * y] = d0 - shift
        part[2 * y] = -d1
        own[2 * y + 1] = d0 - shift
        part[2 * y + 1] = d1
    for y in range(s_step.size):
        own[2 * nd + y] = s_step[y] - shift
    if 2 * nd + s_step.size != n:
        raise ValueError("eigenbasis size does not match the generators")
    return np.concatenate([own, part])


# ---------------------------------------------------------------------------
# Launch
# ---------------------------------------------------------------------------

def lds_instructions(pattern, layout):
    """LDS instructions per Taylor product of a row layout on ``pattern``
    (n <= 64, already in the layout's basis order): the write of the lane's
    component plus one read per gathered column or fetched pair."""
    if layout == "plain":
        return 1 + EllStack(pattern[None].astype(_f64)).gather_slots
    if layout == "quad":
        return 1 + quad_gathers(pattern)
    return 1 + pair_fetches(pattern)


def choose_layout(pattern, layout="auto", pair_units=False):
    """(order, layout) of the ELL table for an n <= 64 pattern.

    ``"plain"`` keeps the order.  ``"quad"`` and ``"pairs"`` return their
    searched order (``quad_order``, ``pair_fetch_order``; with
    ``pair_units`` the rows (2k, 2k + 1) stay together, and the pair layout
    keeps the order).  ``"auto"`` takes the quad layout if it needs fewer
    LDS instructions per product than the plain one, else the plain one.
    Measured on the 7-level NV sweep (4096 instances, same process, unrolled
    degree-11 chain, bank-aware slots): n = 49 plain 476, quad 430, pairs
    477 ms; n = 19 plain 436, quad 409, pairs 412 ms.  The pair layout has
    the fewest instructions, but its scattered 16-byte reads meet more bank
    conflicts when every lane is active.
    """
    n = pattern.shape[0]
    ident = np.arange(n, dtype=np.int64)
    if layout == "plain" or n < 2:
        return ident, "plain"
    if layout == "auto":
        order, _ = choose_layout(pattern, "quad", pair_units)
        better = lds_instructions(pattern[order][:, order], "quad") < \
            lds_instructions(pattern, "plain")
        return (order, "quad") if better else (ident, "plain")
    if layout == "quad":
        return quad_order(pattern, pair_units=pair_units), "quad"
    if layout == "pairs":
        order = ident if pair_units else pair_fetch_order(pattern)
        return order, "pairs"
    raise ValueError(f"unknown ELL layout {layout!r}")


class Propagator:
    """Device copy of one generator stack and its CF scheme.

    ``mats`` is (K, n, n) (non-rotating) or (S, K, n, n) (rotating frame,
    node-wise generators).  ``weights`` (J, S) and ``nodes`` (S,) are the CF
    tables.  ``eig`` is the (128,) lane factor array of the rotating
    eigen-step, or None.  ``layout`` (n <= 64) picks the ELL layout and
    the basis order behind it (``choose_layout``, ``EllStack``);
    ``bank_aware`` orders fetched addresses over the slots against LDS bank
    conflicts.
    """

    def __init__(self, mats, weights, nodes, rotating=False, eig=None,
                 permute=None, layout="auto", bank_aware=True):
        torch = _torch()
        _lib.load()
        mats = np.asarray(mats, dtype=_f64)
        self.rotating = bool(rotating)
        weights = np.asarray(weights, dtype=_f64)
        nodes = np.asarray(nodes, dtype=_f64)
        self.J, self.S = weights.shape
        if nodes.size != self.S:
            raise ValueError("nodes and weights disagree on S")
        if self.J > _lib.MAX_EXP or self.S > _lib.MAX_NODE:
            raise ValueError(f"CF scheme J={self.J}, S={self.S} beyond "
                             "library limits")
        if self.rotating:
            if mats.ndim != 4 or mats.shape[0] != self.S:
                raise ValueError("rotating generators must be (S, K, n, n)")
            self.K = mats.shape[1]
            flat = mats.reshape(-1, mats.shape[-2], mats.shape[-1])
        else:
            if mats.ndim != 3:
                raise ValueError("generators must be (K, n, n)")
            self.K = mats.shape[0]
            flat = mats
        if self.K > _lib.MAX_GEN:
            raise ValueError(f"{self.K} generators > {_lib.MAX_GEN}")
        # The rotating eigenbasis already holds its 2x2 blocks at
        # (2k, 2k + 1); other bases are reordered so that coupled pairs are.
        if permute is None:
            permute = not self.rotating
        n0 = flat.shape[-1]
        if permute:
            perm = pair_permutation(np.any(flat != 0, axis=0))
        else:
            perm = np.arange(n0)
        self.layout = "plain"
        pat = np.any(flat != 0, axis=0)[perm][:, perm]
        if n0 > _lib.MAX_N:
            # several wavefronts per instance: widest pairs first, so that
            # each wavefront reads only the slots its own rows use
            perm = perm[width_order(pat)]
        elif layout != "plain":
            order, self.layout = choose_layout(pat, layout,
                                               pair_units=not permute)
            perm = perm[order]
        if eig is not None and not np.array_equal(perm, np.arange(n0)):
            eig = np.asarray(eig, dtype=_f64)
            half = eig.size // 2
            own, part = eig[:half].copy(), eig[half:].copy()
            own[:n0], part[:n0] = own[perm], part[perm]
            eig = np.concatenate([own, part])
        self.perm = perm
        self.inv_perm = np.argsort(perm)
        flat = flat[:, perm][:, :, perm]
        self.ell = EllStack(flat, partner=True, layout=self.layout,
                            bank_aware=bank_aware)
        self.n = self.ell.n
        norms = self.ell.norms
        self.gen_norms = norms.reshape(self.S, self.K) if self.rotating \
            else norms
        self.weights = weights
        self.nodes = nodes
        self._dense = mats.copy()
        self.d_perm = device_tensor(perm)
        self.d_inv_perm = device_tensor(self.inv_perm)
        self.d_col = device_tensor(self.ell.col)
        self.d_val = device_tensor(self.ell.val)
        self.d_w = device_tensor(weights)
        self.d_nodes = device_tensor(nodes)
        self.d_eig = device_tensor(eig) if eig is not None else None
        self.torch = torch

    def sample_exponents(self, coef_samples, h):
        """Dense A_j for node-coefficient samples (count, S, K), all j."""
        out = []
        mats = self._dense
        for c in np.asarray(coef_samples, dtype=_f64):
            for j in range(self.J):
                if self.rotating:
                    f = c * (h * self.weights[j][:, None])     # (S, K)
                    out.append(np.tensordot(f, mats, axes=([0, 1], [0, 1])))
                else:
                    cw = self.weights[j] @ c                   # (K,)
                    out.append(h * np.tensordot(cw, mats, axes=(0, 0)))
        return out

    def plan(self, coef_samples, h, tol=1e-17, max_substeps=1024,
             max_taylor=_lib.MAX_TAYLOR):
        """Taylor plan from sampled exponents (see taylor_plan_measured)."""
        return taylor_plan_measured(self.sample_exponents(coef_samples, h),
                                    tol=tol, max_substeps=max_substeps,
                                    max_taylor=max_taylor)

    def plan_degrees(self, coef_samples, h, tols, max_taylor=None):
        """One-substep Taylor degrees for several tolerances from one set
        of sampled exponents (see taylor_degrees_measured)."""
        return taylor_degrees_measured(self.sample_exponents(coef_samples, h),
                                       tols, max_taylor=max_taylor)

    def norm_bound(self, coef_abs_max, h):
        """Bound on ||A_j||_1 from per-node maxima |c[s, g]| (S, K)."""
        c = np.asarray(coef_abs_max, dtype=_f64).reshape(self.S, self.K)
        best = 0.0
        for j in range(self.J):
            if self.rotating:
                v = np.sum(np.abs(h * self.weights[j][:, None]) * c
                           * self.gen_norms)
            else:
                cw = np.abs(self.weights[j]) @ c
                v = h * float(np.sum(cw * self.gen_norms))
            best = max(best, float(v))
        return best

    def run(self, *, n_inst, fine, n_out, h, t0, dt, rho0, coef=None,
            coef_inst_stride=0, waves=None, segments=None, taylor=None,
            substeps=1, read="full", obs=None, obs_offset=None, out=None,
            state_out=None, ckpt_every=0, first_out=0, stream=None,
            store_inputs=False):
        """Launch one propagation; returns the output tensor.

        ``rho0`` (n_inst, n) device tensor.  Coefficients: ``coef`` device
        table (..., n_out*fine*S*K) or ``waves``/``segments`` device buffers
        of ``ssmx_wave``/``ssmx_segment`` bytes.  ``read`` is "full",
        "obs" or "integral".  ``state_out`` (device, float64) receives the
        final state, or with ``ckpt_every = K > 0`` the state before every
        K-th step, (n_inst, ceil(n_out * fine / K), n); both in the
        propagator's internal basis (``perm`` order).
        """
        torch = self.torch
        flags = _lib.SSMX_ROTATING if self.rotating else 0
        n_obs = 0
        if read == "full":
            flags |= _lib.SSMX_READ_FULL
            shape = (n_inst, n_out, self.n)
        else:
            if obs is None:
                raise ValueError("obs readout needs obs rows")
            n_obs = obs.shape[0]
            if read == "integral":
                flags |= _lib.SSMX_READ_INTEGRAL
                shape = (n_inst, n_obs)
            elif read == "obs":
                shape = (n_inst, n_out, n_obs)
            else:
                raise ValueError(f"unknown readout {read!r}")
        if waves is not None:
            flags |= _lib.SSMX_COEF_WAVE
        if store_inputs:
            # every exponential's input into state_out, for the adjoint
            flags |= _lib.SSMX_STORE_INPUTS
            want = (n_inst, n_out * fine, self.J, _lib.MAX_N)
            if state_out is None or tuple(state_out.shape) != want:
                raise ValueError(f"stored inputs need a state_out buffer of "
                                 f"shape {want}")
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device="cuda")
        if taylor is None:
            raise ValueError("taylor degree required")
        desc = _lib.PropagateDesc(
            n=self.n, n_inst=n_inst, n_gen=self.K, slots=self.ell.slots,
            n_exp=self.J, n_node=self.S, fine=fine, n_out=n_out,
            taylor=int(taylor), substeps=int(substeps), n_obs=max(n_obs, 1),
            flags=flags, first_out=int(first_out),
            partner=1 if self.ell.partner else 0,
            ckpt_every=int(ckpt_every), reserved=0,
            coef_inst_stride=int(coef_inst_stride), h=float(h),
            t0=float(t0), dt=float(dt))
        obs_rows = None
        if obs is not None:
            rows = np.zeros((n_obs, self.ell.npad), dtype=_f64)
            rows[:, :self.n] = np.asarray(obs)[:, self.perm]
            obs_rows = device_tensor(rows)
        rho0 = rho0[:, self.d_perm].contiguous()
        if ckpt_every and state_out is None:
            raise ValueError("checkpoints need a state_out buffer")
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        lib = _lib.load()
        rc = lib.ssmx_propagate(
            ctypes.byref(desc), _ptr(self.d_col), _ptr(self.d_val),
            _ptr(self.d_w), _ptr(self.d_nodes),
            _ptr(waves if waves is not None else coef), _ptr(segments),
            _ptr(rho0), _ptr(self.d_eig), _ptr(obs_rows), _ptr(obs_offset),
            _ptr(out), _ptr(state_out), ctypes.c_void_p(stream))
        _lib.check(rc)
        # keep the readout rows alive until the stream has consumed them
        self._keep = (obs_rows, rho0)
        if read == "full":
            out = out[..., self.d_inv_perm]
        return out
