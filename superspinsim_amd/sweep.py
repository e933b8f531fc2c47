"""Batched sweeps: many instances of one model in one kernel launch.

The reference simulates one sweep point per ``simspins`` call, each with its
own numba-compiled coefficient closure (generate_generators.py:2325-2339).
Here every instance carries a small waveform descriptor instead, evaluated on
the device at the same node times, so a whole ODMR sweep or protocol batch is
one launch.

A waveform assigns each generator coefficient

    c_g(t) = base_g + sum_k [t_start_k <= t < t_end_k] amp_k cos(omega_k t
                                                                 + phase_k)

which covers rendered microwave tones (``amp cos(omega t)``, the form the
reference's examples use), piecewise-constant amplitude tables
(``omega = 0``) and gated optical excitation.
"""

import ctypes
import math

import numpy as np

from superspinsim_amd import _lib, quadratures
from superspinsim_amd.propagator import (
    Propagator, _ptr, coefficient_samples, device_tensor, eig_lane_factors,
    rotating_tables)

_f64 = np.float64


_NO_ROWS = np.zeros((0, 5), dtype=_f64)
_NO_PARAMS = np.zeros(0, dtype=np.int64)
_NO_DPARAMS = np.zeros(0, dtype=_f64)


class Waveform:
    """Coefficient program of one instance for ``n_gen`` generators.

    Each generator's segments are held as arrays: rows (t_start, t_end,
    amp, omega, phase) in insertion order and, per row, a gradient tag
    (parameter index or -1, d amp / d parameter).  ``segments`` and
    ``params`` show them as tuples of tuples; change them through ``add``,
    ``add_segments``, ``set_segments`` or ``copy_segments``.
    """

    def __init__(self, n_gen, base=None):
        if n_gen > _lib.MAX_GEN:
            raise ValueError(f"{n_gen} generators > {_lib.MAX_GEN}")
        self.n_gen = n_gen
        self.base = np.zeros(n_gen, dtype=_f64)
        if base is not None:
            self.base[:] = base
        # merged arrays per generator (never changed in place, so copies
        # may share them) and rows added one at a time since the last merge
        self._rows = [_NO_ROWS] * n_gen
        self._param = [_NO_PARAMS] * n_gen
        self._dparam = [_NO_DPARAMS] * n_gen
        self._pending = [[] for _ in range(n_gen)]
        self._amp_max = np.zeros(n_gen, dtype=_f64)
        self._views = None

    def _changed(self, g, amp_max):
        self._amp_max[g] = amp_max
        self._views = None

    def add(self, g, t_start, t_end, amp, omega=0.0, phase=0.0, param=-1,
            dparam=1.0):
        """Add ``amp cos(omega t + phase)`` on [t_start, t_end) to c_g.

        ``param``/``dparam`` tag the segment for gradients: its ``amp``
        depends on parameter ``param`` with derivative ``dparam``.
        """
        self._pending[g].append((float(t_start), float(t_end), float(amp),
                                 float(omega), float(phase), int(param),
                                 float(dparam)))
        self._changed(g, max(self._amp_max[g], abs(float(amp))))
        return self

    def add_segments(self, g, t_start, t_end, amp, omega=0.0, phase=0.0,
                     param=-1, dparam=1.0):
        """``add`` for many segments of c_g at once: the arguments are
        arrays (or scalars) broadcast to one length, segment k taking
        element k of each."""
        cols = np.broadcast_arrays(
            *(np.asarray(v, dtype=_f64)
              for v in (t_start, t_end, amp, omega, phase, dparam)),
            np.asarray(param, dtype=np.int64))
        rows, param_g, dparam_g = self.segment_arrays(g)
        add = np.stack([c.ravel() for c in cols[:5]], axis=1)
        self._rows[g] = np.concatenate([rows, add])
        self._param[g] = np.concatenate([param_g, cols[6].ravel()])
        self._dparam[g] = np.concatenate([dparam_g, cols[5].ravel()])
        a = float(np.max(np.abs(add[:, 2]))) if add.shape[0] else 0.0
        self._changed(g, max(self._amp_max[g], a))
        return self

    def set_segments(self, g, segments, params=None):
        """Replace c_g's segments by ``segments`` ((m, 5) rows) and their
        gradient tags by ``params`` ((m, 2) rows, untagged by default)."""
        rows = np.array(segments, dtype=_f64).reshape(-1, 5)
        if params is None:
            param = np.full(rows.shape[0], -1, dtype=np.int64)
            dparam = np.ones(rows.shape[0], dtype=_f64)
        else:
            tags = list(params)
            if len(tags) != rows.shape[0]:
                raise ValueError("one gradient tag per segment")
            param = np.array([int(t[0]) for t in tags], dtype=np.int64)
            dparam = np.array([float(t[1]) for t in tags], dtype=_f64)
        self._rows[g], self._param[g], self._dparam[g] = rows, param, dparam
        self._pending[g] = []
        self._changed(g, float(np.max(np.abs(rows[:, 2])))
                      if rows.shape[0] else 0.0)
        return self

    def copy_segments(self, g, other, other_g):
        """c_g's segments and tags become those of ``other``'s c_other_g."""
        rows, param, dparam = other.segment_arrays(other_g)
        self._rows[g], self._param[g], self._dparam[g] = rows, param, dparam
        self._pending[g] = []
        self._changed(g, float(other._amplitude_max()[other_g]))
        return self

    def segment_arrays(self, g):
        """(rows (m, 5), parameter index (m,), d amp / d parameter (m,)) of
        c_g's segments in insertion order (shared; do not modify)."""
        pend = self._pending[g]
        if pend:
            arr = np.array(pend, dtype=_f64).reshape(-1, 7)
            self._rows[g] = np.concatenate([self._rows[g], arr[:, :5]])
            self._param[g] = np.concatenate(
                [self._param[g], np.array([r[5] for r in pend],
                                          dtype=np.int64)])
            self._dparam[g] = np.concatenate([self._dparam[g], arr[:, 6]])
            self._pending[g] = []
        return self._rows[g], self._param[g], self._dparam[g]

    def segment_count(self, g):
        return self._rows[g].shape[0] + len(self._pending[g])

    def _views_of(self):
        if self._views is None:
            segs, tags = [], []
            for g in range(self.n_gen):
                rows, param, dparam = self.segment_arrays(g)
                segs.append(tuple(map(tuple, rows.tolist())))
                tags.append(tuple(zip(param.tolist(), dparam.tolist())))
            self._views = (tuple(segs), tuple(tags))
        return self._views

    @property
    def segments(self):
        """Per generator, the (t_start, t_end, amp, omega, phase) tuples
        (immutable: change them through the methods above)."""
        return self._views_of()[0]

    @property
    def params(self):
        """Per generator and segment, (parameter index or -1, d amp / d
        parameter) (immutable)."""
        return self._views_of()[1]

    def _amplitude_max(self):
        """max |amp| per generator (0 without segments)."""
        return self._amp_max

    def tone(self, g, amp, frequency, t_start=-math.inf, t_end=math.inf,
             phase=0.0):
        """Rendered tone ``amp cos(tau f t + phase)`` (omega = tau * f)."""
        return self.add(g, t_start, t_end, amp, math.tau * frequency, phase)

    def __call__(self, t):
        """Python evaluation (same formula as the device)."""
        out = self.base.copy()
        for g, segs in enumerate(self.segments):
            for ts, te, amp, om, ph in segs:
                if ts <= t < te:
                    out[g] += amp * math.cos(om * t + ph)
        return out

    def bounds(self):
        """(lower, upper) bounds of c_g(t) over all t."""
        a = self._amplitude_max()
        return self.base - a, self.base + a

    def abs_max(self):
        """Upper bound of |c_g(t)| over all t (segments of one generator
        do not overlap)."""
        return np.abs(self.base) + self._amplitude_max()

    def shifted(self, offset):
        """The waveform seen from a clock started at ``offset``:
        ``shifted(T)(t) = self(t + T)`` up to rounding.  Segment windows move
        by -T and each tone's phase gains omega T (reduced mod 2 pi)."""
        out = Waveform(self.n_gen, self.base)
        for g in range(self.n_gen):
            rows, param, dparam = self.segment_arrays(g)
            moved = rows.copy()
            moved[:, 0] -= offset
            moved[:, 1] -= offset
            moved[:, 4] = [math.remainder(ph + om * offset, math.tau)
                           for om, ph in rows[:, 3:5].tolist()]
            out.set_segments(g, moved, zip(param.tolist(), dparam.tolist()))
        return out

    def sampler(self, rotating_offset=0):
        """Reference-style ``sampler(t, row)`` of this waveform."""
        def sample(t, row):
            row[rotating_offset:rotating_offset + self.n_gen] = self(t)
        return sample


WAVE_DTYPE = np.dtype([("base", _f64, _lib.MAX_GEN),
                       ("seg_begin", np.int32, _lib.MAX_GEN),
                       ("seg_count", np.int32, _lib.MAX_GEN)])


def waveform_records(waves, n_gen, with_tags=False):
    """Host arrays matching ``ssmx_wave[]`` and ``ssmx_segment[]``.

    Each generator's segments are stored in ascending (t_start, t_end, amp,
    omega, phase) order; overlapping windows are refused.  With
    ``with_tags`` also returns, per segment record, the parameter index (-1
    if untagged) and d amp / d parameter.  Every instance gets its own
    records (the adjoint accumulates per record).
    """
    recs = np.zeros(len(waves), dtype=WAVE_DTYPE)
    segs = []
    tags = []
    count = 0
    for i, w in enumerate(waves):
        if w.n_gen != n_gen:
            raise ValueError("waveforms must cover the same generators")
        recs["base"][i, :n_gen] = w.base
        for g in range(n_gen):
            recs["seg_begin"][i, g] = count
            arr, param, dparam = w.segment_arrays(g)
            nseg = arr.shape[0]
            recs["seg_count"][i, g] = nseg
            if not nseg:
                continue
            # lexicographic order of the rows (np.lexsort: last key first);
            # rows already in strictly ascending t_start keep their order
            if nseg > 1 and not np.all(arr[1:, 0] > arr[:-1, 0]):
                order = np.lexsort(arr.T[::-1])
                arr, param, dparam = arr[order], param[order], dparam[order]
            if np.any(arr[1:, 0] < arr[:-1, 1]):
                raise ValueError("overlapping segments on generator "
                                 f"{g} of instance {i}")
            segs.append(arr)
            if with_tags:
                tags.append((param, dparam))
            count += nseg
    if not segs:
        segs = [np.zeros((1, 5), dtype=_f64)]
        tags = [(np.array([-1], dtype=np.int64), np.zeros(1, dtype=_f64))]
    seg_arr = np.concatenate(segs)
    if not with_tags:
        return recs, seg_arr
    index = np.concatenate([t[0] for t in tags])
    scale = np.concatenate([t[1] for t in tags])
    return recs, seg_arr, index, scale


def pack_waveforms(waves, n_gen):
    """Device buffers (ssmx_wave array, ssmx_segment array) for instances."""
    recs, seg_arr = waveform_records(waves, n_gen)
    torch = __import__("torch")
    rec_bytes = np.frombuffer(recs.tobytes(), dtype=np.uint8).copy()
    d_waves = torch.from_numpy(rec_bytes).to("cuda")
    d_segs = device_tensor(seg_arr)
    return d_waves, d_segs


class BatchSimulator:
    """One model, many instances with their own waveforms.

    Parameters mirror ``generate_simulator``: ``generators`` (G, n, n) in the
    order of the coefficient rows, the CF scheme, fine divisions and the
    optional rotating frame (eigen-data from ``make_rotating``; generators
    then exclude L0).  ``layout`` is the propagators' ELL layout
    (``Propagator``).  ``pack`` lets ``run`` put several instances of a
    small model (n <= 32, lab frame) into one block-diagonal system per
    wavefront (``pack_factor``); an integer asks for that many per system
    (within the kernel's limits) whatever the instance count.
    """

    def __init__(self, generators, number_of_exponentials=2,
                 number_of_fine_divisions=1, use_rotating=False,
                 vectors_real=None, inv_vectors_real=None, doubles=None,
                 singles=None, tolerance=1e-16, layout="auto",
                 pack=True):
        self.gens = np.asarray(generators, dtype=_f64)
        self.layout = layout
        self.pack = pack if isinstance(pack, bool) else int(pack)
        self.G, self.n = self.gens.shape[0], self.gens.shape[1]
        self.nodes, self.weights, self.label = quadratures.scheme(
            number_of_exponentials)
        self.F = int(number_of_fine_divisions)
        self.rotating = bool(use_rotating)
        self.tolerance = tolerance
        if self.rotating:
            self.V = np.asarray(vectors_real, dtype=_f64)
            self.Vinv = np.asarray(inv_vectors_real, dtype=_f64)
            d = np.asarray(doubles, dtype=_f64)
            self.doubles = d.reshape((0, 2)) if d.shape[0] == 0 else d
            self.singles = np.asarray(singles, dtype=_f64)
        self._prop_cache = {}

    def _propagator(self, active, const, time_step):
        """Propagator over [folded constants] + varying generators.

        ``const`` maps generator -> its time-independent coefficient (same
        for every instance); those generators are summed into one matrix
        with coefficient 1, which the kernel assembles like any other.
        """
        key = (tuple(active), tuple(sorted(const.items())), float(time_step))
        if key not in self._prop_cache:
            varying = [g for g in active if g not in const]
            if self.rotating:
                mats, d_step, s_step = rotating_tables(
                    self.gens, self.doubles, self.singles, self.nodes,
                    time_step, self.F, True)
                eig = eig_lane_factors(d_step, s_step, self.n, True)
                parts = []
                if const:
                    parts.append(sum(v * mats[:, g] for g, v in
                                     const.items())[:, None])
                if varying:
                    parts.append(mats[:, varying])
                mats = np.concatenate(parts, axis=1)
            else:
                parts = []
                if const:
                    parts.append(sum(v * self.gens[g] for g, v in
                                     const.items())[None])
                if varying:
                    parts.append(self.gens[varying])
                mats = np.concatenate(parts, axis=0)
                eig = None
            self._prop_cache[key] = Propagator(
                mats, self.weights, self.nodes, rotating=self.rotating,
                eig=eig, layout=self.layout)
        return self._prop_cache[key]

    def _packed_propagator(self, active, const, time_step, k):
        """Propagator of k instances in one block-diagonal system.

        Block j (rows j n .. j n + n - 1) is instance j of the pack.  The
        folded constant matrix C becomes diag(C, .., C) with coefficient 1;
        each varying generator G becomes k matrices, G in block j only,
        whose coefficient is member j's.  A row of block j therefore sums
        exactly the terms of the single-instance row (the other blocks'
        matrices add exact zeros).
        """
        key = ("packed", k, tuple(active), tuple(sorted(const.items())),
               float(time_step))
        if key not in self._prop_cache:
            n = self.n
            varying = [g for g in active if g not in const]
            parts = []
            if const:
                c = sum(v * self.gens[g] for g, v in const.items())
                parts.append(np.kron(np.eye(k), c)[None])
            for g in varying:
                for j in range(k):
                    m = np.zeros((k * n, k * n))
                    m[j * n:(j + 1) * n, j * n:(j + 1) * n] = self.gens[g]
                    parts.append(m[None])
            self._prop_cache[key] = Propagator(
                np.concatenate(parts, axis=0), self.weights, self.nodes,
                layout=self.layout)
        return self._prop_cache[key]

    def pack_factor(self, n_inst, program, read="obs", n_obs=1):
        """Instances per block-diagonal system for one ``run``.

        A wavefront computes 64 rows whatever n is, so with n <= 32 several
        instances share one: k <= 64 / n, within the kernel's generator
        count (the constant block plus k copies of each varying generator)
        and readout rows (k n_obs <= 2; a full-state readout has no such
        limit).  k also keeps at least one system per SIMD (four per CU),
        below which packing would only idle SIMDs.  1 in the rotating
        frame or when ``pack`` is off.
        """
        if not self.pack or self.rotating or self.n > 32 or n_inst < 2:
            return 1
        n_const = 1 if None in program else 0
        n_var = len(program) - n_const
        J = self.weights.shape[0]
        k = _lib.MAX_N // self.n
        if n_var:
            # generators per system, and one coefficient batch of 64 lanes
            # holding J of them per step
            k = min(k, (_lib.MAX_GEN - n_const) // n_var,
                    (_lib.MAX_N // J - n_const) // n_var)
        if read != "full":
            k = min(k, _lib.MAX_OBS // max(n_obs, 1))
        if self.pack is True:
            torch = __import__("torch")
            simds = 4 * torch.cuda.get_device_properties(
                0).multi_processor_count
            k = min(k, max(1, n_inst // simds))
        else:
            k = min(k, int(self.pack))          # a fixed request
        return max(1, k)

    def program(self, waves):
        """(active, const, program) of a batch of waveforms.

        ``active`` lists generators with a non-zero coefficient somewhere,
        ``const`` maps generators whose coefficient is one constant for
        every instance to that constant, and ``program`` is the kernel's
        matrix list: None for the folded sum of constant generators
        (coefficient 1), an int for a varying generator.
        """
        amax = np.max([w.abs_max() for w in waves], axis=0)
        active = [g for g in range(self.G) if amax[g] != 0] or [0]
        const = {}
        for g in active:
            bases = {float(w.base[g]) for w in waves}
            if len(bases) == 1 and not any(w.segment_count(g)
                                           for w in waves):
                const[g] = bases.pop()
        program = ([None] if const else []) + \
            [g for g in active if g not in const]
        return active, const, program

    def coefficient_box(self, waves, program):
        """Node-coefficient samples (count, S, K) covering the batch."""
        lo = np.array([1.0 if g is None else
                       np.min([w.bounds()[0][g] for w in waves])
                       for g in program])
        hi = np.array([1.0 if g is None else
                       np.max([w.bounds()[1][g] for w in waves])
                       for g in program])
        S = self.nodes.size
        return coefficient_samples(np.broadcast_to(lo, (S, lo.size)),
                                   np.broadcast_to(hi, (S, hi.size)))

    def plan(self, waves, time_step):
        """(propagator, coefficient program, Taylor degree, substeps)."""
        active, const, program = self.program(waves)
        prop = self._propagator(active, const, time_step)
        h = time_step / self.F
        m, s = prop.plan(self.coefficient_box(waves, program), h,
                         tol=self.tolerance)
        return prop, program, m, s

    @staticmethod
    def kernel_waves(waves, program):
        """Per-instance waveforms over the kernel's generator list."""
        out = []
        for w in waves:
            base = [1.0 if g is None else w.base[g] for g in program]
            nw = Waveform(len(program), base)
            for k, g in enumerate(program):
                if g is not None:
                    nw.copy_segments(k, w, g)
            out.append(nw)
        return out

    def prepare(self, waves, rho0, time_start, time_end, time_step,
                obs=None, read="obs", taylor=None, substeps=None):
        """Plan, pack and stage one ``run`` on the device; returns a
        ``PreparedRun`` whose ``launch`` only enqueues the kernel."""
        N = int((time_end - time_start) / time_step)
        prop0, program, m, s = self.plan(waves, time_step)
        prop = prop0
        if taylor is not None:
            m = taylor
        if substeps is not None:
            s = substeps
        n_inst = len(waves)
        rho0 = np.atleast_2d(np.asarray(rho0, dtype=_f64))
        if rho0.shape[0] == 1 and n_inst > 1:
            rho0 = np.repeat(rho0, n_inst, axis=0)
        x0 = rho0 @ self.Vinv.T if self.rotating else rho0
        obs_eff = None
        if obs is not None:
            obs = np.atleast_2d(np.asarray(obs, dtype=_f64))
            obs_eff = obs @ self.V if self.rotating else obs
        n_obs = obs_eff.shape[0] if obs_eff is not None else 0
        k = self.pack_factor(n_inst, program, read, n_obs)
        kwaves = self.kernel_waves(waves, program)
        if k > 1:
            # systems of k consecutive instances; the last is filled up with
            # copies of the last instance, whose results are dropped
            n_sys = -(-n_inst // k)
            fill = [min(i, n_inst - 1) for i in range(n_sys * k)]
            active, const, _ = self.program(waves)
            prop = self._packed_propagator(active, const, time_step, k)
            n_const = 1 if None in program else 0
            kwaves = [self._pack_waves([kwaves[fill[q * k + j]]
                                        for j in range(k)], n_const)
                      for q in range(n_sys)]
            x0 = x0[fill].reshape(n_sys, k * self.n)
            if obs_eff is not None:
                blk = np.zeros((k * n_obs, k * self.n))
                for j in range(k):
                    blk[j * n_obs:(j + 1) * n_obs,
                        j * self.n:(j + 1) * self.n] = obs_eff
                obs_eff = blk
        d_waves, d_segs = pack_waveforms(kwaves, kwaves[0].n_gen)
        run = PreparedRun(self, prop, k, n_inst, read, N, m, s, time_start,
                          time_step, device_tensor(x0), d_waves, d_segs,
                          obs_eff)
        self.last_plan = dict(taylor=m, substeps=s, slots=prop.ell.slots,
                              nnz=prop0.ell.nnz, active=program, n_out=N,
                              steps=N * self.F, n=self.n, packed=k)
        return run

    @staticmethod
    def _pack_waves(members, n_const):
        """Kernel waveform of one block-diagonal system: the constant slot,
        then for each varying generator one slot per member."""
        k = len(members)
        n_var = members[0].n_gen - n_const
        w = Waveform(n_const + k * n_var)
        if n_const:
            w.base[0] = 1.0
        for v in range(n_var):
            for j, mw in enumerate(members):
                g, slot = n_const + v, n_const + v * k + j
                w.base[slot] = mw.base[g]
                w.copy_segments(slot, mw, g)
        return w

    def run(self, waves, rho0, time_start, time_end, time_step, obs=None,
            read="obs", taylor=None, substeps=None):
        """Propagate every instance; returns a torch tensor on the device.

        ``rho0``: (n,) or (n_inst, n) vectorised initial states (standard
        basis).  ``obs``: (n_obs, n) readout rows in the standard basis;
        ``read`` is "obs" (every output), "integral" (trapezoid over all
        propagator steps) or "full" (state vectors, standard basis).
        """
        run = self.prepare(waves, rho0, time_start, time_end, time_step,
                           obs=obs, read=read, taylor=taylor,
                           substeps=substeps)
        run.launch()
        self._keep = run
        return run.result()

    # ------------------------------------------------------------------
    # time-chunked runs (small systems, few instances)

    def chunk_plan(self, n_inst, N, resident=None):
        """(chunks P, outputs per chunk L, pack k) of a chunked run.

        Every instance's N outputs are cut into P chunks of L; each chunk
        propagates the n basis vectors, k per wavefront.  P minimises the
        launch's length in wavefront rounds times chunk length,
        ceil(waves / resident) L, smallest P first.
        """
        if self.rotating or self.n > 32:
            raise ValueError("chunked runs need the lab frame and n <= 32")
        k = max(1, min(_lib.MAX_N // self.n, _lib.MAX_GEN - 1))
        if resident is None:
            torch = __import__("torch")
            resident = 16 * torch.cuda.get_device_properties(
                0).multi_processor_count
        best = None
        for P in range(1, N + 1):
            L = -(-N // P)
            if (P - 1) * L >= N:        # the same split as a smaller P
                continue
            waves = -(-(n_inst * P * self.n) // k)
            cost = -(-waves // resident) * L
            if best is None or cost < best[0]:
                best = (cost, P, L)
        return best[1], best[2], k

    def prepare_chunked(self, waves, rho0, time_start, time_end, time_step,
                        obs=None, read="obs", chunks=None):
        """``prepare`` for a run cut into time chunks (``ChunkedRun``).

        Each instance's outputs are split into P chunks; every chunk
        propagates the n basis vectors from the chunk's start (waveforms
        time-shifted to it, ``Waveform.shifted``), all chunks of all
        instances in one launch, and ``ssmx_chunk_combine`` chains the chunk
        propagators from rho0.  Same readouts as ``run`` ("obs" or "full");
        the node times of a chunk are evaluated from the chunk's start, so
        results equal ``run``'s to rounding, not bit for bit.
        """
        if read not in ("obs", "full"):
            raise ValueError("chunked runs read out 'obs' or 'full'")
        N = int((time_end - time_start) / time_step)
        n_inst = len(waves)
        if chunks is None:
            P, L, k = self.chunk_plan(n_inst, N)
        else:
            P = max(1, min(int(chunks), N))
            L = -(-N // P)
            P = -(-N // L)
            k = max(1, min(_lib.MAX_N // self.n, _lib.MAX_GEN - 1))
        n = self.n
        members = []
        for w in waves:
            for p in range(P):
                shifted = w.shifted(time_start + p * L * time_step)
                members.extend([shifted] * n)
        basis = np.tile(np.eye(n), (n_inst * P, 1))
        saved = self.pack
        self.pack = k
        try:
            run = self.prepare(members, basis, 0.0,
                               L * time_step * (1 + 1e-12), time_step,
                               read="full")
        finally:
            self.pack = saved
        rho0 = np.atleast_2d(np.asarray(rho0, dtype=_f64))
        if rho0.shape[0] == 1 and n_inst > 1:
            rho0 = np.repeat(rho0, n_inst, axis=0)
        obs_rows = None
        if read == "obs":
            obs_rows = np.atleast_2d(np.asarray(obs, dtype=_f64))
        self.last_plan = dict(self.last_plan, n_out=N,
                              steps=N * self.F, chunks=P, chunk_out=L,
                              chunk_pack=run.k)
        return ChunkedRun(run, n_inst, N, P, L, device_tensor(rho0),
                          None if obs_rows is None else
                          device_tensor(obs_rows))

    def run_chunked(self, waves, rho0, time_start, time_end, time_step,
                    obs=None, read="obs", chunks=None):
        """``run`` by time chunks (see ``prepare_chunked``)."""
        run = self.prepare_chunked(waves, rho0, time_start, time_end,
                                   time_step, obs=obs, read=read,
                                   chunks=chunks)
        run.launch()
        self._keep = run
        return run.result()

    def step_flops(self, plan=None):
        """Algorithmic FP64 FLOPs of one propagator step of one instance.

        J exponentials, each: assembly 2 K nnz (+ h scaling n), and s m
        Taylor terms of one sparse product (2 nnz) plus scale and add
        (2 n); rotating frame adds the 2x2-block step (4 n).
        """
        p = plan or self.last_plan
        J = self.weights.shape[0]
        K = len(p["active"]) * (self.nodes.size if self.rotating else 1)
        nnz, n = p["nnz"], p["n"]
        per_exp = 2 * K * nnz + n + p["substeps"] * p["taylor"] * (
            2 * nnz + 2 * n)
        return J * per_exp + (4 * n if self.rotating else 0)


class ChunkedRun:
    """Device inputs of a time-chunked run (``prepare_chunked``): ``launch``
    propagates every chunk's basis vectors, then chains the chunks."""

    def __init__(self, chunk_run, n_inst, N, P, L, d_rho0, d_obs):
        torch = __import__("torch")
        self.chunk_run, self.n_inst, self.N = chunk_run, n_inst, N
        self.P, self.L = P, L
        self.d_rho0, self.d_obs = d_rho0, d_obs
        n = chunk_run.sim.n
        n_obs = 0 if d_obs is None else d_obs.shape[0]
        self.desc = _lib.ChunkDesc(n=n, n_inst=n_inst, n_chunks=P,
                                   chunk_out=L, n_out=N, pack=chunk_run.k,
                                   n_obs=n_obs, reserved=0)
        shape = (n_inst, N, n_obs if n_obs else n)
        self.out = torch.empty(shape, dtype=torch.float64, device="cuda")

    def launch(self):
        torch = __import__("torch")
        ops = self.chunk_run.launch()
        systems = -(-self.n_inst * self.P * self.desc.n // self.desc.pack)
        want = (systems, self.L, self.desc.pack * self.desc.n)
        if tuple(ops.shape) != want or not ops.is_contiguous():
            raise RuntimeError(f"chunk propagators {tuple(ops.shape)}, "
                               f"expected contiguous {want}")
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().ssmx_chunk_combine(
            ctypes.byref(self.desc), _ptr(ops), _ptr(self.d_rho0),
            _ptr(self.d_obs), _ptr(self.out), ctypes.c_void_p(stream)))
        self._keep = ops
        return self.out

    def result(self):
        return self.out


class PreparedRun:
    """Device inputs of one ``BatchSimulator.run`` (see ``prepare``):
    ``launch`` enqueues the kernel on the current stream, ``result`` maps
    the output of the last launch back to one row per instance (standard
    basis for full states)."""

    def __init__(self, sim, prop, k, n_inst, read, N, taylor, substeps, t0,
                 dt, d_x0, d_waves, d_segs, obs):
        torch = __import__("torch")
        self.sim, self.prop, self.k, self.n_inst = sim, prop, k, n_inst
        self.read, self.N = read, N
        self.args = dict(n_inst=d_x0.shape[0], fine=sim.F, n_out=N,
                         h=dt / sim.F, t0=t0, dt=dt, rho0=d_x0,
                         waves=d_waves, segments=d_segs, taylor=taylor,
                         substeps=substeps, read=read, obs=obs)
        n_obs = obs.shape[0] if obs is not None else 0
        shape = {"full": (d_x0.shape[0], N, prop.n),
                 "obs": (d_x0.shape[0], N, n_obs),
                 "integral": (d_x0.shape[0], n_obs)}[read]
        self.out = torch.empty(shape, dtype=torch.float64, device="cuda")
        self.last = None

    def launch(self):
        self.last = self.prop.run(out=self.out, **self.args)
        return self.last

    def result(self):
        out, k, n = self.last, self.k, self.sim.n
        if k > 1:
            S = out.shape[0]
            if self.read == "full":
                out = out.reshape(S, self.N, k, n).permute(0, 2, 1, 3)
                out = out.reshape(S * k, self.N, n)
            elif self.read == "obs":
                n_obs = out.shape[2] // k
                out = out.reshape(S, self.N, k, n_obs).permute(0, 2, 1, 3)
                out = out.reshape(S * k, self.N, n_obs)
            else:
                out = out.reshape(S * k, -1)
            out = out[:self.n_inst]
        if self.read == "full" and self.sim.rotating:
            torch = __import__("torch")
            out = out @ torch.from_numpy(self.sim.V.T.copy()).to(out.device)
        return out
