"""Gradients of integrated readouts through the HIP adjoint kernel.

A ``GradientProblem`` is a batch of instances of one parametrised model:

* ``generators(theta)`` returns the (G, n, n) generator stack in the order
  of the waveform coefficients (for the NV models ``[L0, Gx, Gy, Gz, Gr]``,
  generate_generators.py:2296-2304);
* ``waveforms(amplitudes)`` returns one ``Waveform`` per instance; segments
  whose ``amp`` depends on ``amplitudes[k]`` are added with ``param=k`` and
  ``dparam = d amp / d amplitudes[k]``;
* ``rho0``, ``obs`` and the time grid are those of ``BatchSimulator.run``.

The readout is the trapezoid integral over every propagator step,
``F[i, o] = sum_k w_k obs_o . x_ik`` (``read="integral"``), e.g. the
integrated fluorescence of an ODMR or pulsed protocol.  ``forward`` runs
``ssmx_propagate`` and keeps state checkpoints every ``checkpoint_every``
steps in HBM; ``backward`` runs ``ssmx_adjoint``, which recomputes each
segment from its checkpoint and returns, per instance, the gradient with
respect to the physical parameters and to every tagged segment amplitude.

Parameters may enter the generators whose coefficients are the same
constant for every instance (these are folded into one matrix) and the
tagged segment amplitudes; ``d generators / d theta`` is evaluated by
central differences of ``generators`` with step
``fd_rel * max(|theta_k|, theta_scale_k)``.  ``IntegratedReadout`` wraps
forward and backward as a ``torch.autograd.Function``.

The reference has no gradients; this differentiates the path that
``ssmx_propagate`` replaces (generate_simulator.py:1945-2116).
"""

import ctypes

import numpy as np

from superspinsim_amd import _lib
from superspinsim_amd.propagator import _bank_slots, _ptr, _torch, device_tensor
from superspinsim_amd.sweep import BatchSimulator, waveform_records

_f64 = np.float64


def shared_adjoint_ell(fwd_col, mats, dmats):
    """One ELL pattern for the adjoint's forward and transposed products.

    ``fwd_col`` is the forward kernel's column table (slots, 64), ``mats``
    (M, n, n) and ``dmats`` (n_theta, n, n) are in the same (permuted)
    basis.  Row r lists the forward columns of row r in the forward
    kernel's slot order (so a product on this pattern adds the same terms
    in the same order, plus exact zeros), then the other columns c with
    A[c][r] or D[c][r] non-zero.  Returns (col, val_f, val_t, dval_t) with
    val_f[m, s, r] = G_m[r][c], val_t[m, s, r] = G_m[c][r] and
    dval_t[t, s, r] = D_t[c][r] for c = col[s, r].
    """
    M, n, _ = mats.shape
    union = np.any(mats != 0, axis=0)
    extra = union.T.copy()
    if dmats.shape[0]:
        extra |= np.any(dmats != 0, axis=0).T
    rows = []
    for r in range(n):
        fwd = [int(c) for c in fwd_col[2:int(fwd_col[0, r]), r]]
        special = {r, r ^ 1}
        more = [int(c) for c in np.nonzero(extra[r])[0]
                if c not in special and c not in fwd]
        rows.append(fwd + more)
    slots = 2 + max(len(c) for c in rows)
    col = np.tile(np.arange(64, dtype=np.int32), (slots, 1))
    for r in range(n):
        col[1, r] = r ^ 1
        for k, c in enumerate(rows[r]):
            col[2 + k, r] = c
    val_f = np.zeros((M, slots, 64))
    val_t = np.zeros((M, slots, 64))
    dval_t = np.zeros((dmats.shape[0], slots, 64))
    for r in range(n):
        for sl in range(slots):
            c = int(col[sl, r])
            if c >= n or (sl > 1 and sl >= 2 + len(rows[r])):
                continue
            val_f[:, sl, r] = mats[:, r, c]
            val_t[:, sl, r] = mats[:, c, r]
            if dmats.shape[0]:
                dval_t[:, sl, r] = dmats[:, c, r]
    for r in range(n):
        col[0, r] = 2 + len(rows[r])
    return col, val_f, val_t, dval_t


def transposed_adjoint_ell(mats, dmats):
    """The adjoint chain's own ELL pattern (n <= 64).

    Row r lists r (slot 0), its partner r ^ 1 (slot 1, exchanged by DPP)
    and then every other column c with A[c][r] or D[c][r] non-zero, in an
    order chosen against LDS bank conflicts (``_bank_slots``: the chain
    gathers these columns one at a time, and the forward kernel's slot
    order, which the forward terms must keep, does not constrain it).  A
    padding slot may repeat another column of the row; each column carries
    its values once.  Returns (col_t, val_t, dval_t) with
    val_t[m, s, r] = G_m[c][r] and dval_t[t, s, r] = D_t[c][r] for
    c = col_t[s, r].
    """
    M, n, _ = mats.shape
    tr = np.any(mats != 0, axis=0).T
    if dmats.shape[0]:
        tr = tr | np.any(dmats != 0, axis=0).T
    entries = [[int(c) for c in np.nonzero(tr[r])[0] if c not in (r, r ^ 1)]
               if r < n else [] for r in range(64)]
    G = max(len(e) for e in entries)
    slots = 2 + G
    col = np.empty((slots, 64), dtype=np.int32)
    lanes = np.arange(64, dtype=np.int32)
    col[0] = lanes
    col[1] = lanes ^ 1
    if G:
        col[2:] = _bank_slots(entries, G, lambda r: r, 16, 16).T
    val_t = np.zeros((M, slots, 64))
    dval_t = np.zeros((dmats.shape[0], slots, 64))
    for r in range(n):
        seen = set()
        for sl in range(slots):
            c = int(col[sl, r])
            if c >= n or c in seen:
                continue
            seen.add(c)
            val_t[:, sl, r] = mats[:, c, r]
            if dmats.shape[0]:
                dval_t[:, sl, r] = dmats[:, c, r]
    col[0, :] = slots
    return col, val_t, dval_t


def wide_adjoint_ell(ell, mats, dmats):
    """Adjoint tables for n > 64 (several wavefronts per instance).

    ``ell`` is the propagator's ``EllStack`` (the matrices' own pattern,
    row stride NP); ``mats`` (M, n, n) and ``dmats`` (n_theta, n, n) are in
    its basis.  Returns (col_e, val_e, col_t, dval_e, slots_f):

    * ``col_e`` (slots_e, NP): row r lists the matrices' columns in the
      propagator's slot order, then the other columns with a non-zero
      derivative entry D_t[r][c]; slot 0 holds the used count | the
      matrices' count << 8;
    * ``val_e`` (M, slots_e, NP): G_m[r][c] (zero in derivative-only slots);
    * ``col_t`` (slots_t, NP): the transposed pattern of the matrices
      (slot 0 diagonal with the used count, slot 1 the partner r ^ 1);
    * ``dval_e`` (n_theta, slots_e, NP): D_t[r][c] at the slots of col_e;
    * ``slots_f``: slots of the matrices' own pattern (``ell.slots``).
    """
    M, n, _ = mats.shape
    NP = ell.npad
    nt = dmats.shape[0]
    dpat = np.any(dmats != 0, axis=0) if nt else np.zeros((n, n), bool)
    rows = []
    for r in range(n):
        own = [int(c) for c in ell.col[2:int(ell.col[0, r]), r]]
        taken = set(own) | {r, r ^ 1}
        extra = [int(c) for c in np.nonzero(dpat[r])[0]
                 if int(c) not in taken]
        rows.append(own + extra)
    slots_e = max(ell.slots, 2 + max(len(c) for c in rows))
    col_e = np.tile(np.arange(NP, dtype=np.int32), (slots_e, 1))
    col_e[0, :] = 2 | (2 << 8)
    val_e = np.zeros((M, slots_e, NP))
    val_e[:, :ell.slots] = ell.val
    dval_e = np.zeros((nt, slots_e, NP))
    for r in range(n):
        col_e[1, r] = r ^ 1
        for k, c in enumerate(rows[r]):
            col_e[2 + k, r] = c
        col_e[0, r] = (2 + len(rows[r])) | (int(ell.col[0, r]) << 8)
        if nt:
            dval_e[:, 0, r] = dmats[:, r, r]
            if (r ^ 1) < n:
                dval_e[:, 1, r] = dmats[:, r, r ^ 1]
            for k, c in enumerate(rows[r]):
                # a slot naming the row itself pads the matrices' pattern
                # (its diagonal is slot 0)
                if c != r:
                    dval_e[:, 2 + k, r] = dmats[:, r, c]
    union_t = np.any(mats != 0, axis=0).T
    rows_t = [[int(c) for c in np.nonzero(union_t[r])[0]
               if int(c) not in (r, r ^ 1)] for r in range(n)]
    slots_t = 2 + max(len(c) for c in rows_t)
    col_t = np.tile(np.arange(NP, dtype=np.int32), (slots_t, 1))
    col_t[0, :] = 2
    for r in range(n):
        col_t[1, r] = r ^ 1
        for k, c in enumerate(rows_t[r]):
            col_t[2 + k, r] = c
        col_t[0, r] = 2 + len(rows_t[r])
    return col_e, val_e, col_t, dval_e, ell.slots


class GradientProblem:
    """Batch of parametrised instances with an integrated readout."""

    def __init__(self, generators, waveforms, rho0, obs, time_start,
                 time_end, time_step, number_of_fine_divisions=1,
                 number_of_exponentials=2, checkpoint_every=1000,
                 tolerance=1e-16, adjoint_tolerance=1e-12, fd_rel=1e-6,
                 theta_scale=None, store_inputs="auto"):
        self.generators = generators
        self.waveforms = waveforms
        self.rho0 = np.asarray(rho0, dtype=_f64)
        self.obs = np.atleast_2d(np.asarray(obs, dtype=_f64))
        if self.obs.shape[0] > _lib.MAX_OBS:
            raise ValueError(f"at most {_lib.MAX_OBS} readout rows")
        self.t0 = float(time_start)
        self.t1 = float(time_end)
        self.dt = float(time_step)
        self.F = int(number_of_fine_divisions)
        self.J = int(number_of_exponentials)
        self.K = int(checkpoint_every)
        if self.K < 1:
            raise ValueError("checkpoint_every must be >= 1")
        self.tolerance = tolerance
        self.adjoint_tolerance = adjoint_tolerance
        self.fd_rel = fd_rel
        self.theta_scale = None if theta_scale is None else \
            np.asarray(theta_scale, dtype=_f64)
        self.n_out = int((self.t1 - self.t0) / self.dt)
        if store_inputs not in (True, False, "auto"):
            raise ValueError("store_inputs is True, False or 'auto'")
        self.store_inputs = store_inputs
        self._cache = {}

    # ------------------------------------------------------------------
    def stored_input_bytes(self, n_inst):
        """Device bytes of every exponential's input (stride 64 rows)."""
        return n_inst * self.n_out * self.F * self.J * _lib.MAX_N * 8

    def _use_stored_inputs(self, state):
        """Whether ``forward`` keeps every exponential's input, so that
        ``backward`` needs no recomputation: supported shapes (n <= 64,
        degree 11, at most 7 forward and 10 adjoint slots, 2 or 3 matrices)
        and, with ``store_inputs="auto"``, room for them in 80 % of the free
        device memory."""
        prop = state["prop"]
        if not state.get("derivatives", True):
            return False                 # a plain forward: no backward
        ok = (prop.n <= _lib.MAX_N and int(state["taylor"]) == 11 and
              state["slots"] <= 7 and state["slots_t"] <= 10 and
              prop.K in (2, 3))
        if self.store_inputs is False:
            return False
        if self.store_inputs is True:
            if not ok:
                raise NotImplementedError(
                    "stored inputs need n <= 64, Taylor degree 11, at most "
                    "7 / 10 slots and 2 or 3 matrices")
            return True
        if not ok:
            return False
        torch = _torch()
        need = self.stored_input_bytes(state["n_inst"])
        have = state.get("inputs")
        free, _ = torch.cuda.mem_get_info()
        if have is not None:
            free += have.numel() * 8
        return need <= 0.8 * free

    # ------------------------------------------------------------------
    def _fd_steps(self, theta):
        scale = np.abs(theta)
        if self.theta_scale is not None:
            scale = np.maximum(scale, self.theta_scale)
        scale = np.where(scale > 0, scale, 1.0)
        return self.fd_rel * scale

    def derivative_generators(self, theta):
        """d generators / d theta_k, (n_theta, G, n, n), central FD."""
        theta = np.asarray(theta, dtype=_f64)
        steps = self._fd_steps(theta)
        out = []
        for k in range(theta.size):
            tp = theta.copy()
            tm = theta.copy()
            tp[k] += steps[k]
            tm[k] -= steps[k]
            hk = tp[k] - tm[k]        # the step actually taken
            gp = np.asarray(self.generators(tp), dtype=_f64)
            gm = np.asarray(self.generators(tm), dtype=_f64)
            out.append((gp - gm) / hk)
        return np.array(out)

    def setup(self, theta, amplitudes, derivatives=True):
        """Plan, device tables and waveforms for one parameter point.

        ``derivatives=False`` skips the derivative tables (forward only).
        """
        theta = np.asarray(theta, dtype=_f64).ravel()
        amplitudes = np.asarray(amplitudes, dtype=_f64).ravel()
        gens = np.asarray(self.generators(theta), dtype=_f64)
        waves = self.waveforms(amplitudes)
        # plain layout: the gradient kernel gathers the shared forward /
        # transposed pattern one column at a time, and the fetch layouts
        # would widen that pattern
        sim = BatchSimulator(gens, self.J, self.F, tolerance=self.tolerance,
                             layout="plain")
        active, const, program = sim.program(waves)
        prop = sim._propagator(active, const, self.dt)
        h = self.dt / self.F
        samples = sim.coefficient_box(waves, program)
        m, ma = prop.plan_degrees(
            samples, h, (self.tolerance, self.adjoint_tolerance),
            max_taylor=(_lib.MAX_TAYLOR, _lib.ADJ_MAX_TAYLOR))

        # derivative tables of the folded constant matrix
        dgen = self.derivative_generators(theta) \
            if theta.size and derivatives else np.zeros((0,) + gens.shape)
        varying = [g for g in program if g is not None]
        for g in varying:
            if np.any(dgen[:, g] != 0):
                raise NotImplementedError(
                    f"generator {g} has a time-dependent coefficient and "
                    "depends on theta; only constant-coefficient generators "
                    "may depend on parameters")
        if const:
            dconst = sum(c * dgen[:, g] for g, c in const.items())
            theta_gen = 0
        else:
            if np.any(dgen != 0):
                raise NotImplementedError("theta enters no constant generator")
            dconst = np.zeros((dgen.shape[0],) + gens.shape[1:])
            theta_gen = -1

        kwaves = BatchSimulator.kernel_waves(waves, program)
        tagged = sorted({k for k, w in enumerate(program)
                         if any(np.any(wv.segment_arrays(k)[1] >= 0)
                                for wv in kwaves)})
        if len(tagged) > 1:
            raise NotImplementedError(
                "amplitude parameters on more than one generator")
        grad_gen = tagged[0] if tagged else -1
        recs, seg_arr, tag_index, tag_scale = waveform_records(
            kwaves, len(program), with_tags=True)

        perm = prop.perm
        mats = prop._dense                                   # (M, n, n)
        mp = mats[:, perm][:, :, perm]
        dp = dconst[:, perm][:, :, perm]
        torch = _torch()
        if not derivatives:
            # forward only: the adjoint's tables (whose pattern may exceed
            # the adjoint kernels' limits where the forward's does not) are
            # not built; backward refuses such a state
            tables = dict(d_col=None, d_val_f=None, d_col_t=None,
                          d_val_t=None, d_dval_t=None, slots=prop.ell.slots,
                          slots_t=0, slots_f=0)
        elif prop.n > _lib.MAX_N:
            if grad_gen >= 0:
                raise NotImplementedError(
                    "amplitude gradients need n <= 64")
            col_e, val_e, col_t, dval_e, slots_f = wide_adjoint_ell(
                prop.ell, mp, dp)
            if max(slots_f, col_t.shape[0]) > _lib.ADJ_WIDE_MAX_SLOTS or \
                    col_e.shape[0] > _lib.ADJ_MAX_SLOTS:
                raise NotImplementedError(
                    f"wide adjoint pattern {slots_f}/{col_t.shape[0]}/"
                    f"{col_e.shape[0]} slots beyond the kernel's limits")
            tables = dict(
                d_col=device_tensor(col_e), d_val_f=device_tensor(val_e),
                d_col_t=device_tensor(col_t), d_val_t=None,
                d_dval_t=device_tensor(dval_e) if dconst.shape[0] else None,
                slots=col_e.shape[0], slots_t=col_t.shape[0],
                slots_f=slots_f)
        else:
            col, val_f, _, _ = shared_adjoint_ell(prop.ell.col, mp, dp)
            # the adjoint chain gathers its own, bank-ordered pattern
            col_t, val_t, dval_t = transposed_adjoint_ell(mp, dp)
            if max(col.shape[0], col_t.shape[0]) > _lib.ADJ_MAX_SLOTS:
                raise NotImplementedError(
                    f"adjoint pattern of {col.shape[0]} / {col_t.shape[0]} "
                    f"slots > {_lib.ADJ_MAX_SLOTS}")
            tables = dict(
                d_col=device_tensor(col),
                # forward values are zero beyond the forward kernel's slots
                d_val_f=device_tensor(val_f[:, :prop.ell.slots].copy()),
                d_col_t=device_tensor(col_t), d_val_t=device_tensor(val_t),
                d_dval_t=device_tensor(dval_t) if dconst.shape[0] else None,
                slots=prop.ell.slots, slots_t=col_t.shape[0], slots_f=0)
        state = dict(
            theta=theta, amplitudes=amplitudes, prop=prop, program=program,
            derivatives=bool(derivatives),
            taylor=m, taylor_adj=ma, theta_gen=theta_gen, grad_gen=grad_gen,
            n_theta=dconst.shape[0] if theta_gen >= 0 else 0,
            d_waves=torch.from_numpy(np.frombuffer(
                recs.tobytes(), dtype=np.uint8).copy()).to("cuda"),
            d_segs=device_tensor(seg_arr),
            tag_index=tag_index, tag_scale=tag_scale,
            n_records=seg_arr.shape[0], n_inst=len(waves), dconst=dconst,
            kernel_waves=kwaves, **tables)
        return state

    # ------------------------------------------------------------------
    def forward(self, theta, amplitudes, state=None):
        """(F (n_inst, n_obs) device tensor, state for ``backward``)."""
        torch = _torch()
        if state is None:
            state = self.setup(theta, amplitudes)
        prop = state["prop"]
        n_inst = state["n_inst"]
        rho0 = np.atleast_2d(self.rho0)
        if rho0.shape[0] == 1 and n_inst > 1:
            rho0 = np.repeat(rho0, n_inst, axis=0)
        total = self.n_out * self.F
        P = (total + self.K - 1) // self.K
        stored = self._use_stored_inputs(state)
        if stored:
            # every exponential's input (the backward's forward terms start
            # from them); the buffer is reused across calls
            shape = (n_inst, total, prop.J, _lib.MAX_N)
            ckpt = state.get("inputs")
            if ckpt is None or tuple(ckpt.shape) != shape:
                state.pop("inputs", None)
                state.pop("ckpt", None)
                ckpt = torch.empty(shape, dtype=torch.float64, device="cuda")
            state["inputs"] = ckpt
        else:
            state.pop("inputs", None)
            ckpt = torch.empty((n_inst, P, prop.n), dtype=torch.float64,
                               device="cuda")
        out = prop.run(n_inst=n_inst, fine=self.F, n_out=self.n_out,
                       h=self.dt / self.F, t0=self.t0, dt=self.dt,
                       rho0=device_tensor(rho0), waves=state["d_waves"],
                       segments=state["d_segs"], taylor=state["taylor"],
                       substeps=1, read="integral", obs=self.obs,
                       state_out=ckpt, ckpt_every=0 if stored else self.K,
                       store_inputs=stored)
        state["ckpt"] = ckpt
        state["n_ckpt"] = P
        state["stored"] = stored
        return out, state

    def adjoint_desc(self, state):
        prop = state["prop"]
        return _lib.AdjointDesc(
            n=prop.n, n_inst=state["n_inst"], n_gen=prop.K,
            slots=state["slots"], slots_t=state["slots_t"], n_exp=prop.J,
            n_node=prop.S, fine=self.F, n_out=self.n_out,
            taylor=int(state["taylor"]), taylor_adj=int(state["taylor_adj"]),
            n_obs=self.obs.shape[0],
            flags=_lib.SSMX_COEF_WAVE | (_lib.SSMX_STORE_INPUTS
                                         if state.get("stored") else 0),
            first_out=0,
            partner=1, ckpt_every=self.K, n_theta=int(state["n_theta"]),
            theta_gen=int(state["theta_gen"]),
            grad_gen=int(state["grad_gen"]), reserved=int(state["slots_f"]),
            coef_inst_stride=0, h=self.dt / self.F, t0=self.t0, dt=self.dt)

    def backward(self, state, seed, lambda_out=False):
        """Per-instance gradients for readout seeds ``seed`` (n_inst, n_obs).

        Returns a dict with ``theta`` (n_inst, n_theta), ``amplitudes``
        (n_amplitudes,) summed over instances, ``records`` (per segment
        record) and optionally ``lambda`` (n_inst, n), dLoss/d rho0 in the
        standard basis.
        """
        if not state.get("derivatives", True):
            raise ValueError("this state was set up without derivatives "
                             "(setup(..., derivatives=False) or readout); "
                             "use forward(theta, amplitudes) for gradients")
        torch = _torch()
        prop = state["prop"]
        n_inst = state["n_inst"]
        desc = self.adjoint_desc(state)
        lib = _lib.load()
        nbytes = ctypes.c_uint64(0)
        _lib.check(lib.ssmx_adjoint_workspace_bytes(ctypes.byref(desc),
                                                   ctypes.byref(nbytes)))
        work = torch.empty(int(nbytes.value) // 8, dtype=torch.float64,
                           device="cuda")
        seed = torch.as_tensor(seed, dtype=torch.float64).to("cuda")
        seed = seed.reshape(n_inst, self.obs.shape[0]).contiguous()
        rows = np.zeros((self.obs.shape[0], prop.ell.npad), dtype=_f64)
        rows[:, :prop.n] = self.obs[:, prop.perm]
        d_obs = device_tensor(rows)
        n_theta = int(state["n_theta"])
        g_theta = torch.zeros((n_inst, max(n_theta, 1)), dtype=torch.float64,
                              device="cuda")
        g_seg = torch.zeros(state["n_records"], dtype=torch.float64,
                            device="cuda")
        lam = torch.empty((n_inst, prop.n), dtype=torch.float64,
                          device="cuda") if lambda_out else None
        stream = torch.cuda.current_stream().cuda_stream
        rc = lib.ssmx_adjoint(
            ctypes.byref(desc), _ptr(state["d_col"]), _ptr(state["d_val_f"]),
            _ptr(state["d_col_t"]), _ptr(state["d_val_t"]),
            _ptr(state["d_dval_t"]), _ptr(prop.d_w), _ptr(prop.d_nodes),
            _ptr(state["d_waves"]), _ptr(state["d_segs"]),
            _ptr(state["ckpt"]), _ptr(d_obs), _ptr(seed), _ptr(work),
            _ptr(g_theta) if n_theta else None,
            _ptr(g_seg) if state["grad_gen"] >= 0 else None,
            _ptr(lam), ctypes.c_void_p(stream))
        _lib.check(rc)
        self._keep = (work, d_obs, seed)
        result = dict(theta=g_theta[:, :n_theta], records=g_seg)
        n_amp = state["amplitudes"].size
        if "amp_sum" not in state:
            # records grouped by amplitude (stable: ascending record order),
            # so that each amplitude's sum has a fixed order
            idx = state["tag_index"]
            sel = np.nonzero(idx >= 0)[0]
            sel = sel[np.argsort(idx[sel], kind="stable")]
            state["amp_sum"] = (
                torch.from_numpy(sel).to("cuda"),
                torch.from_numpy(state["tag_scale"][sel]).to("cuda"),
                torch.from_numpy(np.bincount(idx[sel], minlength=n_amp))
                .to("cuda"))
        d_sel, d_scale, d_len = state["amp_sum"]
        if d_sel.numel():
            g_amp = torch.segment_reduce(g_seg[d_sel] * d_scale, "sum",
                                         lengths=d_len)
        else:
            g_amp = torch.zeros(n_amp, dtype=torch.float64, device="cuda")
        result["amplitudes"] = g_amp
        if lambda_out:
            result["lambda"] = lam[:, prop.d_inv_perm]
        return result

    # ------------------------------------------------------------------
    def readout(self, theta, amplitudes):
        """Forward only: the integrated readouts (n_inst, n_obs)."""
        state = self.setup(theta, amplitudes, derivatives=False)
        out, _ = self.forward(theta, amplitudes, state)
        return out


class IntegratedReadout:
    """``torch.autograd.Function`` over ``GradientProblem``.

    ``IntegratedReadout.apply(theta, amplitudes, problem)`` returns the
    (n_inst, n_obs) integrated readouts; gradients flow to ``theta`` and
    ``amplitudes`` (float64 tensors on any device).
    """

    _fn = None

    @classmethod
    def function(cls):
        if cls._fn is None:
            torch = __import__("torch")

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, theta, amplitudes, problem):
                    out, state = problem.forward(
                        theta.detach().cpu().numpy(),
                        amplitudes.detach().cpu().numpy())
                    ctx.problem = problem
                    ctx.state = state
                    ctx.devices = (theta.device, amplitudes.device)
                    return out

                @staticmethod
                def backward(ctx, grad_out):
                    res = ctx.problem.backward(ctx.state, grad_out)
                    gt = res["theta"].sum(dim=0).to(ctx.devices[0])
                    ga = res["amplitudes"].to(ctx.devices[1])
                    return gt, ga, None

            cls._fn = _Fn
        return cls._fn

    @classmethod
    def apply(cls, theta, amplitudes, problem):
        return cls.function().apply(theta, amplitudes, problem)


def step_flops_adjoint(state, slots_nnz_t):
    """Algorithmic FP64 FLOPs of one backward step of one instance.

    Recomputation: J (2 K nnz + m 2 (nnz + n)); adjoint: J [(m_a - 1)
    (2 nnz + 2 n) forward terms + m_a (2 nnz_t + 2 n) adjoint terms +
    m_a 2 nnz_t Frechet accumulation + m_a^2 n coefficient sums].
    """
    prop = state["prop"]
    J, n = prop.J, prop.n
    nnz = prop.ell.nnz
    nnz_t = slots_nnz_t
    m, ma = state["taylor"], state["taylor_adj"]
    K = prop.K
    recompute = J * (2 * K * nnz + m * 2 * (nnz + n))
    adj = J * ((ma - 1) * (2 * nnz + 2 * n) + ma * (2 * nnz_t + 2 * n)
               + ma * 2 * nnz_t + ma * ma * n + 2 * K * (nnz + nnz_t))
    return recompute + adj


__all__ = ["GradientProblem", "IntegratedReadout", "wide_adjoint_ell"]
