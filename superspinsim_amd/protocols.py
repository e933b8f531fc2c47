"""Batches of protocols with uneven lengths in one persistent launch.

``ProtocolBatch`` runs many instances of one model whose waveforms (pulse
sequences, free precession, readout windows) have different durations.  All
instances share the propagator step h; instance i runs ``steps[i]`` steps
and reports the trapezoid integral of its readout rows over steps
``read_from[i] .. steps[i]`` (e.g. the fluorescence collected during a final
optical readout window).

The launch is ``ssmx_propagate_queue``: a grid of about as many wavefronts as
the device holds, each taking the next instance from an atomic counter.
Instances are queued longest first (greedy longest-processing-time), so the
tail of the launch is filled with short protocols.  ``capture`` records the
counter reset and the launch in a HIP graph (``torch.cuda.CUDAGraph``) for
repeated replay; the reset is a fill kernel, not a memset node.

Launch shape (wavefronts per workgroup) and the Taylor plan (degree m and
substeps s: the norm above which an exponential is split) are chosen by
measurement with ``superspinsim_amd.autotune.tune(batch)``.
"""

import ctypes

import numpy as np

from superspinsim_amd import _lib
from superspinsim_amd.propagator import _ptr, _torch, device_tensor
from superspinsim_amd.sweep import BatchSimulator, waveform_records

_f64 = np.float64

#: wavefronts per SIMD at 4 waves/SIMD occupancy (kernels up to 8 slots)
WAVES_PER_CU = 16


def resident_wavefronts(waves_per_cu=WAVES_PER_CU):
    """Wavefronts the device holds at once (CUs x waves per CU)."""
    torch = _torch()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    return int(cus * waves_per_cu)


class ProtocolBatch:
    """Uneven protocols of one model (see module docstring).

    ``generators`` (G, n, n) in coefficient-row order, CF scheme and
    tolerance as ``BatchSimulator`` (non-rotating frame).
    """

    def __init__(self, generators, number_of_exponentials=2,
                 tolerance=1e-16):
        self.sim = BatchSimulator(generators, number_of_exponentials, 1,
                                  tolerance=tolerance)
        self.tolerance = tolerance
        self.prepared = False

    # ------------------------------------------------------------------
    def prepare(self, waves, rho0, time_step, steps, read_from, obs,
                taylor=None, substeps=None, waves_per_block=4, n_waves=None):
        """Plan and stage every input on the device (nothing is timed).

        ``steps``/``read_from``: (n_inst,) integers; ``rho0`` (n,) or
        (n_inst, n); ``obs`` (n_obs, n) readout rows (standard basis).
        """
        torch = _torch()
        steps = np.asarray(steps, dtype=np.int64)
        read_from = np.asarray(read_from, dtype=np.int64)
        n_inst = len(waves)
        if steps.shape != (n_inst,) or read_from.shape != (n_inst,):
            raise ValueError("steps and read_from need one entry per "
                             "instance")
        if np.any(steps < 0) or np.any(read_from < 0):
            raise ValueError("negative step counts")
        self.h = float(time_step)
        self.waves_host = waves
        prop, program, m, s = self.sim.plan(waves, self.h)
        self.prop = prop
        self.program = program
        self.plan_default = (m, s)
        self.taylor = int(taylor if taylor is not None else m)
        self.substeps = int(substeps if substeps is not None else s)
        kwaves = self.sim.kernel_waves(waves, program)
        recs, seg_arr = waveform_records(kwaves, len(program))
        self.d_waves = torch.from_numpy(np.frombuffer(
            recs.tobytes(), dtype=np.uint8).copy()).to("cuda")
        self.d_segs = device_tensor(seg_arr)
        rho0 = np.atleast_2d(np.asarray(rho0, dtype=_f64))
        if rho0.shape[0] == 1 and n_inst > 1:
            rho0 = np.repeat(rho0, n_inst, axis=0)
        self.d_rho0 = device_tensor(rho0[:, prop.perm])
        obs = np.atleast_2d(np.asarray(obs, dtype=_f64))
        rows = np.zeros((obs.shape[0], 64), dtype=_f64)
        rows[:, :prop.n] = obs[:, prop.perm]
        self.n_obs = obs.shape[0]
        self.d_obs = device_tensor(rows)
        self.steps = steps
        self.read_from = read_from
        self.order = np.argsort(-steps, kind="stable").astype(np.int32)
        self.d_order = device_tensor(self.order)
        self.d_steps = device_tensor(steps)
        self.d_read_from = device_tensor(read_from)
        self.counter = torch.zeros(1, dtype=torch.int32, device="cuda")
        self.out = torch.empty((n_inst, self.n_obs), dtype=torch.float64,
                               device="cuda")
        self.n_inst = n_inst
        self.waves_per_block = int(waves_per_block)
        self.n_waves = int(n_waves or min(n_inst, resident_wavefronts()))
        self.prepared = True
        return self

    def total_steps(self):
        return int(self.steps.sum())

    def _desc(self):
        prop = self.prop
        flags = _lib.SSMX_COEF_WAVE | _lib.SSMX_READ_INTEGRAL
        return _lib.PropagateDesc(
            n=prop.n, n_inst=self.n_inst, n_gen=prop.K, slots=prop.ell.slots,
            n_exp=prop.J, n_node=prop.S, fine=1, n_out=0,
            taylor=self.taylor, substeps=self.substeps, n_obs=self.n_obs,
            flags=flags, first_out=0, partner=1, ckpt_every=0, reserved=0,
            coef_inst_stride=0, h=self.h, t0=0.0, dt=self.h)

    def launch(self, stream=None):
        """Reset the queue counter and launch (asynchronous)."""
        if not self.prepared:
            raise RuntimeError("prepare() first")
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
            self.counter.zero_()
        else:
            # the reset (a fill kernel) must run on the launch's stream, or
            # the kernel could read the previous launch's final count
            with torch.cuda.stream(torch.cuda.ExternalStream(stream)):
                self.counter.zero_()
        desc = self._desc()
        prop = self.prop
        rc = _lib.load().ssmx_propagate_queue(
            ctypes.byref(desc), _ptr(prop.d_col), _ptr(prop.d_val),
            _ptr(prop.d_w), _ptr(prop.d_nodes), _ptr(self.d_waves),
            _ptr(self.d_segs), _ptr(self.d_rho0), None, _ptr(self.d_obs),
            _ptr(self.d_order), _ptr(self.d_steps), _ptr(self.d_read_from),
            _ptr(self.counter), ctypes.c_int32(self.n_waves),
            ctypes.c_int32(self.waves_per_block), _ptr(self.out),
            ctypes.c_void_p(stream))
        _lib.check(rc)
        return self.out

    def launch_noalloc(self, stream=None):
        """Launch through ``ssmx_propagate_queue_noalloc``: the library's
        own per-device counter, cleared on ``stream``; no allocation on
        this path.  Overlapping calls on one device race."""
        if not self.prepared:
            raise RuntimeError("prepare() first")
        torch = _torch()
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        desc = self._desc()
        prop = self.prop
        rc = _lib.load().ssmx_propagate_queue_noalloc(
            ctypes.byref(desc), _ptr(prop.d_col), _ptr(prop.d_val),
            _ptr(prop.d_w), _ptr(prop.d_nodes), _ptr(self.d_waves),
            _ptr(self.d_segs), _ptr(self.d_rho0), None, _ptr(self.d_obs),
            _ptr(self.d_order), _ptr(self.d_steps), _ptr(self.d_read_from),
            ctypes.c_int32(self.n_waves),
            ctypes.c_int32(self.waves_per_block), _ptr(self.out),
            ctypes.c_void_p(stream))
        _lib.check(rc)
        return self.out

    def run(self):
        """Eager launch; returns the (n_inst, n_obs) readouts (device)."""
        return self.launch()

    def capture(self):
        """HIP graph of (counter reset, launch); ``replay()`` reruns it and
        refreshes ``self.out``."""
        torch = _torch()
        self.launch()                   # warm: first launch outside capture
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self.launch()
        self.graph = graph
        return graph

    def step_flops(self):
        """Algorithmic FP64 FLOPs of one propagator step of one instance."""
        prop = self.prop
        J, n, nnz, K = prop.J, prop.n, prop.ell.nnz, prop.K
        per_exp = 2 * K * nnz + n + self.substeps * self.taylor * (
            2 * nnz + 2 * n)
        return J * per_exp


def trapezoid_window(series, read_from, steps, h):
    """Host reference of the windowed trapezoid for a per-step series.

    ``series`` (n_inst, T + 1) holds obs . x_k for k = 0..T (k = 0 is
    rho0); returns sum_{k=read_from}^{steps-1} (h/2)(s_k + s_{k+1}).
    """
    out = np.zeros(series.shape[0])
    for i in range(series.shape[0]):
        a, b = int(read_from[i]), int(steps[i])
        if b > a:
            seg = series[i, a:b + 1]
            out[i] = 0.5 * h * float(np.sum(seg[:-1] + seg[1:]))
    return out


__all__ = ["ProtocolBatch", "resident_wavefronts", "trapezoid_window"]
