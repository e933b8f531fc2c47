"""The allocation-free queue entry point ``ssmx_propagate_queue_noalloc``.

It is ``ssmx_propagate_queue`` with the work-queue counter taken from the
library's own per-device scratch array instead of an argument.  On the GPU
its readouts must equal those of the existing entry point bit for bit (same
kernels, same grid), on two calls in a row with the second batch smaller,
so a counter left over from the first call would show.  On the CPU: the
symbol, its argument checks and its size limit, all before any HIP call.
"""

import ctypes

import numpy as np
import pytest

from superspinsim_amd import _lib, configs

H = 20e-12


def _lib_or_build():
    import os
    if not os.path.exists(_lib.LIB_PATH):
        from superspinsim_amd.build import build_library
        build_library()
    return _lib.load()


def _qdesc(**kw):
    base = dict(n=4, n_inst=3, n_gen=1, slots=2, n_exp=1, n_node=1, fine=1,
                n_out=0, taylor=4, substeps=1, n_obs=1,
                flags=_lib.SSMX_COEF_WAVE | _lib.SSMX_READ_INTEGRAL,
                first_out=0, partner=1, coef_inst_stride=0, h=1e-11, t0=0.0,
                dt=1e-11)
    base.update(kw)
    return _lib.PropagateDesc(**base)


def _call(lib, d, ptr, n_waves=4, wpb=2):
    return lib.ssmx_propagate_queue_noalloc(
        ctypes.byref(d), *([ptr] * 12), ctypes.c_int32(n_waves),
        ctypes.c_int32(wpb), ptr, None)


def test_noalloc_rejects_bad_arguments_before_any_hip_call():
    lib = _lib_or_build()
    dummy = ctypes.c_void_p(8)
    assert _call(lib, _qdesc(n=0), dummy) == -1
    assert _call(lib, _qdesc(flags=_lib.SSMX_READ_FULL), dummy) == -1
    assert _call(lib, _qdesc(), dummy, wpb=9) == -1
    assert _call(lib, _qdesc(), None) == -1
    assert b"null" in lib.ssmx_last_error()
    # the counter ends at n_inst + n_waves, which must fit the scratch int32
    d = _qdesc(n_inst=_lib.NOALLOC_MAX_QUEUE)
    assert _call(lib, d, dummy, n_waves=2) == -1
    assert b"scratch" in lib.ssmx_last_error()


def _batch(count):
    wl = configs.pulsed7(count=count, seed=5, t_p_range=(2e-9, 20e-9),
                         tau_range=(10e-9, 100e-9), readout=50e-9)
    return wl.batch().prepare(wl.waves, wl.rho0, H, wl.steps, wl.read_from,
                              wl.obs, n_waves=7, waves_per_block=2)


@pytest.mark.gpu
def test_noalloc_equals_counter_argument_twice_second_smaller(gpu):
    import torch
    for count in (24, 9):
        b = _batch(count)
        print(f"queue of {count} protocols: counter argument", flush=True)
        want = b.launch().cpu().numpy().copy()
        assert np.all(np.isfinite(want)) and np.all(want > 0)
        b.out.fill_(np.nan)
        print(f"queue of {count} protocols: no-alloc entry point",
              flush=True)
        got = b.launch_noalloc().cpu().numpy()
        torch.cuda.synchronize()
        assert np.array_equal(got, want), count
