"""Waveform coefficients for the oracle, vectorised over node times.

Test infrastructure only (see ``oracle/__init__.py``).  A
``superspinsim_amd.sweep.Waveform`` assigns each generator coefficient

    c_g(t) = base_g + sum_k [t_start_k <= t < t_end_k] amp_k cos(omega_k t
                                                                 + phase_k)

and the oracle's ``vectorised_sampler`` hook wants ``ts -> (len(ts), G)``.
This evaluates that formula with numpy over all node times at once, so a
full-length run (hundreds of thousands of nodes) does not pay one Python
call per node.
"""

import numpy as np


def vectorised_sampler(wave):
    """``ts -> (len(ts), n_gen)`` coefficient rows of ``wave``."""
    base = np.asarray(wave.base, dtype=np.float64)
    segments = [list(s) for s in wave.segments]

    def sample(ts):
        ts = np.asarray(ts, dtype=np.float64)
        out = np.empty((ts.size, base.size), dtype=np.float64)
        out[:] = base
        for g, segs in enumerate(segments):
            for t_start, t_end, amp, omega, phase in segs:
                on = (t_start <= ts) & (ts < t_end)
                out[on, g] += amp * np.cos(omega * ts[on] + phase)
        return out

    return sample
