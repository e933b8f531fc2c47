"""RGB view of complex matrices (``superspinsim/util.py:5-52``).

Plotting helper only; it is not on the simulation path.  The colour map is
the reference's: the (re, im) plane is projected onto three RGB axes 120
degrees apart, brightened by ``1 - |z|`` and clipped to [0, 1], so that zero
is white and unit-modulus entries are saturated with hue = phase.
"""

import math

import numpy as np

# rows: R, G, B; columns: weight of (re, im)
_PROJECTION = np.array([[2 / math.sqrt(6), 0.0],
                        [-1 / math.sqrt(6), 1 / math.sqrt(2)],
                        [-1 / math.sqrt(6), -1 / math.sqrt(2)]])


def colour_complex_matrix(inp):
    """(N, M) complex/real or (N, M, 2) re/im matrix -> (N, M, 3) RGB."""
    arr = np.asarray(inp)
    if arr.ndim == 2:
        pair = np.stack([np.real(arr), np.imag(arr)], axis=-1)
    else:
        pair = arr[..., :2]
    pair = pair.astype(np.float64)
    rgb = pair @ _PROJECTION.T
    rgb += (1 - np.sqrt(pair[..., 0] ** 2 + pair[..., 1] ** 2))[..., None]
    return np.clip(rgb, 0, 1)
