"""MI355X-native propagator for superspinsim-style open spin systems.

Drop-in surface of ``superspinsim/__init__.py:1-4``: ``generate_simulator``,
``simspins``, ``mesolve``, ``colour_complex_matrix`` and ``models``.  The
simulation runs in hand-written HIP kernels (``lib/libssmx.so``) on the
first HIP device; host-side model construction is numpy.
"""

from superspinsim_amd.generate_simulator import generate_simulator
from superspinsim_amd.sugar import mesolve, simspins
from superspinsim_amd.util import colour_complex_matrix
from superspinsim_amd import models

__all__ = ["generate_simulator", "simspins", "mesolve",
           "colour_complex_matrix", "models"]
