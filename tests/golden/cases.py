"""Input recipes of the reference's golden density histories.

The seven CSV files under ``tests/golden/csv`` are the reference's own
fixtures (``test/reference/*.csv``): every 29th element of the flattened
``(T, N_h, N_h, 2)`` density history of a 10 us run with 1 ns output and
``number_of_fine_divisions=100`` (``test/tests/util.py:12-30``).  Each
``GOLDEN`` entry restates the spin description, fields, initial state and
flags of the script that wrote that file (``test/tests/<writer>.py``).
``READERS`` lists the flag variants of the other scripts that compare
against the same file with the reference's loose metric
(``test/tests/util.py:42-59``).
"""

import math
import os

import numpy as np

from superspinsim_amd.golden import density_metric as reference_metric
from superspinsim_amd.golden import flatten_density, load_density

CSV_DIR = os.path.join(os.path.dirname(__file__), "csv")


def load_csv(name):
    return load_density(name, CSV_DIR)


__all__ = ["CSV_DIR", "GOLDEN", "READERS", "case", "flatten_density",
           "load_csv", "reference_metric"]


def _fields(frequency, amplitude, cutoff=None):
    def mag_x(time):
        if cutoff is not None and not time < cutoff:
            return 0
        return amplitude * math.cos(math.tau * frequency * time)

    def zero(time):
        return 0.0

    return [mag_x, zero, zero, zero]


def _qubit(extra=None):
    atom = {"S": 1 / 2, "g": 2.0, "B0": np.array([0, 0, 1e-3])}
    atom.update(extra or {})
    return [[atom]]


def _qutrit(extra=None):
    atom = {"S": 1, "g": -1 / 2, "D": 72, "B0": np.array([0, 0, 1e-3])}
    atom.update(extra or {})
    return [[atom]]


def _rho(size, diag):
    rho = np.zeros((size, size), dtype=np.float64)
    for k, v in diag.items():
        rho[k, k] = v
    return rho


_RABI_10MT = 10e-3 * 28e9


def case(name):
    """Return ``(coefficients, spins, spin_int, group_int, rho0)``."""
    if name == "qubit-couple":
        return (_fields(28e6, 10e-6), _qubit(), [{}], {}, _rho(2, {0: 1}))
    if name == "qubit-hyperfine":
        # The reference dict repeats the key "g"; the later value 1.0 wins.
        spins = [[{"S": 1 / 2, "g": 1.0, "B0": np.array([0, 0, 1e-3]),
                   "I": 1 / 2, "A": 2e9}]]
        return (_fields(28e6, 10e-6), spins, [{}], {}, _rho(4, {0: 1}))
    if name == "qubit-t1-0K":
        return (_fields(28e6, 10e-3, 1 / (2 * _RABI_10MT)),
                _qubit({"TS1": 1e-6, "T": 0}), [{}], {}, _rho(2, {1: 1}))
    if name == "qubit-t1-mK":
        return (_fields(28e6, 10e-3, 1 / (2 * _RABI_10MT)),
                _qubit({"TS1": 1e-6, "T": 1e-3}), [{}], {}, _rho(2, {1: 1}))
    if name == "qutrit-couple":
        return (_fields(7e6, 10e-6), _qutrit(), [{}], {}, _rho(3, {2: 1}))
    if name == "qutrit-couple-unital":
        return (_fields(7e6, 10e-6), _qutrit(), [{}], {},
                _rho(3, {0: 1 / 3, 1: 1 / 3, 2: 1 / 3}))
    if name == "qutrit-hyperfine":
        # "gI" is not a key the nucleus reads; the default gN = 2 applies.
        spins = _qutrit({"I": 1 / 2, "gI": 1.0, "A": 2e9})
        return (_fields(7e6, 10e-6), spins, [{}], {}, _rho(6, {0: 1}))
    raise KeyError(name)


_BASE = dict(number_of_fine_divisions=100, use_kernel=False,
             use_unitary=False)

#: writer flags of each golden file
GOLDEN = {
    "qubit-couple": dict(_BASE, use_rotating=False, use_residual=False,
                         number_of_exponentials=1),
    "qubit-hyperfine": dict(_BASE, use_rotating=True, use_residual=True,
                            number_of_exponentials=1),
    "qubit-t1-0K": dict(_BASE, use_rotating=False, use_residual=True,
                        number_of_exponentials=1),
    "qubit-t1-mK": dict(_BASE, use_rotating=False, use_residual=True,
                        number_of_exponentials=1),
    "qutrit-couple": dict(_BASE, use_rotating=False, use_residual=True,
                          number_of_exponentials=1),
    "qutrit-couple-unital": dict(_BASE, use_rotating=False,
                                 use_residual=True, number_of_exponentials=1),
    "qutrit-hyperfine": dict(_BASE, use_rotating=True, use_residual=True,
                             number_of_exponentials=1),
}


def _v(golden, **flags):
    out = dict(_BASE)
    out.update(flags)
    return golden, out


#: (script, golden file, flags) of the reference's reader scripts
READERS = {
    "qubit-couple-res": _v("qubit-couple", use_rotating=False,
                           use_residual=True, number_of_exponentials=1),
    "qubit-couple-res-ker": _v("qubit-couple", use_rotating=False,
                               use_residual=True, number_of_exponentials=1,
                               use_kernel=True),
    "qubit-couple-res-ker-rot": _v("qubit-couple", use_rotating=True,
                                   use_residual=True,
                                   number_of_exponentials=1, use_kernel=True),
    "qubit-couple-res-ker-rot-e5": _v("qubit-couple", use_rotating=True,
                                      use_residual=True,
                                      number_of_exponentials=5,
                                      use_kernel=True),
    "qubit-couple-res-rot": _v("qubit-couple", use_rotating=True,
                               use_residual=True, number_of_exponentials=1),
    "qubit-couple-res-rot-e2": _v("qubit-couple", use_rotating=True,
                                  use_residual=True,
                                  number_of_exponentials=2),
    "qubit-couple-res-rot-e3": _v("qubit-couple", use_rotating=True,
                                  use_residual=True,
                                  number_of_exponentials=5),
    "qubit-couple-res-rot-e5": _v("qubit-couple", use_rotating=True,
                                  use_residual=True,
                                  number_of_exponentials=5),
    "qubit-couple-res-rot-e6": _v("qubit-couple", use_rotating=True,
                                  use_residual=True,
                                  number_of_exponentials=6),
    "qubit-couple-res-rot-nh": _v("qubit-couple", use_rotating=True,
                                  use_residual=True,
                                  number_of_exponentials=1,
                                  use_hermitian=False),
    "qubit-t1-0K-res-rot": _v("qubit-t1-0K", use_rotating=True,
                              use_residual=True, number_of_exponentials=1),
    "qubit-t1-mK-res-rot": _v("qubit-t1-mK", use_rotating=True,
                              use_residual=True, number_of_exponentials=1),
    "qubit-t1-mK-res-rot-e2": _v("qubit-t1-mK", use_rotating=True,
                                 use_residual=True,
                                 number_of_exponentials=2),
    "qubit-t1-mK-res-rot-e5": _v("qubit-t1-mK", use_rotating=True,
                                 use_residual=True,
                                 number_of_exponentials=5),
    "qutrit-couple-ker": _v("qutrit-couple", use_rotating=False,
                            use_residual=True, number_of_exponentials=1,
                            use_kernel=True),
    "qutrit-couple-ker-rot": _v("qutrit-couple", use_rotating=True,
                                use_residual=True, number_of_exponentials=1,
                                use_kernel=True),
    "qutrit-couple-ker-rot-e5": _v("qutrit-couple", use_rotating=True,
                                   use_residual=True,
                                   number_of_exponentials=5,
                                   use_kernel=True),
    "qutrit-couple-rot": _v("qutrit-couple", use_rotating=True,
                            use_residual=True, number_of_exponentials=1),
    "qutrit-couple-unital-ker": _v("qutrit-couple-unital",
                                   use_rotating=False, use_residual=True,
                                   number_of_exponentials=1,
                                   use_kernel=True),
    "qutrit-couple-unitary": _v("qutrit-couple", use_rotating=True,
                                use_residual=True, number_of_exponentials=5,
                                use_unitary=True),
}
