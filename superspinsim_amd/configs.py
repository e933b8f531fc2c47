"""Frozen benchmark and parity workloads (NV ODMR sweeps).

* ``odmr3`` - spin-1 NV ground triplet, 9x9 superoperator: D = 2.87 GHz,
  Bz = 5 mT, T1 = 1 ms, T2* = 1 us, drive Bx(t) = B1 cos(2 pi f t) with
  Rabi frequency 5 MHz, 256 frequencies in 2.77-2.97 GHz, 1 us at 1 ns
  readout with 50 fine steps (h = 20 ps), CF4:2, readout of the ms = 0
  population.
* ``odmr7`` - 7-level NV (ground triplet D = 2.87 GHz, excited triplet
  D = 1.42 GHz, singlet), Bz = 5 mT, radiative decay 65 MHz, ISC 50 MHz
  (excited +-1) and 8 MHz (excited 0), singlet -> 0 at 1 MHz and -> +-1 at
  0.7 MHz, optical pumping 5 MHz, Rabi 3 MHz, 4096 frequencies in
  2.77-2.97 GHz, 2 us at 2 ns readout with 100 fine steps (h = 20 ps),
  CF4:2, readout of the fluorescence rho33 + rho44 + rho55.  Full
  (coherent) vectorisation, n = 49 (``use_incoherent=False``); n = 19 with
  the compressed one.

Drive amplitude convention (the reference docs' ``examples.rst``):
``B1 = sqrt(2) * Omega / 28 GHz/T`` for a Rabi frequency Omega (Hz).
"""

import copy
import math

import numpy as np

from superspinsim_amd import models, params
from superspinsim_amd.generate_generators import generate_atoms
from superspinsim_amd.sweep import BatchSimulator, Waveform

GYRO_HZ_PER_T = 28e9


def drive_amplitude(rabi_hz):
    return math.sqrt(2) * rabi_hz / GYRO_HZ_PER_T


def _nv7_description(bz=5e-3):
    spins, spin_int, groups = models.nv_7(np.array([0.0, 0.0, bz]), params)
    spins[0][0]["D"] = math.tau * 2.87e9
    spins[1][0]["D"] = math.tau * 1.42e9
    groups[((0, 0), (1, 0))] = {"rel": 65e6}
    groups[((1, 0), (2, 0))] = {"s_gets_0": 8e6, "s_gets_1": 50e6}
    groups[((2, 0), (0, 0))] = {"0_gets_s": 1e6, "1_gets_s": 0.7e6}
    return spins, spin_int, groups


def _nv3_description(bz=5e-3):
    atom = {"S": 1, "g": params.nv.room.ground.g_longitudinal,
            "D": math.tau * 2.87e9, "TS1": 1e-3, "TS2": 1e-6,
            "B0": np.array([0.0, 0.0, bz])}
    return [[atom]], [{}], {}


class Workload:
    """Generators, vectorisation, waveforms and readout of one sweep."""

    def __init__(self, name, description, frequencies, rabi_hz, excitation,
                 time_end, time_step, fine, obs_levels, rho0_diag,
                 use_incoherent=False, number_of_exponentials=2,
                 use_rotating=False):
        self.name = name
        spins, spin_int, groups = description
        ops, vmap = generate_atoms(copy.deepcopy(spins), spin_int, groups,
                                   use_incoherent=use_incoherent)
        g = ops["generators"]
        self.generators = np.array([g["L0"], g["Gx"], g["Gy"], g["Gz"],
                                    g["Gr"]])
        self.vmap = vmap
        self.n = vmap.shape[0]
        self.hilbert = int(max(vmap[:, 0].max(), vmap[:, 1].max())) + 1
        self.frequencies = np.asarray(frequencies, dtype=np.float64)
        self.b1 = drive_amplitude(rabi_hz)
        self.excitation = excitation
        self.time_start = 0.0
        self.time_end = time_end
        self.time_step = time_step
        self.fine = fine
        self.number_of_exponentials = number_of_exponentials
        self.use_rotating = use_rotating
        rho = np.zeros((self.hilbert, self.hilbert))
        for k, v in rho0_diag.items():
            rho[k, k] = v
        self.rho0_matrix = rho
        self.rho0 = self._flat(rho)
        self.obs = np.zeros((1, self.n))
        for k in obs_levels:
            self.obs[0, self._index(k, k)] = 1.0
        self.n_out = int((time_end - self.time_start) / time_step)

    def _index(self, y, x):
        for k, (a, b, c) in enumerate(self.vmap):
            if a == y and b == x and c == 0:
                return k
        raise KeyError((y, x))

    def _flat(self, rho):
        out = np.zeros(self.n)
        for k, (y, x, c) in enumerate(self.vmap):
            out[k] = rho[y, x].real if c == 0 else rho[y, x].imag
        return out

    def waveform(self, frequency):
        """Coefficient rows [1, Bx, By, Bz, r] of one sweep point."""
        w = Waveform(5, base=[1.0, 0.0, 0.0, 0.0, self.excitation])
        w.tone(1, self.b1, frequency)
        return w

    def waveforms(self, count=None):
        freqs = self.frequencies if count is None else \
            self.frequencies[:count]
        return [self.waveform(f) for f in freqs]

    def simulator(self):
        return BatchSimulator(self.generators, self.number_of_exponentials,
                              self.fine)

    def steps_per_instance(self):
        return self.n_out * self.fine


def odmr3(count=256):
    return Workload(
        "odmr3-nv-ground-n9", _nv3_description(),
        np.linspace(2.77e9, 2.97e9, 256)[:count], 5e6, 0.0,
        1e-6, 1e-9, 50, obs_levels=[1], rho0_diag={1: 1.0})


def odmr7(count=4096, use_incoherent=False):
    return Workload(
        "odmr7-nv-n49" if not use_incoherent else "odmr7-nv-n19",
        _nv7_description(), np.linspace(2.77e9, 2.97e9, 4096)[:count],
        3e6, 5e6 / 65e6, 2e-6, 2e-9, 100, obs_levels=[3, 4, 5],
        rho0_diag={0: 1 / 3, 1: 1 / 3, 2: 1 / 3},
        use_incoherent=use_incoherent)
