"""Spin descriptions of NV-centre models.

``nv_7`` follows ``superspinsim/models.py:7-87``: ground triplet, excited
triplet and metastable singlet as three incoherently coupled blocks with
spin-conserving (and non-conserving) optical transitions and the ISC paths.
Basis order: ground m=+1,0,-1 (0..2), excited m=+1,0,-1 (3..5), singlet (6).

``nv_21`` has no counterpart in the reference models module; it is the same
three blocks with a 14N nucleus (I = 1, quadrupole P, hyperfine A) added to
every atom, i.e. 21 levels.
"""

import math

import numpy as np

from superspinsim_amd import params as s3p


def nv_7(quiescent_magnetic_field: np.ndarray = None, params=None):
    """Spin description of the 7-level NV model (models.py:7-87).

    Returns ``(spins, spin_interactions, group_interactions)`` for
    ``simspins``.
    """
    if quiescent_magnetic_field is None:
        quiescent_magnetic_field = np.zeros(3)
    if params is None:
        params = s3p
    room = params.nv.room
    ground = {
        "S": 1,
        "g": room.ground.g_longitudinal,
        "g_perp": room.ground.g_transverse,
        "D": math.tau * room.ground.zfs_longitudinal,
        "TS1": room.ground.thermalisation_time,
        "TS2": room.ground.dephasing_time,
        "B0": quiescent_magnetic_field,
        "T": params.standards.lab.ntp.temperature,
    }
    excited = {
        "S": 1,
        "g": room.excited.g_longitudinal,
        "D": math.tau * room.excited.zfs_longitudinal,
        "TS1": room.excited.thermalisation_time,
        "TS2": room.excited.dephasing_time,
        "B0": quiescent_magnetic_field,
        "T": params.standards.lab.ntp.temperature,
    }
    singlet = {"S": 0}
    orbitals = {
        ((0, 0), (1, 0)): {"rel": room.optical.conserving,
                           "rel_n": room.optical.nonconserving},
        ((1, 0), (2, 0)): {"s_gets_0": room.isc.s_gets_z,
                           "s_gets_1": room.isc.s_gets_pm},
        ((2, 0), (0, 0)): {"0_gets_s": room.isc.z_gets_s,
                           "1_gets_s": room.isc.pm_gets_s},
    }
    return [[ground], [excited], [singlet]], [{}, {}, {}], orbitals


def nv_21(quiescent_magnetic_field: np.ndarray = None, params=None,
          quadrupole=None, hyperfine_longitudinal=None,
          hyperfine_transverse=None, nuclear_g=None):
    """7-level NV with a 14N nucleus on every orbital block (21 levels).

    Frequencies are in Hz and converted to rad/s here.  Defaults come from
    ``params.nv.room.ground`` (quadrupole, A_par, A_perp, nuclear g).  The
    nuclear g factor is expressed in nuclear magnetons, as the reference's
    ``gN`` key expects.
    """
    spins, interactions, orbitals = nv_7(quiescent_magnetic_field, params)
    if params is None:
        params = s3p
    g = params.nv.room.ground
    quad = g.nuclear_quadrupole_14N_longitudinal if quadrupole is None \
        else quadrupole
    a_par = g.hyperfine_14N_longitudinal if hyperfine_longitudinal is None \
        else hyperfine_longitudinal
    a_perp = g.hyperfine_14N_transverse if hyperfine_transverse is None \
        else hyperfine_transverse
    g_n = g.g_N_14N_longitudinal if nuclear_g is None else nuclear_g
    for block in spins:
        atom = block[0]
        atom["I"] = 1
        atom["gN"] = g_n
        atom["P"] = math.tau * quad
        if atom.get("S", 0) > 0:
            atom["A"] = math.tau * a_par
            atom["A_perp"] = math.tau * a_perp
        if "B0" not in atom:
            atom["B0"] = spins[0][0]["B0"]
    return spins, interactions, orbitals
