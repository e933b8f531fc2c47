"""Physical constants used by the propagator path, as class namespaces.

The namespace layout (``fundamental``, ``standards.lab.ntp``,
``nv.room.{optical,ground,excited,isc}``) and the values follow
``superspinsim/params.py`` (fundamental :5-116, standards :117-153,
nv.room :208-1316) so that ``models.nv_7(params=...)`` accepts either module.
Only the constants the spin models read are restated; the Rb-87, sensing and
point-group tables of the reference are not on the simulation path.
Units: Hz for rates and frequencies, Hz/T for gyromagnetic ratios, s, K.
"""


class fundamental:
    planck = 6.62607015e-34
    boltzmann = 1.3806489999999998e-23
    # Bohr magneton / h (Hz/T), params.py:44
    bohr_gyro = 13996244917.1
    # nuclear magneton / h (Hz/T), params.py:63
    nuclear_gyro = 7622593.2188
    # k_B / h (Hz/K), params.py:82
    boltzmann_gyro = 20836619120.0

    class electron:
        g = 2.00231930436092


class standards:
    class lab:
        class ntp:
            temperature = 293.15
            pressure = 101325.0


class nv:
    class room:
        class optical:
            conserving = 66000000.0
            nonconserving = 3547020.0564265894
            ratio = 0.01
            life_total = 1.3e-08

        class ground:
            thermalisation_time = 0.01
            dephasing_time = 9.999999999999999e-05
            thermalisation_time_14N = 10.0
            dephasing_time_14N = 9.999999999999999e-06
            gyro_longitudinal = 28000000000.0
            zfs_longitudinal = 2872000000.0
            gyro_14N_longitudinal = -3076999.9999999995
            nuclear_quadrupole_14N_longitudinal = -4945000.0
            hyperfine_14N_longitudinal = -2162000.0
            hyperfine_14N_transverse = -2620000.0
            g_longitudinal = 2.000536581621305
            g_transverse = 2.0031
            g_N_14N_longitudinal = 0.40366839834155777

        class excited:
            thermalisation_time = 0.001
            dephasing_time = 1e-08
            thermalisation_time_14N = 0.1
            dephasing_time_14N = 0.001
            zfs_longitudinal = 1423000000.0
            g_longitudinal = 2.0019655363224635
            hyperfine_14N_transverse = -23000000.0
            nuclear_quadrupole_14N_longitudinal = -4945000.0
            gyro_14N_longitudinal = -3080000.0
            gyro_longitudinal = 28020000000.0
            hyperfine_14N_longitudinal = -40000000.0
            g_N_14N_longitudinal = 0.4040619651907696

        class isc:
            s_gets_pm = 60400000.0
            s_gets_z = 9390000.0
            z_gets_s = 9600000.0
            pm_gets_s = 2110000.0
