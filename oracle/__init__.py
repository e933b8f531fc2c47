"""Oracle: CPU restatement of the reference propagator, for testing only.

Everything under ``oracle/`` is test infrastructure.  It is imported by
``tests/``, by ``__graft_entry__.smoke()`` (as the checker) and by the CPU
baseline leg of ``bench.py``; the product package ``superspinsim_amd`` never
imports, links or runs it.

* ``reference_simulate`` - numpy restatement of ``generate_simulator`` /
  ``simulate`` (``superspinsim/generate_simulator.py:17-2202``), following the
  reference's own algorithm step by step: parallel-in-time over output
  intervals, first-order Taylor step scaled by 4^-R, 2R residual squarings,
  commutator-free product, rotating-frame eigen-step, serial prefix product,
  global sandwiches and application to rho0.
* ``cpu_ref.c`` - the same algorithm in C with OpenMP (built into
  ``oracle/_build/libcpuref.so`` by ``oracle/Makefile``); used for the larger
  parity cases and as the timed CPU baseline.
* ``brute`` - brute-force checks from the mathematics: scipy ``expm`` per
  step and ``solve_ivp`` (DOP853) on the same vectorised Lindbladian.
"""
