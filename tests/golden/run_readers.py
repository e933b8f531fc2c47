"""Run the reference's golden checks through the drop-in and write JUnit.

The counterpart of the reference's ``test/__main__.py``: every writer and
reader recipe of ``tests/golden/cases.py`` (the reference's
``test/tests/*.py`` scripts) runs through
``superspinsim_amd.generate_simulator`` on the GPU over the whole 10 us
history, is compared with its golden CSV by the reference's metric
(``superspinsim_amd.golden.compare_density``, 1e-5), and the outcome goes to
a JUnit XML report.

    python tests/golden/run_readers.py [report.xml]
"""

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

from golden.cases import CSV_DIR, GOLDEN, READERS, case  # noqa: E402
from superspinsim_amd import golden  # noqa: E402
from superspinsim_amd.generate_simulator import (  # noqa: E402
    generate_simulator)
from superspinsim_amd.sugar import prepare_simspins  # noqa: E402


def _check(golden_name, flags):
    def run():
        coeffs, spins, si, gi, rho = case(golden_name)
        sampler, gens, vmap, kw = prepare_simspins(coeffs, spins, si, gi,
                                                   **flags)
        _, dens = generate_simulator(sampler, gens, vmap, **kw)(
            rho, 0, 10e-6, 1e-9)
        metric = golden.compare_density(golden_name, dens, CSV_DIR)
        print(f"{golden_name}: reference metric {metric:.3e}")
    return run


def main(path="test-results.xml"):
    checks = [(name, _check(name, flags)) for name, flags in
              sorted(GOLDEN.items())]
    checks += [(name, _check(g, flags)) for name, (g, flags) in
               sorted(READERS.items())]
    suite = "superspinsim_amd golden"
    report = {suite: {}}
    for check in checks:
        one = golden.run_checks([check], suite=suite)[suite]
        report[suite].update(one)
        for name, r in one.items():
            print(f"{name}: {r['outcome']} ({r['time']:.1f} s) "
                  f"{r.get('error', r['out'].strip())}", flush=True)
    failed = golden.write_junit(report, path)
    print(f"{len(checks) - failed} passed, {failed} failed -> {path}")
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main(*sys.argv[1:2]))
