"""The golden-history format and JUnit report (superspinsim_amd.golden).

The seven golden CSVs are the reference's own files: flattening the oracle's
history with ``flatten_density`` and comparing with ``compare_density``
must pass, a perturbed or NaN history must fail, a round trip through
``save_density``/``load_density`` must be exact, and ``write_junit`` must
write well-formed XML with the right counts.
"""

import math
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from golden.cases import CSV_DIR, GOLDEN, case
from oracle.reference_simulate import generate_simulator as oracle_gen
from superspinsim_amd import golden
from superspinsim_amd.sugar import prepare_simspins


def _oracle_history(name):
    coeffs, spins, si, gi, rho = case(name)
    sampler, gens, vmap, kw = prepare_simspins(coeffs, spins, si, gi,
                                               **GOLDEN[name])
    return oracle_gen(sampler, gens, vmap, backend="c", **kw)(
        rho, 0, 10e-6, 1e-9)[1]


def test_flatten_matches_the_reference_layout():
    d = np.arange(2 * 3 * 3).reshape(2, 3, 3) + 1j * np.ones((2, 3, 3))
    flat = golden.flatten_density(d)
    expanded = np.zeros((2, 3, 3, 2))
    expanded[..., 0] = d.real
    expanded[..., 1] = d.imag
    assert np.array_equal(flat, expanded.flatten()[::29])


def test_oracle_history_passes_the_golden_check():
    dens = _oracle_history("qubit-t1-mK")
    assert golden.compare_density("qubit-t1-mK", dens, CSV_DIR) < 1e-12


def test_compare_fails_on_mismatch_and_nan():
    dens = _oracle_history("qubit-t1-mK")
    bad = dens.copy()
    bad[5000:] *= 1.01
    with pytest.raises(golden.GoldenMismatch):
        golden.compare_density("qubit-t1-mK", bad, CSV_DIR)
    bad = dens.copy()
    bad[0, 0, 0] = np.nan             # flat index 0 is sampled
    with pytest.raises(golden.GoldenMismatch):
        golden.compare_density("qubit-t1-mK", bad, CSV_DIR)
    with pytest.raises(golden.GoldenMismatch):
        golden.compare_density("qubit-t1-mK", dens[:100], CSV_DIR)


def test_save_load_round_trip_and_no_overwrite(tmp_path):
    d = np.exp(1j * np.linspace(0, 3, 300)).reshape(25, 2, 2, 3)[..., 0]
    path = golden.save_density("h", d, str(tmp_path))
    back = golden.load_density("h", str(tmp_path))
    assert np.array_equal(back, golden.flatten_density(d))
    golden.save_density("h", 2 * d, str(tmp_path))   # keeps the first file
    assert np.array_equal(np.loadtxt(path, delimiter=","), back)
    assert golden.compare_density("h", d, str(tmp_path)) == 0.0


def test_junit_report(tmp_path):
    def ok():
        print("history within 1e-5")

    def bad():
        raise golden.GoldenMismatch("x: metric 1e-3 > 1e-05 <&>")

    report = golden.run_checks([("ok", ok), ("bad", bad)], suite="readers")
    path = tmp_path / "test-results.xml"
    assert golden.write_junit(report, str(path)) == 1
    root = ET.parse(str(path)).getroot()
    suite = root.find("testsuite")
    assert suite.get("tests") == "2" and suite.get("failures") == "1"
    cases = {c.get("name"): c for c in suite.findall("testcase")}
    assert "history within 1e-5" in cases["ok"].find("system-out").text
    assert "<&>" in cases["bad"].find("failure").text
    assert math.isfinite(float(cases["ok"].get("time")))
