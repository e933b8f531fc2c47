"""Golden density histories and the JUnit-style check report.

The reference keeps regression fixtures as CSV files of a flattened density
history and checks new runs against them (``test/tests/util.py:12-59``); its
test driver runs every check script and writes ``test-results.xml``
(``test/__main__.py:146-198``).  This module provides the same surface for
this package:

* ``flatten_density`` - every 29th element of the C-ordered
  ``(T, N, N, 2)`` (real, imaginary) expansion of a ``(T, N, N)`` complex
  history (``util.py:12-23``);
* ``save_density`` / ``load_density`` - one CSV file per history, written
  only if it does not exist yet (``util.py:26-40``);
* ``compare_density`` - ``||this - reference||_2 / size`` against a
  precision (default 1e-5, ``util.py:9, 43-59``); a NaN anywhere fails;
* ``run_checks`` / ``write_junit`` - run named check callables, record
  outcome, wall time and output, and write the JUnit XML report.
"""

import contextlib
import io
import os
import time
import traceback
from xml.sax.saxutils import escape, quoteattr

import numpy as np

PRECISION = 1e-5
DELIMITER = ","


class GoldenMismatch(AssertionError):
    """A history differs from its golden file by more than the precision."""


def flatten_density(density):
    """Every 29th (re, im) entry of the C-ordered history."""
    density = np.asarray(density)
    expanded = np.stack([density.real, density.imag], axis=-1)
    return expanded.ravel()[::29]


def _path(name, directory):
    return os.path.join(directory, f"{name}.csv")


def save_density(name, density, directory):
    """Write the golden file of ``name`` unless it already exists."""
    os.makedirs(directory, exist_ok=True)
    path = _path(name, directory)
    if not os.path.exists(path):
        np.savetxt(path, flatten_density(density), delimiter=DELIMITER)
    return path


def load_density(name, directory):
    return np.loadtxt(_path(name, directory), dtype=np.float64,
                      delimiter=DELIMITER)


def density_metric(this, reference):
    """``||this - reference||_2 / this.size`` (NaN gives NaN)."""
    this = np.asarray(this, dtype=np.float64)
    reference = np.asarray(reference, dtype=np.float64)
    if this.shape != reference.shape:
        raise GoldenMismatch(f"history has {this.size} values, the golden "
                             f"file {reference.size}")
    with np.errstate(all="ignore"):
        return float(np.sqrt(np.sum((this - reference) ** 2)) / this.size)


def compare_density(name, density, directory, precision=PRECISION):
    """Raise ``GoldenMismatch`` unless ``density`` matches the golden file
    of ``name`` within ``precision``; returns the metric."""
    path = _path(name, directory)
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    this = flatten_density(density)
    metric = density_metric(this, load_density(name, directory))
    if not np.isfinite(metric) or np.any(np.isnan(this)) or \
            metric > precision:
        raise GoldenMismatch(f"{name}: metric {metric:.3e} > {precision:g}")
    return metric


def run_checks(checks, suite="superspinsim_amd"):
    """Run ``checks`` (an iterable of ``(name, callable)``) and return the
    report dict: ``{suite: {name: {outcome, time, out[, error]}}}``.  A check
    passes when it returns without raising; what it prints is kept."""
    results = {}
    for name, check in checks:
        buf = io.StringIO()
        start = time.time()
        entry = {}
        try:
            with contextlib.redirect_stdout(buf):
                ret = check()
            if ret is not None:
                buf.write(f"{ret}\n")
            entry["outcome"] = "Pass"
        except Exception as exc:  # a failed check is a report entry
            entry["outcome"] = "Fail"
            entry["error"] = "".join(traceback.format_exception_only(
                type(exc), exc)).strip()
        entry["time"] = time.time() - start
        entry["out"] = buf.getvalue()
        results[name] = entry
    return {suite: results}


def write_junit(report, path):
    """JUnit XML with one ``testsuite`` per suite of ``report`` (the layout
    of ``test/__main__.py:146-198``, with text and attributes escaped).
    Returns the number of failed checks."""
    failed = 0
    lines = ["<testsuites>"]
    for suite, results in report.items():
        fails = sum(r["outcome"] == "Fail" for r in results.values())
        failed += fails
        lines.append(f"<testsuite name={quoteattr(str(suite))} "
                     f"tests=\"{len(results)}\" failures=\"{fails}\">")
        for name, r in results.items():
            lines.append(f"<testcase classname={quoteattr(str(suite))} "
                         f"name={quoteattr(str(name))} "
                         f"time=\"{r['time']:.6f}\">")
            if r["outcome"] == "Pass":
                lines.append(f"<system-out>{escape(r.get('out', ''))}"
                             "</system-out>")
            else:
                lines.append(f"<failure>{escape(r.get('error', ''))}"
                             "</failure>")
            lines.append("</testcase>")
        lines.append("</testsuite>")
    lines.append("</testsuites>")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return failed


__all__ = ["PRECISION", "GoldenMismatch", "flatten_density", "save_density",
           "load_density", "density_metric", "compare_density", "run_checks",
           "write_junit"]
