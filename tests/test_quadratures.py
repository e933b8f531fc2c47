"""CF/GL tables: reference literals, Gauss-Legendre roots, order conditions."""

import json
import os

import numpy as np
import pytest

from superspinsim_amd import quadratures

FIXTURE = os.path.join(os.path.dirname(__file__), "golden",
                       "quadratures.json")


def test_tables_equal_reference_literals():
    ref = json.load(open(FIXTURE))
    for key, vals in ref["samples"].items():
        assert np.array_equal(quadratures.samples[key], np.array(vals))
    for key, vals in ref["weights"].items():
        assert np.array_equal(quadratures.weights[key], np.array(vals))


@pytest.mark.parametrize("points", [1, 2, 3])
def test_nodes_are_gauss_legendre(points):
    x, _ = np.polynomial.legendre.leggauss(points)
    nodes = quadratures.samples[f"{points}_gl"]
    assert np.allclose(np.sort(nodes), (x + 1) / 2, atol=2e-16)


@pytest.mark.parametrize("j", [1, 2, 3, 5, 6])
def test_weights_sum_to_one(j):
    """Time-independent generator: the CF product is exp(h A) exactly."""
    _, w, _ = quadratures.scheme(j)
    assert abs(w.sum() - 1.0) < 1e-15


@pytest.mark.parametrize("j,order", [(2, 4), (3, 4), (5, 6), (6, 6)])
def test_cf_order_conditions(j, order):
    """Scalar linear test: sum_j sum_s w[j,s] c_s^k = 1/(k+1) for k < order.

    Exact for the time-polynomial part of commuting problems; a dropped
    weight or swapped node fails it.
    """
    nodes, w, _ = quadratures.scheme(j)
    for k in range(order):
        assert abs(np.sum(w * nodes[None, :] ** k) - 1 / (k + 1)) < 1e-14


def test_unknown_scheme_raises():
    with pytest.raises(ValueError):
        quadratures.scheme(4)
