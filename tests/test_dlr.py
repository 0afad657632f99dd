"""DLR grid and operations against closed forms (test/python/pycppdlr.py:60-141,
test/python/volterra_integral.py:12-34) — these pin the restated cppdlr layer of the oracle and the
product's node construction."""
import numpy as np
import pytest

import pyoracle as po
from xca_amd.dlr import DLRGrid, kernel, rel2abs


def free_g(tau, beta, e):
    """-K(tau/beta, beta e), the free propagator of a level at energy e (pycppdlr.py:31-51)."""
    return kernel(np.asarray(tau) / beta - (np.asarray(tau) / beta > 0.5), beta * e)


@pytest.mark.parametrize("lam,eps", [(200.0, 1e-8), (1000.0, 1e-10), (2000.0, 1e-10)])
def test_interpolation_and_reflection(lam, eps):
    beta = 13.0
    g = DLRGrid(beta, lam, eps)
    e = 1.0
    vals = free_g(g.tau, beta, e)[:, None, None]
    tf = np.linspace(0, 1, 101)
    trel = np.where(tf > 0.5, tf - 1, tf)
    got = po.interp(g.it, g.rf, vals, trel)[:, 0, 0].real
    ref = kernel(trel, beta * e)
    assert np.abs(got - ref).max() < 5 * eps
    refl = po.reflect(g.it, g.rf, vals)[:, 0, 0].real
    assert np.abs(refl - kernel(g.it, -beta * e)).max() < 5 * eps


@pytest.mark.parametrize("lam,eps", [(1.0, 1e-12), (100.0, 1e-10), (1000.0, 1e-10), (10000.0, 1e-12)])
def test_volterra_constant(lam, eps):
    """(-1) * (-1) time-ordered = tau (volterra_integral.py:21-31), error well below eps."""
    beta = 1.0
    g = DLRGrid(beta, lam, eps)
    G = -np.ones((g.r, 1, 1))
    GG = po.convolve(beta, g.it, g.rf, G, G)[:, 0, 0].real
    assert np.abs(GG - g.tau).max() < 1e-2 * eps * 10


@pytest.mark.parametrize("a,b", [(0.3, 1.7), (-2.0, 0.5), (4.0, 4.0)])
def test_convolution_of_exponentials(a, b):
    """int_0^tau exp(-a (tau - s)) exp(-b s) ds = (exp(-b tau) - exp(-a tau))/(a - b)."""
    beta = 5.0
    g = DLRGrid(beta, 100.0, 1e-12)
    f = np.exp(-a * g.tau)[:, None, None]
    h = np.exp(-b * g.tau)[:, None, None]
    got = po.convolve(beta, g.it, g.rf, f, h)[:, 0, 0].real
    ana = g.tau * np.exp(-a * g.tau) if a == b else (np.exp(-b * g.tau) - np.exp(-a * g.tau)) / (a - b)
    assert np.abs(got - ana).max() < 1e-11 * max(1.0, np.abs(ana).max())  # eps = 1e-12 grid


def test_matrix_convolution_order():
    """Matrix-valued convolution keeps f on the left: (f * g)_ij = sum_k f_ik * g_kj."""
    beta = 3.0
    g = DLRGrid(beta, 60.0, 1e-12)
    rng = np.random.default_rng(1)
    A = rng.normal(size=(2, 2))
    B = rng.normal(size=(2, 3))
    f = np.exp(-0.7 * g.tau)[:, None, None] * A[None]
    h = np.exp(-1.1 * g.tau)[:, None, None] * B[None]
    got = po.convolve(beta, g.it, g.rf, f, h).real
    s = (np.exp(-1.1 * g.tau) - np.exp(-0.7 * g.tau)) / (0.7 - 1.1)
    assert np.abs(got - s[:, None, None] * (A @ B)[None]).max() < 1e-10


def test_rank_close_to_reference():
    """r at (Lambda, eps) = (200, 1e-10) within one of the reference's r = 36
    (test_block_sparse_backbone_eval.cpp:50)."""
    assert abs(DLRGrid(1.0, 200.0, 1e-10).r - 36) <= 1


def test_nodes_golden():
    """Node construction is deterministic: nodes for the benchmark grids are frozen fixtures."""
    import json
    from pathlib import Path
    gold = json.loads((Path(__file__).parent / "golden" / "dlr_nodes.json").read_text())
    for key, v in gold.items():
        lam, eps = map(float, key.split(","))
        g = DLRGrid(1.0, lam, eps)
        assert g.r == len(v["rf"])
        np.testing.assert_array_equal(g.rf, np.array(v["rf"]))
        np.testing.assert_array_equal(g.it, np.array(v["it"]))
