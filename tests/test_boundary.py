"""The reference's error contract and input invariants at the C-ABI boundary, host side only.

xca_ctx_create validates the problem (extents, symmetry-set coupling) before it touches a device, so
every check here runs without a GPU: a rejected problem returns NULL with the reference's message, an
accepted one gets as far as device selection.
"""
import numpy as np
import pytest

import pyoracle as po
from xca_amd import models
from xca_amd._lib import XcaError
from xca_amd.atom import AtomDiagArrays
from xca_amd.dlr import DLRGrid
from xca_amd.evaluator import DiagramEvaluator

NO_GPU = "no GPU device available"


def _create(P, poles=None, coeffs=None, ad=None, grid=None):
    """DiagramEvaluator construction through xca_ctx_create; returns the error text ('' if a context
    was created)."""
    try:
        ev = DiagramEvaluator(P.poles if poles is None else poles, P.coeffs if coeffs is None else coeffs, P.grid if grid is None else grid,
                              P.ad if ad is None else ad)
    except XcaError as e:
        return str(e)
    del ev
    return ""


def _accepted(msg):
    # the host checks passed: either a context exists (GPU present) or device selection failed
    return msg == "" or NO_GPU in msg


@pytest.fixture(scope="module")
def two_band():
    """Two-orbital Kanamori atom in (N, S_z) blocks: symmetry sets {down 0, down 1} and {up 0, up 1},
    with within-set off-diagonal hybridization."""
    return models.kanamori(2, 4.0, 0.5, 4.75, beta=5.0, lam=50.0, eps=1e-10, offdiag=0.3)


def test_valid_problem_passes_host_checks(two_band):
    assert _accepted(_create(two_band))


def test_cross_set_coefficients_rejected(two_band):
    """test_block_sparse_sym_set_mask.cpp:172-200, test_block_sparse_sym_sets.cpp:87: a coefficient
    that couples different symmetry sets is rejected with a message naming the coupling."""
    c = two_band.coeffs.copy()
    c[:, 0, 2] = 0.25 * np.abs(c).max()
    msg = _create(two_band, coeffs=c)
    assert "couples different symmetry sets" in msg, msg
    assert "symmetry set" in msg


def test_cross_set_below_tolerance_accepted(two_band):
    """0.9 * 1e-12 * max|hyb_coeffs| is on the accepted side of the threshold
    (test_block_sparse_sym_set_mask.cpp:144-147)."""
    c = two_band.coeffs.copy()
    c[:, 0, 2] = c[:, 2, 0] = 0.9e-12 * np.abs(c).max()
    assert _accepted(_create(two_band, coeffs=c))


def test_coefficient_extent_rejected(two_band):
    """test_block_sparse_sym_set_mask.cpp:207-225: (p, n-1, n-1) coefficients -> the extent check,
    whose message names "shape" and "sum(Fq.sym_set_sizes)"."""
    P = two_band
    msg = _create(P, coeffs=P.coeffs[:, :-1, :-1].copy())
    assert "shape" in msg and "sum(Fq.sym_set_sizes)" in msg, msg
    assert "couples" not in msg


def test_pole_extent_rejected(two_band):
    """test_block_sparse_sym_set_mask.cpp:227-238: (p+1, n, n) coefficients against p poles -> the
    extent check naming "hyb_poles.size()"."""
    P = two_band
    c = np.concatenate([P.coeffs, P.coeffs[:1]], axis=0)
    msg = _create(P, coeffs=c)
    assert "shape" in msg and "hyb_poles.size()" in msg, msg


def test_operator_block_shape_rejected(two_band):
    """A c block of the wrong shape is rejected before any device work."""
    ad = AtomDiagArrays.from_atom_diag(two_band.ad)
    o, j = next(k for k in ad.c_blocks)
    bad = dict(ad.c_blocks)
    bad[(o, j)] = np.zeros((ad.c_blocks[(o, j)].shape[0] + 1, ad.c_blocks[(o, j)].shape[1]))
    with pytest.raises(ValueError, match="shape"):
        AtomDiagArrays(ad.fs.n, ad.dims, ad.c_conn, ad.cdag_conn, bad, ad.cdag_blocks)


def test_sub_tolerance_cross_set_leaves_sigma_bitwise_unchanged(two_band):
    """test_block_sparse_sym_set_mask.cpp:107-170 on the oracle: a cross-set coefficient at
    0.9 * 1e-12 * max|hyb_coeffs| leaves Sigma exactly unchanged (the (p_kap, p_mu) loop runs over
    p_mu == p_kap only), while the same perturbation inside a set moves it."""
    P = two_band
    G = P.G0
    t = np.array([[0, 2], [1, 3]], dtype=np.int32)
    delta = 0.9e-12 * np.abs(P.coeffs).max()

    def sigma(c):
        Q = models.ImpurityProblem(grid=P.grid, ad=P.ad, poles=P.poles, coeffs=c, G0=P.G0, eta0=P.eta0)
        return po.OracleEvaluator(Q, complex_arith=False).sigma(G, t)

    s0 = sigma(P.coeffs)
    assert max(np.abs(x).max() for x in s0) > 1e-3
    cross = P.coeffs.copy()
    cross[:, 0, 2] += delta
    cross[:, 3, 1] += delta
    assert all(np.array_equal(a, b) for a, b in zip(sigma(cross), s0))
    within = P.coeffs.copy()
    within[:, 0, 1] += delta
    assert max(np.abs(a - b).max() for a, b in zip(sigma(within), s0)) > 0.0


def test_dlr_grid_from_nodes_round_trip():
    g = DLRGrid(10.0, 100.0, 1e-10)
    # the reference's mesh gives tau / beta in [0, 1]; relative format is recovered
    h = DLRGrid.from_nodes(10.0, np.where(g.it < 0, 1.0 + g.it, g.it), g.rf, lam=100.0, eps=1e-10)
    assert h.r == g.r
    assert np.allclose(h.it, g.it, rtol=0, atol=1e-15)
    assert np.array_equal(h.rf, g.rf)
    assert np.allclose(h.tau, g.tau, rtol=1e-15, atol=1e-14)
    # interpolation through either grid (the coefficients themselves are ill-conditioned in the nodes)
    x = -np.exp(-0.3 * g.tau) / (1 + np.exp(-3.0))
    ts = np.linspace(0.0, 1.0, 7)
    exact = -np.exp(-0.3 * 10.0 * ts) / (1 + np.exp(-3.0))
    assert np.abs(h.eval(h.vals2coefs(x), ts) - exact).max() < 1e-9
    assert np.abs(h.eval(h.vals2coefs(x), ts) - g.eval(g.vals2coefs(x), ts)).max() < 1e-10
    with pytest.raises(ValueError, match="relative"):
        DLRGrid.from_nodes(10.0, g.tau + 2.0, g.rf)
    with pytest.raises(ValueError, match="equal length"):
        DLRGrid.from_nodes(10.0, g.it[:-1], g.rf)


def test_atom_diag_arrays_from_triqs_layout(two_band):
    """AtomDiagArrays.from_triqs reads the attributes of a TRIQS atom_diag that get_operators_impl
    reads (fops, fock_states, c_connection, cdag_connection; atom_diag_utils.cpp:39-148). A stand-in
    exposing them for the two-band atom gives the same operators as the builder, and an evaluator
    accepts it."""
    ad = two_band.ad

    class TriqsLike:
        fops = [("f", o) for o in range(ad.fs.n)]
        fock_states = [np.asarray(f) for f in ad.fock_states]
        c_connection = ad.c_conn.astype(np.int64)
        cdag_connection = ad.cdag_conn.astype(np.int64)
        n_subspaces = ad.nblocks

    arr = AtomDiagArrays.from_triqs(TriqsLike())
    assert np.array_equal(arr.dims, ad.dims)
    for a, b in zip(arr.packed_ops(), ad.packed_ops()):
        assert np.array_equal(a, b)
    assert _accepted(_create(two_band, ad=arr))
    # a grid built from the reference's node arrays is accepted as well
    grid = DLRGrid.from_nodes(two_band.beta, two_band.grid.it, two_band.grid.rf)
    assert _accepted(_create(two_band, grid=grid))
