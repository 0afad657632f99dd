"""Oracle pins of the single-particle Green's function and the one-time correlator.

Closed forms held by the reference's tests:
- test/c++/test_block_sparse_backbone_gf_eval.cpp:35-130 (one fermion, constant and one-pole Delta);
- test/c++/test_block_sparse_two_fermions.cpp:101-163 (two fermions, constant Delta);
- test/python/analytic_diagrams.py:96-122 (n fermions, H = 0: g summed over orders with the solver's
  (-1)^order * s_topo, block_sparse_solver.py:657-677);
- test/python/one_spinful_fermion.py:67-118,212-244 (NESSi g(tau) on 801 points after atom, NCA and OCA
  self-consistency; the reference asserts 6 decimals).
Invariants: f_ix decomposition and operator ordering (test_block_sparse_sparsity_invariance.cpp:40-135),
one block == blocks by N."""
import json
from pathlib import Path

import numpy as np
import pytest

import pyoracle as po
from xca_amd import models
from xca_amd.dlr import rel2abs

GOLDEN = Path(__file__).resolve().parent / "golden"
T3 = [[0, 3], [1, 4], [2, 5]]


def test_one_fermion_const_hyb():
    beta, eps = 2.0, 1e-10
    P = models.one_fermion(beta=beta, lam=20 * beta, eps=eps)
    O = po.OracleEvaluator(P)
    t = rel2abs(P.grid.it)
    assert np.abs(O.spgf(P.G0, [[0, 1]])[:, 0, 0] - 0.5).max() <= eps
    assert np.abs(O.spgf(P.G0, [[0, 2], [1, 3]])).max() <= eps
    ana = (beta / 2) ** 4 * (1 - t) ** 2 * t ** 2 / 2
    assert np.abs(O.spgf(P.G0, T3)[:, 0, 0] - ana).max() <= eps


def test_one_fermion_one_pole():
    beta, eps, om = 1.0, 1e-10, 0.8
    P = models.one_fermion(beta=beta, lam=20 * beta, eps=eps, pole=om)
    O = po.OracleEvaluator(P)
    t = rel2abs(P.grid.it)
    assert np.abs(O.spgf(P.G0, [[0, 1]])[:, 0, 0] - 0.5).max() <= eps
    assert np.abs(O.spgf(P.G0, [[0, 2], [1, 3]])).max() <= eps
    ana = (t + (np.exp(-om * t) - 1) / om) * (t - beta + (np.exp(om * (beta - t)) - 1) / om) \
        / (2 * om * om * (1 + np.exp(-beta * om)) * (np.exp(beta * om) + 1))
    assert np.abs(O.spgf(P.G0, T3)[:, 0, 0] - ana).max() <= eps


def test_two_fermions_const_hyb():
    beta, eps = 2.0, 1e-12
    P = models.two_fermion(beta=beta, lam=20 * beta, eps=eps, U=0.0, mu=0.0)
    O = po.OracleEvaluator(P)
    t = rel2abs(P.grid.it)
    assert np.abs(O.spgf(P.G0, [[0, 1]]) - 0.5 * np.eye(2)[None]).max() <= eps
    oca = O.spgf(P.G0, [[0, 2], [1, 3]])
    ana = 0.25 * (beta * beta * t - beta * beta * t * t)
    for i in range(2):
        assert np.abs(oca[:, i, i] - ana).max() <= 10 * eps
    third = O.spgf(P.G0, T3)
    ana3 = 0.5 * beta ** 4 * t * t * (-t + 0.5 * (1 + t * t)) / 16.0
    for i in range(2):
        assert np.abs(third[:, i, i] - 2 * ana3).max() <= eps


@pytest.mark.parametrize("nflav", [1, 2, 3])
@pytest.mark.parametrize("qn", ["none", "N"])
def test_n_fermions_h0_orders(nflav, qn):
    """g_NCA = -1/2, dg_OCA = -1/2 (beta - tau) tau (n - 1)/2, dg_TCA = -(1/2)^2 (c1 + c2 + c3) with
    the combinatorial factors C_1 = (n-2)n + (n-1)(n-2), C_2 = C_3 = (n-1)^2 (analytic_diagrams.py:108-126)."""
    beta = 2.0
    P = models.n_fermion_h0(nflav, beta=beta, lam=10 * beta, eps=1e-12, qn=qn)
    O = po.OracleEvaluator(P)
    tau = P.grid.tau[:, None, None]
    n, I = nflav, np.eye(nflav)[None]
    g_nca = -0.5 * I + 0 * tau
    dg_oca = -0.5 * (beta - tau) * tau * (n - 1) / 2 * I
    C1, C2 = 0.5 * ((n - 2) * n + (n - 1) * (n - 2)), 0.5 * (n - 1) ** 2
    c1 = (beta - tau) ** 2 * tau ** 2 / 4 * C1
    c2 = (beta - tau) ** 3 * tau / 6 * C2
    c3 = (beta - tau) * tau ** 3 / 6 * C2
    dg_tca = -0.25 * (c1 + c2 + c3) * I
    for order, ana in ((1, g_nca), (2, g_nca + dg_oca), (3, g_nca + dg_oca + dg_tca)):
        g = po.spgf_sum(O, P.G0, order)
        assert np.abs(g - ana).max() < 1e-11, (order, np.abs(g - ana).max())


@pytest.fixture(scope="module")
def nessi():
    return json.loads((GOLDEN / "nessi_one_spinful_fermion.json").read_text())


@pytest.mark.parametrize("case,V,order", [("atom", 0.0, 1), ("nca", 0.1, 1), ("oca", 0.1, 2)])
def test_one_spinful_fermion_matches_nessi(nessi, case, V, order):
    """Self-consistent g(tau) of one spinful level against the NESSi reference on 801 equispaced tau;
    the reference test asserts 6 decimals, the oracle meets 1e-10."""
    P = models.one_spinful_fermion(V=V)
    d = nessi[case]
    ts = np.array(d["tau"]) / P.beta
    ts[-1] = 1.0  # tau = beta in relative time
    delta = V * V * np.array([models.kernel(t, P.beta * 0.1) for t in ts])
    assert np.abs(delta - np.array(d["delta_tau"])).max() < 1e-12
    G, eta, it, diff = po.ppsc_solve(P, order, tol=1e-11, maxiter=100)
    assert diff < 1e-11
    g = po.spgf_sum(po.OracleEvaluator(P), G, order)
    gf = po.interp(P.grid.it, P.grid.rf, g, ts)
    assert np.abs(gf[:, 0, 0].real - np.array(d["g_tau"])).max() < 1e-10
    assert np.abs(gf[:, 1, 1].real - np.array(d["g_tau"])).max() < 1e-10
    assert np.abs(gf[:, 0, 1]).max() == 0.0


def _orbital_ops(prob, dag):
    """c_o (dag False) or c^dag_o per orbital as (conn, packed matrices)."""
    ad = prob.ad
    conn = np.asarray(ad.cdag_conn if dag else ad.c_conn, dtype=np.int32)
    cm, dm = ad.packed_ops()
    return conn, (dm if dag else cm)


def test_fix_decomposition_and_explicit_operators():
    P = models.two_band_discrete_bath(beta=2.0, lam=20.0, eps=1e-8)
    O = po.OracleEvaluator(P)
    G = P.perturbed_G()
    topo = [[0, 2], [1, 3]]
    full = O.spgf(G, topo)
    scale = np.abs(full).max()
    assert scale > 1e-3
    nf = O.num_correlator_backbones(2)
    assert nf == 2 * 4 * len(P.poles)
    parts = sum(O.spgf(G, topo, fix=[f]) for f in range(nf))
    assert np.abs(parts - full).max() <= 1e-13 * scale
    halves = O.spgf(G, topo, fix=np.arange(0, nf, 2)) + O.spgf(G, topo, fix=np.arange(1, nf, 2))
    assert np.abs(halves - full).max() <= 1e-13 * scale
    with pytest.raises(po.OracleError, match="f_ix out of range"):
        O.spgf(G, topo, fix=[nf])
    # explicit c / c^dag operator lists give the same result, and a flavour permutation of them
    # permutes rows and columns of the correlator
    mu, kap = _orbital_ops(P, False), _orbital_ops(P, True)
    expl = O.correlator(G, topo, mu_ops=mu, kap_ops=kap)
    assert np.abs(expl - full).max() <= 1e-14 * scale
    perm = [2, 0, 3, 1]
    ad = P.ad
    cm, dm = ad.packed_ops()

    def permuted(conn, mats):
        # split the packed matrices per operator, then reorder operators
        sizes = [sum(P.dims[t] * P.dims[j] for j, t in enumerate(row) if t >= 0) for row in conn]
        offs = np.concatenate([[0], np.cumsum(sizes)])
        return conn[perm], np.concatenate([mats[offs[o]:offs[o + 1]] for o in perm])

    pc = O.correlator(G, topo, mu_ops=permuted(*mu), kap_ops=permuted(*kap))
    assert np.abs(pc - full[:, perm][:, :, perm]).max() <= 1e-14 * scale
    # and the permutation is visible (the test is not vacuous)
    assert np.abs(full[:, perm][:, :, perm] - full).max() > 1e-3 * scale


@pytest.mark.parametrize("order", [1, 2])
def test_one_block_equals_blocks_by_n(order):
    """The correlator does not depend on the block partition (test_block_sparse_sparsity_invariance.cpp)."""
    A = models.two_band_discrete_bath(beta=2.0, lam=20.0, eps=1e-8, qn="N")
    B = models.two_band_discrete_bath(beta=2.0, lam=20.0, eps=1e-8, qn="none")
    ga = po.spgf_sum(po.OracleEvaluator(A), A.G0, order)
    gb = po.spgf_sum(po.OracleEvaluator(B), B.G0, order)
    assert np.abs(ga).max() > 0.1
    assert np.abs(ga - gb).max() < 1e-12
