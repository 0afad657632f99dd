"""Oracle pins against the closed forms the reference's own tests hold.

test/c++/test_block_sparse_backbone_eval.cpp:142-269 (one fermion), test_block_sparse_two_fermions.cpp:34-94
(two fermions), test/python/analytic_diagrams.py:82-122 (n fermions, H = 0, full Sigma with the solver's
order/topology signs, block_sparse_solver.py:609-638)."""
import numpy as np
import pytest

import pyoracle as po
from xca_amd import models
from xca_amd.diag import all_connected_pairings
from xca_amd.dlr import rel2abs

LN2 = np.log(2.0)


def test_one_fermion_const_hyb():
    beta, eps = 2.0, 1e-10
    P = models.one_fermion(beta=beta, lam=20 * beta, eps=eps, pole=0.0)
    t = rel2abs(P.grid.it)
    G0 = -np.exp(-t * LN2)
    for b in range(2):
        assert np.abs(P.G0[b][:, 0, 0] - G0).max() <= eps
    O = po.OracleEvaluator(P)
    nca = O.sigma(P.G0, [[0, 1]])
    for b in range(2):
        assert np.abs(nca[b][:, 0, 0] - (-G0 / 2)).max() <= eps
    oca = O.sigma(P.G0, [[0, 2], [1, 3]])
    for b in range(2):
        assert np.abs(oca[b]).max() <= eps
    third = O.sigma(P.G0, [[0, 3], [1, 4], [2, 5]])
    ana = (beta * t) ** 4 * np.exp(-t * LN2) / 192.0
    assert np.abs(third[0][:, 0, 0] - ana).max() <= eps


def test_one_fermion_one_pole():
    beta, eps = 1.0, 1e-10
    om = 0.8
    P = models.one_fermion(beta=beta, lam=20 * beta, eps=eps, pole=om)
    t = rel2abs(P.grid.it)
    O = po.OracleEvaluator(P)
    nca = O.sigma(P.G0, [[0, 1]])
    tom = t * om
    ana0 = np.exp(-t * LN2) * np.exp(tom) / (np.exp(beta * om) + 1)
    ana1 = np.exp(-t * LN2) * np.exp(-tom) / (np.exp(-beta * om) + 1)
    assert np.abs(nca[0][:, 0, 0] - ana0).max() <= eps
    assert np.abs(nca[1][:, 0, 0] - ana1).max() <= eps
    oca = O.sigma(P.G0, [[0, 2], [1, 3]])
    assert max(np.abs(b).max() for b in oca) <= eps
    third = O.sigma(P.G0, [[0, 3], [1, 4], [2, 5]])
    denom = 2 * om * (om * (np.exp(beta * om) + 1)) ** 3
    a0 = (tom ** 2 * np.exp(tom) - 4 * tom * np.exp(tom) - 2 * tom + 6 * np.exp(tom) - 6) * np.exp(beta * om)
    a1 = (tom ** 2 + 4 * tom + 2 * (tom - 3) * np.exp(tom) + 6) * np.exp(om * (2 * beta - t))
    a0 *= np.exp(-t * LN2) / denom
    a1 *= np.exp(-t * LN2) / denom
    assert np.abs(third[0][:, 0, 0] - a0).max() <= eps
    assert np.abs(third[1][:, 0, 0] - a1).max() <= eps


def test_two_fermions_const_hyb():
    beta, eps = 2.0, 1e-12
    P = models.two_fermion(beta=beta, lam=20 * beta, eps=eps, U=0.0, mu=0.0, pole=0.0)
    t = rel2abs(P.grid.it)
    G0 = -np.exp(-t * 2 * LN2)
    for b in range(3):
        assert np.abs(P.G0[b][:, 0, 0] - G0).max() <= eps
    O = po.OracleEvaluator(P)
    nca = O.sigma(P.G0, [[0, 1]])
    for b in range(3):
        assert np.abs(nca[b][:, 0, 0] - (-G0)).max() <= eps
    oca = O.sigma(P.G0, [[0, 2], [1, 3]])
    ana = -0.25 * np.exp(-t * 2 * LN2) * t * t * beta * beta
    for b in range(3):
        assert np.abs(oca[b][:, 0, 0] - ana).max() <= eps
    third = O.sigma(P.G0, [[0, 3], [1, 4], [2, 5]])
    ana3 = np.exp(-t * 2 * LN2) * t ** 4 * beta ** 4 / 96.0
    for b in range(3):
        assert np.abs(third[b][:, 0, 0] - ana3).max() <= eps
    assert np.abs(third[1][:, 1, 1] - ana3).max() <= eps


def full_sigma(evaluate, max_order):
    """Sum over orders and connected topologies with the solver's net sign -s_topo
    (block_sparse_solver.py:609-638: (-1)^order * s * (-1)^(order+1))."""
    tot = None
    for order in range(1, max_order + 1):
        for sign, topo in all_connected_pairings(order):
            s = evaluate(np.array(topo, dtype=np.int32))
            fac = (-1) ** order * sign * (-1) ** (order + 1)
            tot = [fac * x for x in s] if tot is None else [a + fac * x for a, x in zip(tot, s)]
    return tot


@pytest.mark.parametrize("nflav", [1, 2, 3])
def test_n_fermions_h0(nflav):
    """analytic_diagrams.py:96-107: Sigma_NCA = n/2 G0, dSigma_OCA = 1/4 tau^2/2 n(n-1) G0,
    dSigma_TCA = (-1/2)^3 tau^4/24 (-K(n)) G0 with K(n) = n(4n^2 - 9n + 4)."""
    beta = 2.0
    P = models.n_fermion_h0(nflav, beta=beta, lam=10 * beta, eps=1e-12)
    O = po.OracleEvaluator(P)
    tau = P.grid.tau[:, None, None]
    G0 = P.G0[0]
    N = 2 ** nflav
    assert np.abs(G0 + np.exp(-np.log(N) / beta * tau) * np.eye(N)).max() < 1e-13
    K = nflav * (4 * nflav ** 2 - 9 * nflav + 4)
    nca = 0.5 * nflav * G0
    oca = 0.25 * (tau ** 2 / 2) * nflav * (nflav - 1) * G0
    tca = (-0.5) ** 3 * tau ** 4 / 24 * (-K) * G0
    for order, ana in ((1, nca), (2, nca + oca), (3, nca + oca + tca)):
        s = full_sigma(lambda topo: O.sigma(P.G0, topo), order)[0]
        assert np.abs(s - ana).max() < 1e-10 * max(1.0, np.abs(ana).max()), order
