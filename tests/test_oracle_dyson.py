"""Oracle pins for the pseudo-particle Dyson solve (dyson_it_ppsc::solve, c++/triqs_xca/dlr_dyson_ppsc.hpp:135-168)
and the classic-normalization self-consistency loop (python/triqs_xca/block_sparse_solver.py:295-346, 427-449).

Closed forms:
  * Sigma = 0: G = G0 exp(-eta tau) (the pseudo-particle chemical potential shifts every level);
  * a block H coupled to bath levels Eb through V, Sigma(tau) = -V exp(-Eb tau) V^T: the Dyson equation
    (d/dtau + H + eta) G = -Sigma * G with G(0) = -1 is the projection of the enlarged system, so
    G(tau) = -[exp(-A tau)]_HH with A = [[H + eta, V], [V^T, Eb]].
H and V V^T do not commute, so exchanging the order of the convolution operands or the sign of Sigma fails.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import pyoracle as po
from xca_amd import models
from xca_amd.diag import all_connected_pairings
from xca_amd.dlr import DLRGrid, rel2abs


def _grid(beta=10.0, lam=100.0, eps=1e-12):
    return DLRGrid(beta, lam, eps)


def _expm_nodes(A, tau):
    return np.array([sl.expm(-A * t) for t in tau])


def test_dyson_eta_shift():
    g = _grid()
    tau = rel2abs(g.it) * g.beta
    H = np.array([[0.3, 0.2], [0.2, 1.4]])
    G0 = -_expm_nodes(H, tau)
    eta = 0.37
    G = po.dyson(g.beta, g.it, g.rf, G0, np.zeros_like(G0), eta).real
    exact = G0 * np.exp(-eta * tau)[:, None, None]
    assert np.abs(G - exact).max() <= 1e-12


@pytest.mark.parametrize("n,m", [(1, 1), (2, 2), (3, 2)])
def test_dyson_bath_projection(n, m):
    rng = np.random.default_rng(7 + n)
    g = _grid()
    tau = rel2abs(g.it) * g.beta
    H = rng.uniform(-0.5, 0.5, (n, n))
    H = 0.5 * (H + H.T) + np.diag(np.linspace(0.0, 1.5, n))
    Eb = np.linspace(0.4, 2.0, m)
    V = 0.6 * rng.standard_normal((n, m))
    eta = 0.15
    G0 = -_expm_nodes(H, tau)
    Sig = -np.array([V @ np.diag(np.exp(-Eb * t)) @ V.T for t in tau])
    G = po.dyson(g.beta, g.it, g.rf, G0, Sig, eta).real
    A = np.block([[H + eta * np.eye(n), V], [V.T, np.diag(Eb)]])
    exact = -_expm_nodes(A, tau)[:, :n, :n]
    err = np.abs(G - exact).max() / np.abs(exact).max()
    assert err <= 1e-10
    # the wrong sign of Sigma or the reversed operand order is far off
    Gw = po.dyson(g.beta, g.it, g.rf, G0, -Sig, eta).real
    assert np.abs(Gw - exact).max() > 1e-3


def test_ppsc_fixed_point_and_normalization():
    P = models.anderson_nca()
    G, eta, it, diff = po.ppsc_solve(P, 1, tol=1e-11, maxiter=60, mix=1.0)
    assert diff < 1e-11 and it < 60
    # Z = -Tr G(beta) = 1 at the fixed point
    Z = -sum(np.trace(po.interp(P.grid.it, P.grid.rf, b, np.array([1.0]))[0].real) for b in G)
    assert abs(Z - 1.0) <= 1e-9
    # one more Dyson step reproduces G (NCA: Sigma of order 1, net factor -s for its single topology)
    (s_topo, topo), = all_connected_pairings(1)
    O = po.OracleEvaluator(P, complex_arith=False)
    S = [-s_topo * x.real for x in O.sigma(G, np.array(topo, dtype=np.int32))]
    Gn = []
    for g0, s in zip(P.G0, S):
        x = po.dyson(P.beta, P.grid.it, P.grid.rf, g0, s, eta).real
        Gn.append(0.5 * (x + x.transpose(0, 2, 1)))
    assert max(np.abs(a - b).max() for a, b in zip(Gn, G)) <= 1e-9
