"""Device Dyson solve and self-consistency loop vs the oracle restatement of the reference solver."""
import numpy as np
import pytest

import pyoracle as po
from xca_amd import models

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return max(np.abs(x - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)


@pytest.mark.parametrize("case", ["anderson_nca", "oca_2orb"])
def test_dyson_single_solve(case):
    from xca_amd.evaluator import DiagramEvaluator
    P = getattr(models, case)()
    ev = DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)
    ev.dyson_init(P.G0)
    S = ev.pseudo_particle_self_energy(P.G0, 1)
    eta = 0.01
    g_dev = ev.dyson_solve(S, eta)
    g_orc = []
    for g0, s in zip(P.G0, S):
        g = po.dyson(P.beta, P.grid.it, P.grid.rf, g0, s, eta).real
        g_orc.append(0.5 * (g + g.transpose(0, 2, 1)))
    e = rel_err(g_dev, g_orc)
    print(f"[gpu] dyson {case} rel_err={e:.2e}", flush=True)
    assert e <= 1e-12
    # Sigma = 0, eta = 0 reproduces G0 (block_sparse_api_two_fermions.py:114-116)
    g_free = ev.dyson_solve([0 * x for x in S], 0.0)
    assert rel_err(g_free, P.G0) <= 1e-12


@pytest.mark.parametrize("case,order", [("anderson_nca", 1), ("oca_2orb", 2)])
def test_ppsc_solve_matches_oracle(case, order):
    from xca_amd.solver import BlockSparseSolver
    P = getattr(models, case)()
    tol, maxiter = 1e-10, 60
    S = BlockSparseSolver(P).solve(order, tol=tol, maxiter=maxiter, mix=1.0)
    G_o, eta_o, it_o, diff_o = po.ppsc_solve(P, order, tol=tol, maxiter=maxiter, mix=1.0)
    e = rel_err(S.G, G_o)
    print(f"[gpu] ppsc {case}: iters {S.n_ppsc_iter} vs {it_o}, eta {S.eta:.15e} vs {eta_o:.15e}, rel_err={e:.2e}", flush=True)
    assert S.n_ppsc_iter == it_o
    assert abs(S.eta - eta_o) <= 1e-12 * max(1.0, abs(eta_o))
    assert e <= 1e-12
    # normalization: Z = -Tr G(beta) = 1 after the last update (block_sparse_api_two_fermions.py:110-112)
    assert S.diff_G < tol
