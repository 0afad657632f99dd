"""Dyson solve of the blocks the one-workgroup kernel does not take (more than 12 states, or r n above
784): the 5-orbital Slater-Condon d shell, whose (N, S_z) blocks reach 100 states, on a grid small
enough for the oracle and resolving the atomic propagator (beta = 0.25, Lambda = 20, r = 19,
beta E_max = 12.5): blocks of 25, 50 and 100 (systems of 475, 950 and 1900) take the device-wide
path, blocks up to 12 states the one-workgroup kernel. Against the oracle's dyson_solve
(hermitised), the closed form Sigma = 0, eta = 0 -> G0, and the fixed point of the self-consistent
loop."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import pyoracle as po
from xca_amd import models

pytestmark = pytest.mark.gpu
TOL = 1e-12


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.fixture(scope="module")
def shell():
    from xca_amd.evaluator import DiagramEvaluator
    P = models.oca_5orb_slater(beta=0.25, lam=20.0, eps=1e-10)
    assert P.grid.r * int(P.dims.max()) > 784
    ev = DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)
    ev.dyson_init(P.G0)
    return P, ev


def test_large_block_dyson_matches_oracle(shell):
    P, ev = shell
    S = ev.pseudo_particle_self_energy(P.G0, 1)  # NCA Sigma of G0 (large-block Sigma engine)
    eta = 0.01
    G = ev.dyson_solve(S, eta)
    dims = list(P.dims)
    picks = [dims.index(10), dims.index(25), dims.index(50), dims.index(100)]

    def oracle(b):
        g = po.dyson(P.beta, P.grid.it, P.grid.rf, P.G0[b], S[b], eta).real
        return 0.5 * (g + g.transpose(0, 2, 1))

    with ThreadPoolExecutor(max_workers=4) as pool:
        refs = list(pool.map(oracle, picks))
    for b, ref in zip(picks, refs):
        e = rel(G[b], ref)
        print(f"[gpu] Dyson block {b} (n = {dims[b]}, N = {P.grid.r * dims[b]}) rel_err={e:.2e}", flush=True)
        assert e <= TOL
    # Sigma = 0, eta = 0 reproduces G0 on every block, small and large
    G_free = ev.dyson_solve([0 * s for s in S], 0.0)
    worst = max(rel(g, g0) for g, g0 in zip(G_free, P.G0))
    print(f"[gpu] Dyson Sigma = 0: worst block {worst:.2e}", flush=True)
    assert worst <= TOL


def test_large_block_self_consistency_fixed_point(shell):
    """NCA self-consistency of the d shell through BlockSparseSolver (host loop around the device's
    Sigma and Dyson): it converges, Z = 1, and one more Sigma + Dyson step reproduces G."""
    from xca_amd.solver import BlockSparseSolver
    P, ev = shell
    tol = 1e-9
    S = BlockSparseSolver(P).solve(1, tol=tol, maxiter=40, hyb_comp=False)
    print(f"[gpu] 5-orbital NCA solve: {S.n_ppsc_iter} iterations, diff {S.diff_G:.2e}, eta {S.eta:.6f}", flush=True)
    assert S.diff_G < tol
    # g(tau) of the converged G on the large-block engine (its parity: test_gpu_wide_corr.py)
    assert S.G_tau.shape == (P.grid.r, P.n, P.n) and np.isfinite(S.G_tau).all()
    g01 = po.interp(P.grid.it, P.grid.rf, S.G_tau, np.array([0.0, 1.0])).real
    print(f"[gpu] 5-orbital NCA g(0) + g(beta) diagonal: {np.diag(g01[0] + g01[1])}", flush=True)
    # <c c^dag> + <c^dag c> = 1 with Z = 1: g(0+) + g(beta-) = -1 on the diagonal
    assert np.abs(np.diag(g01[0] + g01[1]) + 1.0).max() < 1e-6
    Z = S.partition_function_from_ppgf(S.G)
    assert abs(Z - 1.0) < 1e-7
    G2 = S.d.dyson_solve(S.d.pseudo_particle_self_energy(S.G, 1), S.eta)
    assert max(float(np.abs(a - b).max()) for a, b in zip(G2, S.G)) < 10 * tol


def test_device_loop_matches_host_loop(shell):
    """xca_ppsc_solve on the large-block problem (Sigma on the large-block engine, device-wide Dyson,
    device normalisation / convergence / mixing) against the same loop driven from the host around
    the device's Sigma and Dyson (Z, eta, diff and mixing in numpy): same iterates and history."""
    P, ev = shell
    tol = 1e-9
    Gd, etad, itd, diffd = ev.ppsc_solve(1, tol=tol, maxiter=40, mix=0.7)
    hd = ev.ppsc_history()
    Gh, etah, ith, diffh = ev._ppsc_solve_host(1, tol, 40, 0.7, 0.0)
    hh = ev._history
    print(f"[gpu] large-block loop: device {itd} iterations diff {diffd:.2e} eta {etad:.9f}; host {ith} iterations diff {diffh:.2e} "
          f"eta {etah:.9f}", flush=True)
    assert diffd < tol and diffh < tol
    assert itd == ith and len(hd) == itd
    for (a, b, c), (x, y, z) in zip(hd, hh):
        assert abs(b - y) <= 1e-12 * abs(y) and abs(c - z) <= 1e-12 * max(1.0, abs(z))
        assert abs(a - x) <= 1e-10 * max(x, 1e-12) + 1e-14
    assert abs(etad - etah) <= 1e-12
    assert max(rel(a, b) for a, b in zip(Gd, Gh)) <= 1e-12
