"""Complex128 problems on the device (the real embedding of xca_amd.cplx) against the oracle's complex
instantiation, and the reference's complex-rotation covariance (test/python/unitary_invariance_bs.py)
on the device solver."""
import numpy as np
import pytest

import pyoracle as po
from xca_amd import models

pytestmark = pytest.mark.gpu

U_CPLX = np.array([[1.0, 1j], [1.0, -1j]]) / np.sqrt(2)
TH = 0.4
U_REAL = np.array([[np.cos(TH), -np.sin(TH)], [np.sin(TH), np.cos(TH)]], dtype=complex)
TOL = 1e-12


def rel(a, b):
    return max(np.abs(x - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)


@pytest.mark.parametrize("rotation", ["none", "complex"])
@pytest.mark.parametrize("qn", ["auto", "none"])
def test_complex_sigma_and_spgf_match_complex_oracle(rotation, qn):
    from xca_amd.diag import all_connected_pairings
    from xca_amd.evaluator import DiagramEvaluator
    P = models.two_flavour_hopping(None if rotation == "none" else U_CPLX, qn=qn)
    # unrotated, only G is complex (the constructor sees coefficients and operators): ask for it
    ev = DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad, complex_arith=True if rotation == "none" else None)
    assert ev.complex_arith
    orc = po.OracleEvaluator(P, complex_arith=True)
    G = P.perturbed_G(seed=11, amp=1e-2)
    print(f"[gpu] complex {rotation}/{qn}: dims {list(P.dims)} -> device dims {list(ev.dims)}", flush=True)
    for order in (1, 2, 3):
        for s, t in all_connected_pairings(order):
            t = np.array(t, dtype=np.int32)
            a, b = ev.compute_self_energy(G, t), orc.sigma(G, t)
            e = rel(a, b)
            assert all(np.iscomplexobj(x) for x in a)
            assert e <= TOL, (order, t.tolist(), e)
            g, h = ev.compute_single_ptcle_gf(G, t), orc.spgf(G, t)
            eg = np.abs(g - h).max() / np.abs(h).max()
            assert eg <= TOL, (order, t.tolist(), eg)
        print(f"[gpu] complex {rotation}/{qn} order {order}: Sigma and g(tau) per topology within {TOL:.0e}", flush=True)
    # the signed sum over orders, on the device-resident path (one graph-replayed launch sequence)
    ev.set_G(G)
    ev.sigma_order(3)
    tot = ev.get_sigma()
    ref = None
    for order in (1, 2, 3):
        for s, t in all_connected_pairings(order):
            x = [(-1) ** order * s * (-1) ** (order + 1) * y for y in orc.sigma(G, np.array(t, dtype=np.int32))]
            ref = x if ref is None else [u + v for u, v in zip(ref, x)]
    assert rel(tot, ref) <= TOL


def test_complex_correlator_and_dyson_match_complex_oracle():
    from xca_amd.atom import FockSpace
    from xca_amd.evaluator import DiagramEvaluator
    P = models.two_flavour_hopping(U_CPLX)
    ev = DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)
    orc = po.OracleEvaluator(P, complex_arith=True)
    fs = FockSpace(2)
    ops = []
    for o in range(2):  # n_o, and the hopping c+_0 c_1 (complex after the rotation's H)
        N = (fs.cdag(o) @ fs.c(o)).toarray()
        conn = np.arange(P.ad.nblocks, dtype=np.int32)[None, :]
        ops.append((conn, np.concatenate([N[np.ix_(f, f)].ravel() for f in P.ad.fock_states]).astype(complex) * (1 + 0.5j * o)))
    t = np.array([[0, 2], [1, 3]], dtype=np.int32)
    c = ev.compute_one_time_correlator(P.G0, ops[0], ops[1], t)
    c_ref = orc.correlator(P.G0, t, mu_ops=ops[0], kap_ops=ops[1])
    assert np.abs(c - c_ref).max() <= TOL * np.abs(c_ref).max()
    ev.dyson_init(P.G0)
    S = ev.pseudo_particle_self_energy(P.G0, 2)
    eta = 0.01
    g = ev.dyson_solve(S, eta)
    h = []
    for g0, s in zip(P.G0, S):
        x = po.dyson(P.beta, P.grid.it, P.grid.rf, g0, s, eta)
        h.append(0.5 * (x + x.transpose(0, 2, 1).conj()))
    e = rel(g, h)
    print(f"[gpu] complex Dyson rel_err={e:.2e}", flush=True)
    assert e <= TOL


@pytest.mark.parametrize("order", [1, 2])
def test_complex_rotation_covariance_of_the_device_solver(order):
    """unitary_invariance_bs.py:98-136: one self-consistency iteration in two bases related by a
    rotation, the single-particle Green's function rotated back, dense and block-sparse; the complex
    rotation and the real control; and each G_tau against the oracle's complex loop."""
    from xca_amd.solver import BlockSparseSolver
    worst = 0.0
    for U in (U_REAL, U_CPLX):
        for qn in ("none", "auto"):
            g = []
            for rot in (None, U):
                P = models.two_flavour_hopping(rot, qn=qn)
                S = BlockSparseSolver(P).solve(order, tol=1e-12, maxiter=1, hyb_comp=False)
                assert S.complex_arith
                Go, _, _, _ = po.ppsc_solve(P, order, tol=1e-12, maxiter=1, complex_arith=True)
                g_orc = po.spgf_sum(po.OracleEvaluator(P, complex_arith=True), Go, order)
                assert np.abs(S.G_tau - g_orc).max() <= TOL * np.abs(g_orc).max()
                g.append(S.G_tau)
            diff = np.abs(g[0] - np.einsum("ab,tbc,cd->tad", U, g[1], U.conj().T)).max() / np.abs(g[0]).max()
            worst = max(worst, diff)
            assert diff < 1e-10
    print(f"[gpu] covariance order {order}: worst rel {worst:.2e}", flush=True)


def test_complex_solve_with_compressed_hybridization():
    """An OCA solve of the rotated problem with the hybridization compressed (two exact poles found)
    converges to the oracle's complex self-consistent G on the same poles."""
    from xca_amd.models import ImpurityProblem
    from xca_amd.solver import BlockSparseSolver
    P = models.two_flavour_hopping(U_CPLX, qn="auto")
    S = BlockSparseSolver(P).solve(2, tol=1e-10, maxiter=40, hyb_comp=True)
    assert len(S.hyb.poles) == 2 and np.iscomplexobj(S.hyb.coefficients)
    Q = ImpurityProblem(grid=P.grid, ad=P.ad, poles=S.hyb.poles, coeffs=S.hyb.coefficients, G0=P.G0, eta0=P.eta0)
    Go, eta, it, diff = po.ppsc_solve(Q, 2, tol=1e-10, maxiter=40, complex_arith=True)
    print(f"[gpu] complex OCA solve: {S.n_ppsc_iter} iterations (oracle {it}), rel G {rel(S.G, Go):.2e}", flush=True)
    assert S.n_ppsc_iter == it and S.diff_G < 1e-10
    assert rel(S.G, Go) <= 1e-11
    assert abs(S.eta - eta) <= 1e-11 * max(1.0, abs(eta))
