"""The reference's contract at the device boundary: flat-index range errors, zero G blocks (which prune
block paths and re-key the device plans), and Sigma's exact independence of sub-tolerance cross-set
hybridization."""
import numpy as np
import pytest

import pyoracle as po
from parity_util import TOL, rel_err
from xca_amd import models
from xca_amd._lib import XcaError, lib
from xca_amd.diag import all_connected_pairings

pytestmark = pytest.mark.gpu


def _evaluator(P, coeffs=None):
    from xca_amd.evaluator import DiagramEvaluator
    return DiagramEvaluator(P.poles, P.coeffs if coeffs is None else coeffs, P.grid, P.ad)


def test_sigma_rejects_flat_index_out_of_range():
    """block_sparse_backbone.cpp:319: an f_ix outside [0, get_num_self_energy_backbones) throws
    "f_ix out of range"; the context stays usable."""
    P = models.oca_2orb()
    G = P.perturbed_G()
    ev = _evaluator(P)
    t = np.array([[0, 2], [1, 3]], dtype=np.int32)
    nmax = ev.get_num_self_energy_backbones(t)
    for bad in ([nmax], [0, -1], [3, nmax + 7]):
        with pytest.raises(XcaError, match="f_ix out of range"):
            ev.compute_self_energy(G, t, np.array(bad, dtype=np.int32))
    fix = np.array([0] + list(range(5, nmax - 1, 9)) + [nmax - 1], dtype=np.int32)  # the ends included
    ok = ev.compute_self_energy(G, t, fix)
    assert rel_err(ok, [x.real for x in po.OracleEvaluator(P, complex_arith=False).sigma(G, t, fix)]) <= TOL


def test_zero_blocks_prune_paths_and_rekey_plans():
    """G blocks that are exactly zero (block_sparse.cpp:45-49) remove every block path through them.
    Six zero patterns in turn on one evaluator: each matches the oracle per topology and through the
    combined launch, and the plans of at most four patterns stay cached."""
    P = models.oca_2orb()
    G0 = P.perturbed_G()
    ev = _evaluator(P)
    orc = po.OracleEvaluator(P, complex_arith=False)
    topos = [(o, s, np.array(t, dtype=np.int32)) for o in (1, 2) for s, t in all_connected_pairings(o)]
    patterns = [[], [4], [1], [4, 7], [0, 8], [2, 3, 5], [4]]
    for zero in patterns:
        G = [np.zeros_like(g) if b in zero else g for b, g in enumerate(G0)]
        full = None
        for order, sign, t in topos:
            dev = ev.compute_self_energy(G, t)
            cpu = [x.real for x in orc.sigma(G, t)]
            e = rel_err(dev, cpu)
            fac = (-1) ** order * sign * (-1) ** (order + 1)
            full = [fac * x for x in cpu] if full is None else [a + fac * x for a, x in zip(full, cpu)]
            print(f"[gpu] zero blocks {zero} topo={t.tolist()} rel_err={e:.2e}", flush=True)
            assert e <= TOL
        ev.set_G(G)
        ev.sigma_order(2)
        e = rel_err(ev.get_sigma(), full)
        nplans = lib().xca_num_cached_plans(ev._h)
        print(f"[gpu] zero blocks {zero} combined launch rel_err={e:.2e} cached plans={nplans}", flush=True)
        assert e <= TOL
        assert 0 < nplans <= 4 * len(topos)


def test_sub_tolerance_cross_set_leaves_device_sigma_bitwise_unchanged():
    """test_block_sparse_sym_set_mask.cpp:107-170 on the device: a cross-set hybridization coefficient
    at 0.9 * 1e-12 * max|hyb_coeffs| is accepted and leaves Sigma exactly unchanged; the same
    perturbation inside a symmetry set moves it."""
    P = models.kanamori(2, 4.0, 0.5, 4.75, beta=5.0, lam=50.0, eps=1e-10, offdiag=0.3)
    G = P.perturbed_G()
    t = np.array([[0, 2], [1, 3]], dtype=np.int32)
    delta = 0.9e-12 * np.abs(P.coeffs).max()
    s0 = _evaluator(P).compute_self_energy(G, t)
    assert max(np.abs(x).max() for x in s0) > 1e-3
    cross = P.coeffs.copy()
    cross[:, 0, 2] += delta
    cross[:, 3, 1] += delta
    s1 = _evaluator(P, cross).compute_self_energy(G, t)
    assert all(np.array_equal(a, b) for a, b in zip(s1, s0))
    within = P.coeffs.copy()
    within[:, 0, 1] += delta
    s2 = _evaluator(P, within).compute_self_energy(G, t)
    assert max(np.abs(a - b).max() for a, b in zip(s2, s0)) > 0.0
