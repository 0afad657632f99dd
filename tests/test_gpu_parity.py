"""Device path vs the CPU oracle on identical inputs: max |S_gpu - S_cpu| / max |S_cpu| <= 1e-12, per
topology and for the full signed Sigma (metric of the hot path). Where the double-precision oracle is
itself further than 1e-12 from the exact result of the inputs, the bar is the extended-precision one
of parity_util.assert_parity."""
import numpy as np
import pytest

import pyoracle as po
from parity_util import TOL, assert_parity, rel_err
from xca_amd import models
from xca_amd.diag import all_connected_pairings

pytestmark = pytest.mark.gpu


def _evaluator(P):
    from xca_amd.evaluator import DiagramEvaluator
    return DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)


CASES = {
    "one_fermion_pole": lambda: models.one_fermion(beta=1.0, lam=20.0, eps=1e-10, pole=0.8),
    "two_fermion": lambda: models.two_fermion(beta=2.0, lam=40.0, eps=1e-12, U=3.0, mu=0.5, pole=0.7),
    "n3_h0": lambda: models.n_fermion_h0(3, beta=2.0, lam=20.0, eps=1e-12),
    "two_band_discrete_bath": models.two_band_discrete_bath,
    "anderson_nca": models.anderson_nca,
    "oca_2orb": models.oca_2orb,
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("order", [1, 2, 3])
def test_sigma_per_topology(case, order):
    P = CASES[case]()
    if case in ("oca_2orb", "anderson_nca") and order == 3:
        pytest.skip("third order on the benchmark models is covered by the TCA sample test")
    G = P.perturbed_G() if case in ("oca_2orb", "anderson_nca") else P.G0
    print(f"[gpu] case={case} order={order} r={P.grid.r} dims={list(P.dims)}", flush=True)
    ev = _evaluator(P)
    orc = po.OracleEvaluator(P, complex_arith=False)
    for sign, topo in all_connected_pairings(order):
        t = np.array(topo, dtype=np.int32)
        assert_parity(f"{case} topo={topo}", ev.compute_self_energy(G, t), P, G, t, orc=orc)


def test_oca_two_band_discrete_bath_matches_tpz100():
    """Device OCA against the reference's trapezoidal-quadrature fixture (tpz100.ref.h5), compared
    through the per-tau eigenvalues of each block (basis independent), tolerance 4e-4 as
    test_block_sparse_OCA_manual.cpp:237-282."""
    import json
    from pathlib import Path
    from test_oracle_reference_fixtures import FIXTURE_N, block_particle_numbers, equispaced_rel
    fx = json.loads((Path(__file__).resolve().parent / "golden" / "oca_tpz100_blocks.json").read_text())
    P = models.two_band_discrete_bath()
    ev = _evaluator(P)
    S = ev.compute_self_energy(P.G0, np.array([[0, 2], [1, 3]], dtype=np.int32))
    rel = equispaced_rel(fx["ntau"])
    Nb = block_particle_numbers(P.ad)
    worst = 0.0
    for k, fb in enumerate(fx["blocks_real"]):
        s = S[Nb.index(FIXTURE_N[k])]
        seq = po.interp(P.grid.it, P.grid.rf, s, rel).real
        worst = max(worst, np.abs(np.linalg.eigvalsh(seq) - np.linalg.eigvalsh(np.array(fb))).max())
    print(f"[gpu] tpz100 fixture max eigenvalue error {worst:.2e}", flush=True)
    assert worst <= 4e-4


@pytest.mark.parametrize("case,order", [("oca_2orb", 2), ("two_band_discrete_bath", 3)])
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_every_tile_variant(variant, case, order):
    """Each chain-kernel tile width (128/64/32 columns) gives the oracle's Sigma, and so does the
    autotuner's choice (-1). The two-band third order once produced NaN through stale tile columns
    (0 * inf in the padding rows); it runs twice here so every tile reuses a dirty LDS slab."""
    P = CASES[case]()
    G = P.perturbed_G() if case == "oca_2orb" else P.G0
    ev = _evaluator(P)
    orc = po.OracleEvaluator(P, complex_arith=False)
    topos = [t for o in range(1, order + 1) for _, t in all_connected_pairings(o)]
    ref = {str(t): [x.real for x in orc.sigma(G, np.array(t, dtype=np.int32))] for t in topos}
    for v in (variant, -1):
        ev.set_tile_variant(v)
        for rep in range(2):
            for topo in topos:
                t = np.array(topo, dtype=np.int32)
                s = ev.compute_self_energy(G, t)
                used = ev.last_tile()
                assert all(np.isfinite(x).all() for x in s)
                assert_parity(f"{case} variant {v} used {used} topo={topo}", s, P, G, t, ref=ref[str(topo)])
                # a forced width that cannot hold the widest block path falls back to 128 columns
                if v >= 0:
                    assert used[0] in (v, 0)


@pytest.mark.parametrize("chunk_tiles", ["1", "16"])
def test_shared_prefix_third_order(chunk_tiles, monkeypatch):
    """Third order with items that share the first line's pole: the ops before the second line
    enters run once per tile (or are read back from the workgroup's cache when consecutive tiles of
    a chunk share the pole). Long chunks exercise the cache."""
    monkeypatch.setenv("XCA_CHUNK_TILES", chunk_tiles)
    P = models.two_band_discrete_bath()
    ev = _evaluator(P)
    orc = po.OracleEvaluator(P, complex_arith=False)
    for v in (0, 2):
        ev.set_tile_variant(v)
        for sign, topo in all_connected_pairings(3):
            t = np.array(topo, dtype=np.int32)
            assert_parity(f"shared prefix chunk {chunk_tiles} variant {v} topo={topo}", ev.compute_self_energy(P.G0, t), P, P.G0, t, orc=orc)


def test_flat_index_subset_matches_oracle():
    """compute_self_energy(G, topology, f_ix_vec) on a scattered subset (the MPI-chunk contract)."""
    P = models.oca_2orb()
    G = P.perturbed_G()
    ev = _evaluator(P)
    orc = po.OracleEvaluator(P, complex_arith=False)
    topo = np.array([[0, 2], [1, 3]], dtype=np.int32)
    nmax = ev.get_num_self_energy_backbones(topo)
    assert nmax == orc.num_backbones(2) == 4 * 4 * P.grid.r
    fix = np.arange(7, nmax, 13, dtype=np.int32)
    e = rel_err(ev.compute_self_energy(G, topo, fix), [x.real for x in orc.sigma(G, topo, fix)])
    assert e <= TOL
    # decomposition invariant: sum over chunks == full (test_block_sparse_backbone_eval.cpp:810-869)
    chunks = np.array_split(np.arange(nmax, dtype=np.int32), 5)
    parts = [ev.compute_self_energy(G, topo, c) for c in chunks]
    full = ev.compute_self_energy(G, topo)
    tot = [sum(p[b] for p in parts) for b in range(len(full))]
    assert rel_err(tot, full) <= TOL


def test_full_sigma_signed():
    P = models.oca_2orb()
    G = P.perturbed_G()
    ev = _evaluator(P)
    orc = po.OracleEvaluator(P, complex_arith=False)
    s_gpu = ev.pseudo_particle_self_energy(G, 2)
    s_cpu = None
    for order in (1, 2):
        for sign, topo in all_connected_pairings(order):
            s = [x.real for x in orc.sigma(G, np.array(topo, dtype=np.int32))]
            fac = (-1) ** order * sign * (-1) ** (order + 1)
            s_cpu = [fac * x for x in s] if s_cpu is None else [a + fac * x for a, x in zip(s_cpu, s)]
    assert rel_err(s_gpu, s_cpu) <= TOL


@pytest.mark.parametrize("case,order", [("oca_2orb", 2), ("two_band_discrete_bath", 3)])
def test_sigma_order_one_launch_equals_topology_loop(case, order):
    """xca_sigma_order (all topologies in one launch sequence, graph-replayed) against the topology
    loop: bit for bit at one forced tile width, on every replay and eagerly; with the autotuner, the
    combined launch may run a small plan at the costliest plan's tile width, which changes the
    order of the partial sums only (relative difference <= 1e-14)."""
    P = CASES[case]()
    G = P.perturbed_G()
    ev = _evaluator(P)
    ref_auto = ev.pseudo_particle_self_energy(G, order)
    ev.set_G(G)
    ev.sigma_order(order)
    e = rel_err(ev.get_sigma(), ref_auto)
    print(f"[gpu] sigma_order {case} autotuned rel={e:.2e}", flush=True)
    assert e <= 1e-14
    ev.set_tile_variant(0)
    ref = ev.pseudo_particle_self_energy(G, order)
    ev.set_G(G)
    for rep in range(3):
        ev.sigma_order(order)
        got = ev.get_sigma()
        assert all(np.array_equal(a, b) for a, b in zip(got, ref)), rep
    ev.sigma_order(order, timed=True)
    ev.synchronize()
    assert all(np.array_equal(a, b) for a, b in zip(ev.get_sigma(), ref))
    items, macs, ms = ev.last_stats()
    print(f"[gpu] sigma_order {case} items={items} chain_ms={ms:.3f}", flush=True)
    assert items > 0 and ms > 0


@pytest.mark.parametrize("case,order", [("oca_2orb", 2), ("two_band_discrete_bath", 3)])
def test_per_item_launch_form_equals_batched(case, order):
    """Launch mode 1 (one chain launch per chunk; with an explicit flat-index subset, one launch per
    (path, pole tuple) item) gives the batched launch's Sigma: bit for bit over all flat indices, to
    summation order on a subset."""
    P = CASES[case]()
    G = P.perturbed_G()
    ev = _evaluator(P)
    rng = np.random.default_rng(3)
    for _, topo in all_connected_pairings(order):
        t = np.array(topo, dtype=np.int32)
        nf = ev.get_num_self_energy_backbones(t)
        sub = np.sort(rng.choice(nf, size=min(nf, 400), replace=False)).astype(np.int32)
        ev.set_launch_mode(0)
        full0, sub0 = ev.compute_self_energy(G, t), ev.compute_self_energy(G, t, sub)
        ev.set_launch_mode(1)
        full1, sub1 = ev.compute_self_energy(G, t), ev.compute_self_energy(G, t, sub)
        items, _, ms = ev.last_stats()
        print(f"[gpu] per-item launches {case} topo={topo} items={items} ms={ms:.3f}", flush=True)
        assert all(np.array_equal(a, b) for a, b in zip(full0, full1))
        # a subset's items share tiles in mode 0 and run one per launch in mode 1: the same per-item
        # arithmetic, summed in another order
        assert rel_err(sub1, sub0) <= 1e-14
    ev.set_launch_mode(0)
    with pytest.raises(Exception, match="mode must be 0 or 1"):
        ev.set_launch_mode(2)


def test_forced_width_timed_evaluation_after_graph_replays(monkeypatch):
    """A tile width forced at context creation (no autotuner, whose launches would come first): graph
    replays of the combined launch, then an eager, event-timed evaluation of the same sequence. The
    timed result equals the replayed one and the oracle's."""
    monkeypatch.setenv("XCA_TILE_VARIANT", "0")
    P = models.oca_2orb()
    G = P.perturbed_G()
    ev = _evaluator(P)
    ev.set_G(G)
    for _ in range(3):
        ev.sigma_order(2)
    replayed = ev.get_sigma()
    ev.sigma_order(2, timed=True)
    ev.synchronize()  # reads the chain launch's event time
    timed = ev.get_sigma()
    items, macs, ms = ev.last_stats()
    assert items == 5736 and ms > 0
    assert all(np.array_equal(a, b) for a, b in zip(timed, replayed))
    orc = po.OracleEvaluator(P, complex_arith=False)
    ref = None
    for order in (1, 2):
        for s, t in all_connected_pairings(order):
            x = [-s * y.real for y in orc.sigma(G, np.array(t, dtype=np.int32))]
            ref = x if ref is None else [u + v for u, v in zip(ref, x)]
    assert rel_err(timed, ref) <= TOL
