"""Serving (continuous batching) on the device: every request's result is the one the same problem
gives alone, admission respects the memory budget, and graph replays equal eager launches."""
import os

import numpy as np
import pytest

from xca_amd import models

pytestmark = pytest.mark.gpu


def _alone(P, order, tol, maxiter, mix, graphs=True):
    from xca_amd.evaluator import DiagramEvaluator
    ev = DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)
    ev.set_tile_variant(0)
    ev.dyson_init(P.G0)
    old = os.environ.pop("XCA_NO_GRAPH", None)
    try:
        if not graphs:
            os.environ["XCA_NO_GRAPH"] = "1"
        return ev.ppsc_solve(order, tol=tol, maxiter=maxiter, mix=mix)
    finally:
        os.environ.pop("XCA_NO_GRAPH", None)
        if old is not None:
            os.environ["XCA_NO_GRAPH"] = old


def _rel(a, b):
    return max(np.abs(x - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)


def test_graph_replays_equal_eager_launches():
    """Each self-consistency iteration replays one captured graph; the whole run equals the run with
    eager launches bit for bit (the graph holds no memset nodes; flags are cleared by kernels)."""
    P = models.anderson_nca()
    g1, e1, i1, d1 = _alone(P, 1, 1e-10, 60, 1.0, graphs=True)
    g2, e2, i2, d2 = _alone(P, 1, 1e-10, 60, 1.0, graphs=False)
    print(f"[gpu] graph vs eager: iters {i1} {i2} eta {e1!r} {e2!r} max|dG| {max(np.abs(a - b).max() for a, b in zip(g1, g2)):.1e}",
          flush=True)
    assert i1 == i2 and e1 == e2 and d1 == d2
    assert all(np.array_equal(a, b) for a, b in zip(g1, g2))


@pytest.mark.parametrize("budget_factor", [0.0, 1.5])
def test_server_results_equal_alone(budget_factor):
    """A mixed stream (1-3 orbitals, NCA and OCA, different beta and Lambda) through the server equals
    each problem solved alone. budget_factor > 0 caps device memory at that many times the largest
    request: admission then waits for memory, and fewer problems run together than slots allow."""
    from xca_amd.serving import Server, run_stream
    reqs = models.serving_requests(10, seed=11)
    tol, maxiter, mix = 1e-10, 100, 1.0
    budget = 0
    if budget_factor > 0:
        probe = Server(max_active=16, tile_variant=0)
        sizes = [probe.request_bytes(probe.submit(P, o, tol=tol, maxiter=maxiter, mix=mix)) for P, o in reqs]
        shared = probe.shared_bytes()
        del probe
        budget = shared + int(budget_factor * max(sizes))
        assert sum(sizes) > budget - shared
    sv = Server(mem_budget_bytes=budget, max_active=16, tile_variant=0)
    out = run_stream(sv, reqs, rate=1e4, seed=5, tol=tol, maxiter=maxiter, mix=mix, check_footprint=True)
    print(f"[gpu] stream of {len(reqs)}: {out['solves_per_s']:.1f} solves/s p50 {out['p50_s'] * 1e3:.1f} ms "
          f"p99 {out['p99_s'] * 1e3:.1f} ms max active {out['max_active']} max bytes {out['max_bytes_in_use']} stats {out['stats']}",
          flush=True)
    # the admission estimate bounds what an admitted request really holds on the device
    assert 0 < out["max_footprint_ratio"] <= 1.0
    if budget_factor > 0:
        assert out["max_bytes_in_use"] + sv.shared_bytes() <= budget
        assert out["max_active"] < len(reqs)
    worst = 0.0
    for (P, order), res in zip(reqs, out["results"]):
        g, eta, iters, diff = _alone(P, order, tol, maxiter, mix)
        assert res["converged"] and res["iters"] == iters
        e = _rel(res["G"], g)
        worst = max(worst, e, abs(res["eta"] - eta))
        print(f"  {P.name}: iters {iters} eta {eta:.12f} rel diff {e:.1e}", flush=True)
    assert worst <= 1e-12


def test_server_matches_oracle_on_small_problems():
    """Two small requests through the server against the oracle's self-consistency loop."""
    import pyoracle as po
    from xca_amd.serving import Server, run_stream
    reqs = [(models.anderson_nca(), 1), (models.oca_2orb(), 2)]
    sv = Server(max_active=4, tile_variant=0)
    out = run_stream(sv, reqs, rate=1e4, tol=1e-10, maxiter=60, mix=1.0)
    for (P, order), res in zip(reqs, out["results"]):
        G_o, eta_o, it_o, _ = po.ppsc_solve(P, order, tol=1e-10, maxiter=60, mix=1.0)
        e = _rel(res["G"], G_o)
        print(f"[gpu] server vs oracle {P.name}: iters {res['iters']} vs {it_o} rel {e:.1e}", flush=True)
        assert res["iters"] == it_o and e <= 1e-12


def test_full_serving_stream_equals_alone():
    """The benchmark's whole stream (512 requests, seed 7: 1-3 orbitals, NCA or OCA, beta in
    [10, 100], Lambda = 20 beta) through the server, all submitted at once so that the batch runs
    full: every request converges and equals the same problem solved alone (bit for bit by
    construction, asserted <= 1e-12)."""
    import time
    from xca_amd.serving import Server, run_stream
    t0 = time.perf_counter()
    reqs = models.serving_requests(512, seed=7)
    print(f"[gpu] full stream: built {len(reqs)} requests in {time.perf_counter() - t0:.1f} s", flush=True)
    tol, maxiter, mix = 1e-10, 200, 1.0
    sv = Server(max_active=64, tile_variant=0)
    out = run_stream(sv, reqs, rate=1e9, seed=7, tol=tol, maxiter=maxiter, mix=mix, check_footprint=True)
    print(f"[gpu] full stream: {out['solves_per_s']:.1f} solves/s p50 {out['p50_s']:.3f} s p99 {out['p99_s']:.3f} s "
          f"max active {out['max_active']} footprint ratio {out['max_footprint_ratio']:.3f}", flush=True)
    assert 0 < out["max_footprint_ratio"] <= 1.0
    worst = 0.0
    for k, ((P, order), res) in enumerate(zip(reqs, out["results"])):
        assert res["converged"], P.name
        g, eta, iters, diff = _alone(P, order, tol, maxiter, mix)
        assert res["iters"] == iters, P.name
        worst = max(worst, _rel(res["G"], g), abs(res["eta"] - eta))
        if k % 128 == 127:
            print(f"[gpu] full stream: {k + 1} alone solves compared, worst {worst:.1e}", flush=True)
    assert worst <= 1e-12


def test_server_with_large_block_problem_equals_alone():
    """A 5-orbital d-shell request ((N, S_z) blocks up to 100 states: Sigma on the large-block engine,
    device-wide Dyson) served beside small requests: every result equals the same problem solved alone
    (xca_ppsc_solve), the large one included, and the admission estimate bounds its device bytes."""
    from xca_amd.serving import Server, run_stream
    big = models.oca_5orb_slater(beta=0.25, lam=20.0, eps=1e-10)
    assert int(big.dims.max()) == 100
    reqs = [(models.anderson_nca(), 1), (big, 1), (models.oca_2orb(), 2), (models.anderson_nca(), 1)]
    tol, maxiter, mix = 1e-9, 60, 0.7
    sv = Server(max_active=8, tile_variant=0)
    out = run_stream(sv, reqs, rate=1e4, seed=3, tol=tol, maxiter=maxiter, mix=mix, check_footprint=True)
    print(f"[gpu] stream with a 5-orbital request: max active {out['max_active']} footprint ratio {out['max_footprint_ratio']:.3f}",
          flush=True)
    assert 0 < out["max_footprint_ratio"] <= 1.0
    for (P, order), res in zip(reqs, out["results"]):
        g, eta, iters, diff = _alone(P, order, tol, maxiter, mix)
        e = _rel(res["G"], g)
        print(f"  {P.name}: iters {res['iters']} vs {iters} eta {res['eta']:.12f} vs {eta:.12f} rel diff {e:.1e}", flush=True)
        assert res["converged"] and res["iters"] == iters
        assert e == 0.0 and res["eta"] == eta
