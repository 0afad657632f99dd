"""Benchmark: pseudo-particle self-energy evaluations per second on one MI355X.

One step = one full Sigma evaluation (all connected topologies of orders 1..max_order, with the
solver's signs) of the OCA 2-orbital Kanamori workload (U = 4, J = 0.5, half filling, beta = 50,
Lambda = 1000, eps = 1e-10, 16 states in 9 (N, S_z) blocks, semicircular hybridization, G = atomic
propagator with seed-perturbed DLR coefficients). G stays resident in HBM across steps; each step
is one graph-replayed launch sequence (edge tables, one persistent chain launch, one reduction).

Prints one JSON line: metric, value (Sigma-evaluations/s), ms_per_step and the run parameters;
roofline (the chain launch's canonical FP64 rate against the MI355X peak, its HIP-event time and
the HBM bytes per launch from the committed rocprofv3 summary); the plain GPU forms beside the
batched value (plain_gpu); g(tau) evaluations (spgf); NCA solve latency (solve); the served request
stream (serving); and the oracle on the host's cores (cpu_baseline). --workload tca3 / oca5 measure
the 3-orbital TCA and 5-orbital Slater-Condon configurations instead.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "PP Σ evals/s (OCA/TCA) + solve p50/p99 latency at ≤1e-12 rel vs CPU Σ"
PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 (matrix and vector)


def build_workload(name):
    from xca_amd import models
    if name == "oca2":
        return models.oca_2orb(), 2
    if name == "tca3":
        return models.tca_3orb(), 3
    if name == "oca5":
        return models.oca_5orb_slater(), 2
    raise SystemExit(f"unknown workload {name}")


def measured_traffic(workload):
    """HBM bytes per launch of the dominant chain launch, from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE scaled by a plain copy of known size); None if not measured."""
    import re
    key = f"traffic_bytes_per_launch[{workload}]:"
    for f in sorted((ROOT / "profiles").glob("*.md")):
        for line in f.read_text().splitlines():
            if line.startswith(key):
                return float(re.findall(r"[0-9.eE+]+", line[len(key):])[0])
    return None


def cpu_baseline(P, G, max_order, budget_s, cores):
    """Time the oracle (CPU restatement of the reference path, complex128 like the reference) on the
    same workload for about budget_s seconds, on `cores` host threads, each with its own evaluator
    and a contiguous run of the flat indices, as the reference's solver scatters them over MPI ranks
    (block_sparse_solver.py:629-640); wall time is that of the slowest thread. When full Sigma
    evaluations fit the budget they are repeated; otherwise each topology is timed on a seeded
    sample of flat indices (doubled until it uses its share of the budget) and scaled by the sample
    fraction, since path cost does not depend on the poles."""
    from concurrent.futures import ThreadPoolExecutor
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    from xca_amd.diag import all_connected_pairings
    orcs = [po.OracleEvaluator(P, complex_arith=True) for _ in range(cores)]
    topos = [np.array(t, dtype=np.int32) for o in range(1, max_order + 1) for _, t in all_connected_pairings(o)]
    rng = np.random.default_rng(11)
    pool = ThreadPoolExecutor(max_workers=cores)

    def timed(t, fix):  # wall time of one topology over fix, split in contiguous runs over the threads
        parts = [x for x in np.array_split(np.sort(fix).astype(np.int32), cores) if len(x)]
        t0 = time.perf_counter()
        list(pool.map(lambda a: orcs[a[0]].sigma(G, t, a[1]), enumerate(parts)))
        return time.perf_counter() - t0

    share = budget_s / len(topos)
    est, spent, sampled = 0.0, 0.0, []
    for t in topos:
        nmax = orcs[0].num_backbones(len(t))
        perm = rng.permutation(nmax).astype(np.int32)
        k, done, el = cores, 0, 0.0
        while done < nmax:
            k = min(k, nmax - done)
            el += timed(t, perm[done:done + k])
            done += k
            if el >= 0.5 * share:
                break
            k *= 2
        est += el * nmax / done
        spent += el
        if done < nmax:
            sampled.append(f"{done}/{nmax}")
    if not sampled:  # full evaluations fit: repeat them for the rest of the budget
        evals = 1
        while spent < budget_s:
            for t in topos:
                spent += timed(t, np.arange(orcs[0].num_backbones(len(t)), dtype=np.int32))
            evals += 1
        est = spent / evals
        pool.shutdown()
        return {"value": 1.0 / est, "unit": "Σ-evals/s", "cores": cores, "kind": "port",
                "sample": f"{evals} full Sigma evaluations (orders 1-{max_order}, all flat indices, complex128) on {cores} host "
                          f"threads in {spent:.1f} s"}
    pool.shutdown()
    how = ("seeded samples of flat indices (" + ", ".join(sampled) +
           ") scaled by the sample fraction, the other topologies in full")
    return {"value": 1.0 / est, "unit": "Σ-evals/s", "cores": cores, "kind": "port",
            "sample": f"orders 1-{max_order}, complex128, {cores} host threads: {how}; {spent:.1f} s measured, "
                      f"{est:.3g} s per Sigma evaluation"}


def _po():
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as po
    return po


def host_threads(requested):
    """Oracle threads: the requested count, at most the cores this process may run on."""
    return max(1, min(requested, len(os.sched_getaffinity(0))))


def rel_errors(dev, ref):
    """max |dev - ref| / max |ref| globally and per block (absolute where a reference block is 0)."""
    num = max(float(np.abs(a - b).max()) for a, b in zip(dev, ref))
    den = max(float(np.abs(b).max()) for b in ref)
    per = [float(np.abs(a - b).max()) / (float(np.abs(b).max()) or 1.0) for a, b in zip(dev, ref)]
    return {"global": num / den if den > 0 else num, "per_block_max": max(per), "per_block": per}


def signed_oracle_sigma(P, G, max_order, threads):
    """Sigma over orders 1..max_order with the solver's net factor -s per topology, from the oracle
    (double precision) over every flat index, on `threads` host threads."""
    from xca_amd.diag import all_connected_pairings
    po = _po()
    n = po.OracleEvaluator(P, complex_arith=False)
    tot = None
    for o in range(1, max_order + 1):
        for sg, topo in all_connected_pairings(o):
            t = np.array(topo, dtype=np.int32)
            x = po.sigma_threads(P, G, t, np.arange(n.num_backbones(o), dtype=np.int32), threads=threads)
            tot = [-sg * a for a in x] if tot is None else [u - sg * a for u, a in zip(tot, x)]
    return tot


def copy_calibration(mib):
    """Plain copy kernel of known size: FETCH_SIZE / WRITE_SIZE per dispatch of this copy divided by
    mib * 1024 KiB gives the scale factor applied to the same counters of the chain kernel."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "dispatches": 5}))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def solve_latency(n, cpu=False):
    """Self-consistent NCA solve of the single-orbital Anderson impurity (U = 2, mu = 1, beta = 10,
    Lambda = 100, eps = 1e-10, semicircular hybridization, tol 1e-10, mix 1) timed n times, each from
    the atomic propagator: device solve including its host driver and result download."""
    from xca_amd import models
    from xca_amd.evaluator import DiagramEvaluator
    P = models.anderson_nca()
    ev = DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)
    ev.dyson_init(P.G0)
    ev.ppsc_solve(1, tol=1e-10, maxiter=100, mix=1.0)  # plans and autotuning
    times, iters = [], 0
    for _ in range(n):
        t0 = time.perf_counter()
        _, _, iters, _ = ev.ppsc_solve(1, tol=1e-10, maxiter=100, mix=1.0)
        times.append(time.perf_counter() - t0)
    t = np.array(times) * 1e3
    out = {"workload": "NCA 1-orbital Anderson U=2 mu=1 beta=10 Lambda=100 eps=1e-10, full self-consistency to 1e-10",
           "solves": n, "iterations": int(iters), "p50_ms": float(np.percentile(t, 50)), "p99_ms": float(np.percentile(t, 99)),
           "mean_ms": float(t.mean())}
    if cpu:  # the restated reference loop (complex128, one core): Sigma, per-block Dyson, normalisation
        po = _po()
        ct = []
        for _ in range(5):
            t0 = time.perf_counter()
            _, _, it_o, _ = po.ppsc_solve(P, 1, tol=1e-10, maxiter=100, mix=1.0, complex_arith=True)
            ct.append(time.perf_counter() - t0)
        out["cpu_oracle"] = {"p50_ms": float(np.percentile(np.array(ct) * 1e3, 50)), "iterations": int(it_o), "cores": 1,
                             "solves": 5, "kind": "port"}
        out["cpu_over_gpu_p50"] = out["cpu_oracle"]["p50_ms"] / out["p50_ms"]
    return out


def serving(n, seed=7, cpu_sample=0, threads=16, fixed_rate=43.0):
    """Poisson stream of n requests (1-3 orbital Kanamori, U ~ U[1,6], J = U/8, half filling,
    beta ~ U[10,100], Lambda = 20 beta, eps = 1e-10, NCA 70 % / OCA 30 %, full self-consistency to
    1e-10). The arrival rate is 1.5x the single-stream solve rate measured on the first 16 requests
    solved one at a time; the same stream also runs at a fixed arrival rate (its latencies compare
    across builds at one offered load) and with every request at t = 0 (capacity)."""
    from xca_amd import models
    from xca_amd.serving import Server, run_stream
    t0 = time.perf_counter()
    reqs = models.serving_requests(n, seed=seed)
    log(f"serving: built {n} requests in {time.perf_counter() - t0:.1f} s")
    ncal = min(16, n)
    one = Server(max_active=1, tile_variant=0)
    cal = run_stream(one, reqs[:ncal], rate=1e6, seed=seed)
    single_rate = ncal / cal["span_s"]
    del one
    rate = 1.5 * single_rate
    log(f"serving: single-stream {single_rate:.2f} solves/s, arrival rate {rate:.2f}/s")
    sv = Server(max_active=64, tile_variant=0)
    out = run_stream(sv, reqs, rate=rate, seed=seed, log=log)
    iters = [r["iters"] for r in out["results"]]
    conv = sum(r["converged"] for r in out["results"])
    res = {"workload": "serving stream: 1-3 orbital Kanamori, U~U[1,6], J=U/8, half filling, beta~U[10,100], Lambda=20beta, "
                       "eps=1e-10, NCA 70%/OCA 30%, self-consistency to 1e-10, Poisson arrivals",
           "requests": n, "seed": seed, "arrival_rate_per_s": rate, "single_stream_solves_per_s": single_rate,
           "solves_per_s": out["solves_per_s"], "p50_s": out["p50_s"], "p99_s": out["p99_s"], "span_s": out["span_s"],
           "converged": int(conv), "mean_iterations": float(np.mean(iters)), "max_active": out["max_active"],
           "iterations_run": out["stats"]["ticks"], "max_active_slots": 64}
    del sv
    if fixed_rate > 0:
        sv = Server(max_active=64, tile_variant=0)
        fx = run_stream(sv, reqs, rate=fixed_rate, seed=seed)
        del sv
        res["fixed_rate"] = {"arrival_rate_per_s": fixed_rate, "solves_per_s": fx["solves_per_s"], "p50_s": fx["p50_s"],
                             "p99_s": fx["p99_s"], "span_s": fx["span_s"], "max_active": fx["max_active"]}
        log(f"serving at {fixed_rate:.1f} arrivals/s: {fx['solves_per_s']:.2f} solves/s p50 {fx['p50_s']:.3f} s p99 {fx['p99_s']:.3f} s")
    # capacity: every request arrives at t = 0, so the batch runs full from the start
    sv = Server(max_active=64, tile_variant=0)
    sat = run_stream(sv, reqs, rate=1e9, seed=seed)
    del sv
    res["saturated"] = {"arrivals": "all requests at t = 0", "solves_per_s": sat["solves_per_s"], "p50_s": sat["p50_s"],
                        "p99_s": sat["p99_s"], "span_s": sat["span_s"], "max_active": sat["max_active"],
                        "over_single_stream": sat["solves_per_s"] / single_rate}
    log(f"serving saturated: {sat['solves_per_s']:.2f} solves/s p50 {sat['p50_s']:.3f} s p99 {sat['p99_s']:.3f} s")
    if cpu_sample > 0:
        res["cpu_sample"] = serving_cpu_sample(reqs[:cpu_sample], iters[:cpu_sample], threads)
    return res


def serving_cpu_sample(reqs, iters, threads):
    """The restated reference solver on the first requests of the stream, one request per host core
    as independent solver processes would run: per request, the CPU time of one iteration (Sigma over
    its orders in complex128, timed on a seeded sample of at most 64 flat indices per topology and
    scaled by the topology's count; every block's Dyson solve) times the iteration count the device
    took for it (the device and oracle loops take the same iterations)."""
    from concurrent.futures import ThreadPoolExecutor
    from xca_amd.diag import all_connected_pairings
    po = _po()

    def one(a):
        (P, order), n_it = a
        orc = po.OracleEvaluator(P, complex_arith=True)
        rng = np.random.default_rng(17)
        sig = 0.0
        for o in range(1, order + 1):
            for _, topo in all_connected_pairings(o):
                t = np.array(topo, dtype=np.int32)
                nf = orc.num_backbones(o)
                fix = np.arange(nf, dtype=np.int32) if nf <= 64 else np.sort(rng.choice(nf, 64, replace=False)).astype(np.int32)
                t0 = time.perf_counter()
                orc.sigma(P.G0, t, fix)
                sig += (time.perf_counter() - t0) * nf / len(fix)
        t0 = time.perf_counter()
        for g0 in P.G0:
            po.dyson(P.beta, P.grid.it, P.grid.rf, g0, np.zeros_like(g0), 0.0)
        dys = time.perf_counter() - t0
        return n_it * (sig + dys)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        lat = np.array(list(ex.map(one, zip(reqs, iters))))
    return {"requests": len(reqs), "cores_per_request": 1, "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "solves_per_s_16_cores": 16.0 / float(lat.mean()), "kind": "port",
            "sample": f"the first {len(reqs)} requests of the seed-7 stream; per-iteration CPU cost (Sigma sampled on <= 64 flat "
                      "indices per topology, Dyson of every block) times the device's iteration count"}


def plain_gpu_leg(ev, G, max_order, batched_per_s, wide, budget_s=6.0, sample=3000):
    """The plain one-problem-at-a-time GPU forms of the same Sigma evaluation, beside the batched value:
    - per_topology_calls: the solver's topology loop through the drop-in call
      compute_self_energy(G, topology) with host arrays, one topology after another
      (block_sparse_solver.py:629-640);
    - per_item_launches: one chain launch per (path, pole tuple) item (launch mode 1 on explicit
      flat-index subsets), timed on a seeded sample of flat indices per topology and scaled by the
      topology's item count, since path cost does not depend on the poles. Not applicable to the
      large-block engine, whose launches are per program already."""
    from xca_amd.diag import all_connected_pairings
    topos = [(s, np.array(t, dtype=np.int32)) for o in range(1, max_order + 1) for s, t in all_connected_pairings(o)]
    items = []
    for _, t in topos:  # plans built, item counts
        ev.compute_self_energy(G, t)
        items.append(ev.last_stats()[0])
    reps, t0 = 0, time.perf_counter()
    while reps == 0 or (time.perf_counter() - t0 < budget_s and reps < 50):
        for _, t in topos:
            ev.compute_self_energy(G, t)
        reps += 1
    per_topo = reps / (time.perf_counter() - t0)
    out = {"per_topology_calls": {"evals_per_s": per_topo, "reps": reps, "batched_over_plain": batched_per_s / per_topo}}
    if wide:
        out["per_item_launches"] = None
        return out
    rng = np.random.default_rng(13)
    ev.set_launch_mode(1)
    s_per_eval, sampled, launches_s = 0.0, 0, 0.0
    try:
        for (_, t), n_items in zip(topos, items):
            nf = ev.get_num_self_energy_backbones(t)
            fix = np.arange(nf, dtype=np.int32) if nf <= sample else np.sort(rng.choice(nf, size=sample, replace=False)).astype(np.int32)
            ev.compute_self_energy(G, t, fix)  # builds the subset plan
            t1 = time.perf_counter()
            ev.compute_self_energy(G, t, fix)
            dt = time.perf_counter() - t1
            k = ev.last_stats()[0]
            if k > 0:
                s_per_eval += dt / k * n_items
                sampled += k
                launches_s += dt
    finally:
        ev.set_launch_mode(0)
    out["per_item_launches"] = {"evals_per_s": 1.0 / s_per_eval, "s_per_eval": s_per_eval, "items_timed": sampled,
                                "timed_s": launches_s, "batched_over_plain": batched_per_s * s_per_eval,
                                "sample": f"up to {sample} seeded flat indices per topology, scaled by its item count"}
    return out


def spgf_leg(ev, P, G, max_order, n, cpu):
    """Single-particle Green's function g(tau) of the same workload and G: every connected topology of
    orders 1..max_order with the solver's signs (block_sparse_solver.py:657-677), timed over n
    evaluations; the chain launch of the costliest topology is priced on the FP64 peak. With cpu, one
    oracle evaluation (single core, complex128) is timed beside it."""
    from xca_amd.diag import all_connected_pairings
    topos = [(np.array(t, dtype=np.int32), (-1) ** o * s) for o in range(1, max_order + 1) for s, t in all_connected_pairings(o)]

    ev.set_G(G)  # resident in HBM for every evaluation

    def run(record=None):
        tot = 0
        for t, sc in topos:
            tot = tot + sc * ev.compute_single_ptcle_gf(None, t)
            if record is not None:
                record.append(ev.last_stats())
        return tot

    run()
    t0 = time.perf_counter()
    for _ in range(n):
        g = run()
    ms = 1e3 * (time.perf_counter() - t0) / n
    recs = []
    run(recs)
    dom = max(range(len(recs)), key=lambda i: recs[i][2])
    items, macs, kms = recs[dom]
    achieved = 2.0 * macs / (kms * 1e-3) / 1e12
    out = {"workload": f"g(tau) of the benchmark workload, orders 1-{max_order}, all connected topologies, same G",
           "evals_per_s": 1e3 / ms, "ms_per_eval": ms, "items_per_eval": int(sum(x[0] for x in recs)),
           "canonical_macs_per_eval": float(sum(x[1] for x in recs)),
           "dominant": {"kernel": "chain_kernel (store mode)", "topology": topos[dom][0].tolist(), "kernel_ms": kms,
                        "flops_per_launch": 2.0 * macs, "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64},
           "g_max": float(np.abs(g).max())}
    if cpu:
        sys.path.insert(0, str(ROOT / "oracle"))
        import pyoracle as po
        orc = po.OracleEvaluator(P, complex_arith=True)
        t0 = time.perf_counter()
        po.spgf_sum(orc, G, max_order)
        out["cpu_oracle_s_per_eval"] = time.perf_counter() - t0
        out["cpu_cores"] = 1
    return out


WORKLOAD_TEXT = {
    "oca2": "OCA 2-orbital Kanamori U=4 J=0.5 half filling, beta=50, Lambda=1000, eps=1e-10, orders 1-2, one Sigma evaluation per step",
    "tca3": "TCA 3-orbital Kanamori U=5 J=0.7 filling 2, beta=100, Lambda=2000, eps=1e-10, offdiagonal 0.05, orders 1-3, one Sigma "
            "evaluation per step",
    "oca5": "OCA 5-orbital d shell, Slater-Condon F0=4 F2=6 F4=3.75, half filling, beta=40, Lambda=800, eps=1e-10, 36 (N,Sz) blocks "
            "up to 100, orders 1-2, one Sigma evaluation per step (large-block engine)",
    "tca3c": "TCA 3-orbital Kanamori as tca3 with the hybridization compressed to p < r poles (relative fit error 1e-10), "
             "orders 1-3, one Sigma evaluation per step",
}


def workload_leg(name, steps, cpu, threads, cpu_budget, cpu_cores):
    """One of the larger configurations beside the headline: seconds per full Sigma evaluation (all
    orders and topologies with the solver's signs, G resident, one launch sequence), the dominant
    launch's canonical FP64 rate from HIP events, the plain per-topology GPU form, the accuracy of the
    device Sigma against the oracle, and the sampled CPU oracle. The three-orbital workload runs on
    the inputs committed with its extended-precision reference (tests/golden/tca3_extended.npz), so
    that its accuracy is also stated against the long-double evaluation of the same doubles."""
    from xca_amd.diag import all_connected_pairings
    from xca_amd.evaluator import DiagramEvaluator
    fx = None
    if name in ("tca3", "tca3c"):
        sys.path.insert(0, str(ROOT / "tests" / "golden"))
        from make_extended_fixtures import problem_from_fixture
        fx = np.load(ROOT / "tests" / "golden" / "tca3_extended.npz")
        P, G = problem_from_fixture(fx)
        max_order = 3
        if name == "tca3c":
            from xca_amd import models
            t0 = time.perf_counter()
            P = models.compressed(P, 1e-10)
            fx = None
            log(f"tca3c: hybridization compressed to {len(P.poles)} poles in {time.perf_counter() - t0:.2f} s")
    else:
        P, max_order = build_workload(name)
        G = P.perturbed_G()
    ev = DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)
    ev.set_G(G)
    t0 = time.perf_counter()
    ev.sigma_order(max_order)
    ev.synchronize()
    first_s = time.perf_counter() - t0
    times, recs = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        ev.sigma_order(max_order, timed=True)
        ev.synchronize()
        times.append(time.perf_counter() - t0)
        recs.append(ev.last_stats())
    items, macs, _ = recs[0]
    kms = float(np.mean([x[2] for x in recs]))
    s_per = float(np.mean(times))
    achieved = 2.0 * macs / (kms * 1e-3) / 1e12
    out = {"workload": WORKLOAD_TEXT[name], "dlr_rank": P.grid.r, "poles": len(P.poles), "blocks": [int(d) for d in P.dims],
           "s_per_eval": s_per, "evals_per_s": 1.0 / s_per, "timed_evals": steps, "first_eval_s_with_planning": first_s,
           "items_per_eval": int(items), "canonical_macs_per_eval": macs,
           "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                        "traffic": measured_traffic(name), "kernel_ms": kms, "flops_per_launch": 2.0 * macs,
                        "kernel": "chain_kernel, one launch for all topologies" if name != "oca5" else
                                  "large-block launch sequence (wgemm_kernel dominant)"}}
    log(f"{name}: {s_per:.3f} s per Sigma, {achieved:.2f} TFLOP/s ({achieved / PEAK_FP64:.3f} of peak)")
    # one Dyson solve of every block with this Sigma (host arrays in and out); for the 5-orbital shell
    # the blocks above 12 states take the device-wide path (LU of systems up to r n = 4900)
    sig = ev.get_sigma()
    ev.dyson_init(P.G0)
    ev.dyson_solve(sig, 0.0)
    t0 = time.perf_counter()
    ev.dyson_solve(sig, 0.0)
    out["dyson"] = {"s_per_solve": time.perf_counter() - t0, "largest_system": int(P.grid.r * max(P.dims)),
                    "note": "all blocks, host arrays in and out, after one untimed solve"}
    log(f"{name}: Dyson solve of every block {out['dyson']['s_per_solve']:.3f} s (largest system {out['dyson']['largest_system']})")
    topos = [np.array(t, dtype=np.int32) for o in range(1, max_order + 1) for _, t in all_connected_pairings(o)]
    t0 = time.perf_counter()
    for t in topos:
        ev.compute_self_energy(G, t)
    plain = time.perf_counter() - t0
    out["plain_gpu"] = {"per_topology_calls": {"s_per_eval": plain, "batched_over_plain": plain / s_per}}
    if name == "oca5":
        # g(tau) of the same G, orders 1 to max_order, on the large-block engine (chains in store
        # mode, reflection, contraction), after a first-order warm-up evaluation
        ev.single_particle_greens_function(G, 1)
        t0 = time.perf_counter()
        ev.single_particle_greens_function(G, max_order)
        out["spgf"] = {"s_per_eval": time.perf_counter() - t0, "orders": f"1-{max_order}", "engine": "large-block"}
        log(f"{name}: g(tau) orders 1-{max_order} in {out['spgf']['s_per_eval']:.2f} s")
    if cpu:
        po = _po()
        acc = []

        def dev_runs(t, fix, runs):  # the same contiguous runs, one call each, on both sides
            parts = [ev.compute_self_energy(G, t, run) for run in po.split_runs(fix, runs)]
            return [sum(p[b] for p in parts) for b in range(len(parts[0]))]

        if fx is not None:
            runs = int(fx["runs"])
            for k in range(6):
                t, fix = fx[f"topo{k}"], fx[f"fix{k}"]
                ext = P.unpack(fx[f"sigma{k}"])
                dev = dev_runs(t, fix, runs)
                orc = po.sigma_threads(P, G, t, fix, threads=threads, runs=runs)
                acc.append({"topology": t.tolist(), "flat_indices": int(len(fix)), "vs_oracle": rel_errors(dev, orc),
                            "vs_extended": rel_errors(dev, ext)["global"], "oracle_vs_extended": rel_errors(orc, ext)["global"]})
        else:
            if name == "tca3c":
                rng = np.random.default_rng(21)
                samples = [(t, np.sort(rng.choice(ev.get_num_self_energy_backbones(t), 128, replace=False)).astype(np.int32))
                           for t in topos if len(t) == 3]
            else:
                samples = [(np.array([[0, 1]], dtype=np.int32), np.arange(2, dtype=np.int32)),
                           (np.array([[0, 2], [1, 3]], dtype=np.int32), np.array([7, 1203], dtype=np.int32))]
            for t, fix in samples:
                dev = dev_runs(t, fix, 16)
                orc = po.sigma_threads(P, G, t, fix, threads=threads, runs=16)
                ext = po.sigma_threads(P, G, t, fix, threads=threads, extended=True, runs=16)
                acc.append({"topology": t.tolist(), "flat_indices": fix.tolist(), "vs_oracle": rel_errors(dev, orc),
                            "vs_extended": rel_errors(dev, ext)["global"], "oracle_vs_extended": rel_errors(orc, ext)["global"]})
        for a in acc:
            a["vs_oracle"].pop("per_block")
        out["accuracy"] = {"max_rel_err": max(a["vs_oracle"]["global"] for a in acc),
                           "max_rel_err_per_block": max(a["vs_oracle"]["per_block_max"] for a in acc),
                           "max_rel_err_vs_extended": max(a["vs_extended"] for a in acc),
                           "oracle_max_rel_err_vs_extended": max(a["oracle_vs_extended"] for a in acc),
                           "note": "device Sigma on the listed flat indices against the double-precision oracle and against the "
                                   "oracle in long double; at this beta the atomic propagator is not resolved by the DLR grid and "
                                   "the double-precision oracle itself is oracle_max_rel_err_vs_extended from the exact result",
                           "samples": acc}
        log(f"{name}: accuracy vs oracle {out['accuracy']['max_rel_err']:.2e}, vs long double {out['accuracy']['max_rel_err_vs_extended']:.2e} "
            f"(oracle {out['accuracy']['oracle_max_rel_err_vs_extended']:.2e})")
        out["cpu_baseline"] = cpu_baseline(P, G, max_order, cpu_budget, cpu_cores)
        out["cpu_baseline"]["gpu_over_cpu"] = out["cpu_baseline"]["value"] and (1.0 / s_per) / out["cpu_baseline"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="oca2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle CPU work for cpu_baseline")
    ap.add_argument("--cpu-cores", type=int, default=16,
                    help="host threads for the CPU legs (one oracle evaluator each), capped at the cores this process may use")
    ap.add_argument("--solves", type=int, default=50, help="NCA self-consistent solves timed for the latency figures (0: skip)")
    ap.add_argument("--serving", type=int, default=512, help="requests in the served stream (0: skip)")
    ap.add_argument("--serving-fixed-rate", type=float, default=43.0,
                    help="arrival rate of the fixed-load run of the served stream (0: skip)")
    ap.add_argument("--spgf", type=int, default=5, help="timed single-particle Green's function evaluations (0: skip)")
    ap.add_argument("--plain", type=int, default=1, help="time the plain one-problem-at-a-time GPU forms (0: skip)")
    ap.add_argument("--legs", default="tca3,tca3c,oca5", help="larger configurations measured beside the headline ('none' to skip)")
    ap.add_argument("--leg-steps", type=int, default=2, help="timed Sigma evaluations per larger configuration")
    ap.add_argument("--spgf-only", action="store_true", help="run only the single-particle Green's function leg (profiling)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (reference dispatches for the HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (the Sigma evaluation "
              "would shard by backbone index with one all-reduce of the Sigma blocks)", file=sys.stderr)
        sys.exit(2)

    from xca_amd.evaluator import DiagramEvaluator
    from xca_amd.diag import all_connected_pairings

    args.cpu_cores = host_threads(args.cpu_cores)
    P, max_order = build_workload(args.workload)
    G = P.perturbed_G()
    ev = DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)
    if args.spgf_only:
        print(json.dumps({"spgf": spgf_leg(ev, P, G, max_order, max(args.spgf, 1), cpu=False)}))
        return
    ev.set_G(G)

    def step():
        # Sigma over orders 1..max_order with the solver's signs: every topology's edge tables in one
        # launch, all chunks in one persistent chain launch, one reduction; replayed as a HIP graph
        ev.sigma_order(max_order)

    for _ in range(args.warmup):
        step()
    ev.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ev.synchronize()
    el = time.perf_counter() - t0
    ms_per_step = 1e3 * el / args.steps

    # kernel-level measurement of the dominant launch: the chain launch of one evaluation (all
    # topologies), bracketed by HIP events on the launch stream (eager replays of the same sequence)
    recs = []
    for _ in range(min(args.steps, 50)):
        ev.sigma_order(max_order, timed=True)
        ev.synchronize()
        recs.append(ev.last_stats())
    items, macs, _ = recs[0]
    kms = float(np.mean([x[2] for x in recs]))
    achieved = 2.0 * macs / (kms * 1e-3) / 1e12
    total_macs = macs
    topo_list = [t for o in range(1, max_order + 1) for _, t in all_connected_pairings(o)]

    sig = ev.get_sigma()
    accuracy = None
    if not args.no_cpu and args.workload == "oca2":
        # the timed Sigma against the oracle's full evaluation of the same G (double precision)
        accuracy = rel_errors(sig, signed_oracle_sigma(P, G, max_order, args.cpu_cores))
        log(f"accuracy of the timed Sigma: {accuracy['global']:.2e} (worst block {accuracy['per_block_max']:.2e})")
    out = {
        "metric": METRIC,
        "value": 1e3 / ms_per_step,
        "unit": "Σ-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOAD_TEXT[args.workload],
                   "dlr_rank": P.grid.r, "poles": len(P.poles), "blocks": [int(d) for d in P.dims],
                   "items_per_eval": int(items),
                   "canonical_macs_per_eval": total_macs},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64,
                     "traffic": measured_traffic(args.workload),
                     "kernel": ("chain_kernel, one launch for all topologies of the evaluation" if args.workload != "oca5"
                                else "large-block launch sequence of the evaluation (wgemm_kernel dominant)"),
                     "kernel_ms": kms,
                     "flops_per_launch": 2.0 * macs, "topologies": [[list(p) for p in t] for t in topo_list]},
    }
    if accuracy is not None:
        out["max_rel_err"] = accuracy["global"]
        out["max_rel_err_per_block"] = accuracy["per_block"]
    if args.plain:
        out["plain_gpu"] = plain_gpu_leg(ev, G, max_order, out["value"], wide=args.workload == "oca5")
        pg = out["plain_gpu"]
        log(f"plain GPU forms: per-topology calls {pg['per_topology_calls']['evals_per_s']:.2f} evals/s"
            + (f", per-item launches {pg['per_item_launches']['evals_per_s']:.4g} evals/s" if pg["per_item_launches"] else ""))
        ev.set_G(G)
    if args.spgf > 0:
        out["spgf"] = spgf_leg(ev, P, G, max_order, args.spgf, cpu=not args.no_cpu and args.workload == "oca2")
        log(f"spgf {out['spgf']['evals_per_s']:.1f} evals/s, dominant chain {out['spgf']['dominant']['frac']:.3f} of peak")
    if args.solves > 0:
        out["solve"] = solve_latency(args.solves, cpu=not args.no_cpu)
        log(f"solve latency p50 {out['solve']['p50_ms']:.2f} ms p99 {out['solve']['p99_ms']:.2f} ms")
    if args.serving > 0:
        out["serving"] = serving(args.serving, cpu_sample=0 if args.no_cpu else 16, threads=args.cpu_cores,
                                 fixed_rate=args.serving_fixed_rate)
        log(f"serving {out['serving']['solves_per_s']:.2f} solves/s p50 {out['serving']['p50_s']:.3f} s p99 {out['serving']['p99_s']:.3f} s")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(P, G, max_order, args.cpu_budget, args.cpu_cores)
        out["cpu_baseline"]["host_cores_visible"] = len(os.sched_getaffinity(0))
    del ev
    legs = [x for x in args.legs.split(",") if x and x not in ("none", args.workload)]
    if legs:
        out["workloads"] = {}
        for name in legs:
            out["workloads"][name] = workload_leg(name, args.leg_steps, not args.no_cpu, args.cpu_cores, 10.0, args.cpu_cores)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
