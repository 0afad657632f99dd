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
        k, done, el = 16 * cores, 0, 0.0
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


def solve_latency(n):
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
    return {"workload": "NCA 1-orbital Anderson U=2 mu=1 beta=10 Lambda=100 eps=1e-10, full self-consistency to 1e-10",
            "solves": n, "iterations": int(iters), "p50_ms": float(np.percentile(t, 50)), "p99_ms": float(np.percentile(t, 99)),
            "mean_ms": float(t.mean())}


def serving(n, seed=7):
    """Poisson stream of n requests (1-3 orbital Kanamori, U ~ U[1,6], J = U/8, half filling,
    beta ~ U[10,100], Lambda = 20 beta, eps = 1e-10, NCA 70 % / OCA 30 %, full self-consistency to
    1e-10). The arrival rate is 1.5x the single-stream solve rate measured on the first 16 requests
    solved one at a time."""
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
    return {"workload": "serving stream: 1-3 orbital Kanamori, U~U[1,6], J=U/8, half filling, beta~U[10,100], Lambda=20beta, "
                        "eps=1e-10, NCA 70%/OCA 30%, self-consistency to 1e-10, Poisson arrivals",
            "requests": n, "seed": seed, "arrival_rate_per_s": rate, "single_stream_solves_per_s": single_rate,
            "solves_per_s": out["solves_per_s"], "p50_s": out["p50_s"], "p99_s": out["p99_s"], "span_s": out["span_s"],
            "converged": int(conv), "mean_iterations": float(np.mean(iters)), "max_active": out["max_active"],
            "iterations_run": out["stats"]["ticks"], "max_active_slots": 64}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="oca2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of oracle CPU work for cpu_baseline")
    ap.add_argument("--cpu-cores", type=int, default=min(16, len(os.sched_getaffinity(0))),
                    help="host threads for cpu_baseline (one oracle evaluator each)")
    ap.add_argument("--solves", type=int, default=50, help="NCA self-consistent solves timed for the latency figures (0: skip)")
    ap.add_argument("--serving", type=int, default=512, help="requests in the served stream (0: skip)")
    ap.add_argument("--spgf", type=int, default=5, help="timed single-particle Green's function evaluations (0: skip)")
    ap.add_argument("--plain", type=int, default=1, help="time the plain one-problem-at-a-time GPU forms (0: skip)")
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
        "config": {"workload": {"oca2": "OCA 2-orbital Kanamori U=4 J=0.5 half filling, beta=50, Lambda=1000, eps=1e-10, "
                                        "orders 1-2, one Sigma evaluation per step",
                                "tca3": "TCA 3-orbital Kanamori U=5 J=0.7 filling 2, beta=100, Lambda=2000, eps=1e-10, "
                                        "orders 1-3, one Sigma evaluation per step",
                                "oca5": "OCA 5-orbital d shell, Slater-Condon F0=4 F2=6 F4=3.75, half filling, beta=40, "
                                        "Lambda=800, eps=1e-10, 36 (N,Sz) blocks up to 100, orders 1-2, one Sigma evaluation "
                                        "per step (large-block engine)"}[args.workload],
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
        out["solve"] = solve_latency(args.solves)
        log(f"solve latency p50 {out['solve']['p50_ms']:.2f} ms p99 {out['solve']['p99_ms']:.2f} ms")
    if args.serving > 0:
        out["serving"] = serving(args.serving)
        log(f"serving {out['serving']['solves_per_s']:.2f} solves/s p50 {out['serving']['p50_s']:.3f} s p99 {out['serving']['p99_s']:.3f} s")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(P, G, max_order, args.cpu_budget, args.cpu_cores)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
