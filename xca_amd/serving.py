"""Serving: many impurity problems solved to self-consistency on one GPU with continuous batching.

Server wraps the xca_server_* C ABI: submitted problems queue on the host, are admitted as batch slots
and the device-memory budget allow, iterate together in shared batched launches (one HIP graph per
iteration while the set of problems is unchanged) and leave the batch as soon as they converge.

run_stream drives a Server with Poisson arrivals and reports solves/s and the p50/p99 time from
arrival to converged result.
"""
import ctypes as C
import time

import numpy as np

from ._lib import XcaError, check, dptr, i64_p, last_error, lib
from .evaluator import _real, problem_struct


class Server:
    def __init__(self, device=0, mem_budget_bytes=0, max_active=64, tile_variant=0):
        h = lib().xca_server_create(int(device), int(mem_budget_bytes), int(max_active), int(tile_variant))
        if not h:
            raise XcaError(last_error())
        self._h = h
        self._req = {}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().xca_server_destroy(h)
            self._h = None

    def submit(self, prob, max_order, tol=1e-10, maxiter=100, mix=1.0, eta0=0.0):
        """Queue an ImpurityProblem (xca_amd.models); returns the request id. eta0 is the initial
        pseudo-particle chemical potential on top of the normalised atomic G0 (0, as the reference's
        solver starts, block_sparse_solver.py:103)."""
        keep = []

        def k(a):
            keep.append(a)
            return a

        d = problem_struct(prob.poles, prob.coeffs, prob.grid, prob.ad, k)
        g0 = k(np.ascontiguousarray(np.concatenate([_real(g, "G0").ravel() for g in prob.G0])))
        rid = lib().xca_server_submit(self._h, C.byref(d), dptr(g0), float(eta0), int(max_order), float(tol),
                                      int(maxiter), float(mix))
        if rid < 0:
            raise XcaError(last_error())
        self._req[rid] = (prob.grid.r, [int(b) for b in prob.dims])
        return rid

    def step(self):
        a, q, f = C.c_int(0), C.c_int(0), C.c_int(0)
        check(lib().xca_server_step(self._h, C.byref(a), C.byref(q), C.byref(f)))
        return a.value, q.value, f.value

    def poll(self, max_ids=4096):
        ids = np.zeros(max_ids, dtype=np.int64)
        n = check(lib().xca_server_poll(self._h, ids.ctypes.data_as(i64_p), max_ids))
        return [int(x) for x in ids[:n]]

    def result(self, rid):
        r, dims = self._req.pop(rid)
        out = np.zeros(sum(r * b * b for b in dims))
        eta, diff = C.c_double(0), C.c_double(0)
        iters, conv = C.c_int(0), C.c_int(0)
        times = np.zeros(3)
        check(lib().xca_server_result(self._h, int(rid), dptr(out), C.byref(eta), C.byref(iters), C.byref(diff), C.byref(conv), dptr(times)))
        G, off = [], 0
        for b in dims:
            G.append(out[off:off + r * b * b].reshape(r, b, b))
            off += r * b * b
        return {"G": G, "eta": eta.value, "iters": iters.value, "diff": diff.value, "converged": bool(conv.value),
                "queued_s": times[0], "run_s": times[1], "total_s": times[2]}

    def request_bytes(self, rid):
        """Estimated device footprint of a submitted request (an upper bound, used for admission)."""
        return int(lib().xca_server_request_bytes(self._h, int(rid)))

    def shared_bytes(self):
        """Bytes of the budget reserved for the launch buffers all requests share."""
        return int(lib().xca_server_shared_bytes(self._h))

    def request_device_bytes(self, rid):
        """Device bytes an admitted request holds now; -1 while queued or after it finished."""
        return int(lib().xca_server_request_device_bytes(self._h, int(rid)))

    def stats(self):
        t, u, b = C.c_int64(0), C.c_int64(0), C.c_int64(0)
        check(lib().xca_server_stats(self._h, C.byref(t), C.byref(u), C.byref(b)))
        return {"ticks": t.value, "bytes_in_use": u.value, "budget": b.value}


def run_stream(server, requests, rate, seed=7, tol=1e-10, maxiter=200, mix=1.0, log=None, check_footprint=False):
    """Submit `requests` [(ImpurityProblem, order), ...] with Poisson arrivals at `rate` per second and
    step the server until every result is back. Latency = time from a request's arrival to its
    converged result (queueing included). Returns a dict of results and summary figures;
    check_footprint also records the largest ratio of a request's measured device bytes to its
    admission estimate over every step."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(requests)))
    arrivals -= arrivals[0]
    t0 = time.perf_counter()
    nxt, done = 0, 0
    rid_of = {}
    results = [None] * len(requests)
    finish = np.zeros(len(requests))
    max_active_seen = 0
    max_bytes = 0
    max_ratio = 0.0
    while done < len(requests):
        now = time.perf_counter() - t0
        while nxt < len(requests) and arrivals[nxt] <= now:
            prob, order = requests[nxt]
            rid_of[server.submit(prob, order, tol=tol, maxiter=maxiter, mix=mix)] = nxt
            nxt += 1
        a, q, f = server.step()
        max_active_seen = max(max_active_seen, a + f)
        max_bytes = max(max_bytes, server.stats()["bytes_in_use"])
        if check_footprint:
            for rid in rid_of:
                used = server.request_device_bytes(rid)
                if used >= 0:
                    max_ratio = max(max_ratio, used / server.request_bytes(rid))
        for rid in server.poll():
            i = rid_of.pop(rid)
            results[i] = server.result(rid)
            finish[i] = time.perf_counter() - t0
            done += 1
            if log and done % 64 == 0:
                log(f"serving: {done}/{len(requests)} done, {a} active, {q} queued")
        if a == 0 and q == 0 and nxt < len(requests):
            time.sleep(max(0.0, min(0.01, arrivals[nxt] - (time.perf_counter() - t0))))
    lat = finish - arrivals
    span = finish.max() - arrivals[0]
    return {"results": results, "latency_s": lat, "arrivals_s": arrivals, "finish_s": finish, "solves_per_s": len(requests) / span,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)), "span_s": float(span),
            "max_active": max_active_seen, "max_bytes_in_use": max_bytes, "max_footprint_ratio": max_ratio, "stats": server.stats()}
