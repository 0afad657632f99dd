"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so; liboracle_ld.so is the same source
computed in long double, for reference values where double rounding limits the result).

ORACLE — test infrastructure only: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the checker; never by the product package.
"""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
LIB_LD = HERE / "_build" / "liboracle_ld.so"

d_p = C.POINTER(C.c_double)
i_p = C.POINTER(C.c_int)


class OrcProblem(C.Structure):
    _fields_ = [
        ("beta", C.c_double), ("r", C.c_int), ("it", d_p), ("rf", d_p),
        ("p", C.c_int), ("poles", d_p), ("coeffs", d_p),
        ("n", C.c_int), ("nblocks", C.c_int), ("dims", i_p),
        ("c_conn", i_p), ("cdag_conn", i_p), ("c_mats", d_p), ("cdag_mats", d_p),
    ]


_libs = {}


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib(extended=False):
    """The oracle library: double arithmetic, or long double when extended."""
    path = LIB_LD if extended else LIB
    if path not in _libs:
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        L.orc_last_error.restype = C.c_char_p
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcProblem), C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_num_backbones.restype = C.c_long
        L.orc_num_backbones.argtypes = [C.c_void_p, C.c_int]
        L.orc_num_sym_sets.argtypes = [C.c_void_p]
        L.orc_sigma.argtypes = [C.c_void_p, d_p, C.c_int, i_p, i_p, C.c_long, d_p]
        L.orc_vals2coefs.argtypes = [C.c_int, d_p, d_p, d_p, C.c_int]
        L.orc_convolve.argtypes = [C.c_double, C.c_int, d_p, d_p, d_p, d_p, C.c_int, C.c_int, C.c_int, d_p]
        L.orc_reflect.argtypes = [C.c_int, d_p, d_p, d_p, C.c_int, d_p]
        L.orc_interp.argtypes = [C.c_int, d_p, d_p, d_p, C.c_int, d_p, C.c_int, d_p]
        L.orc_dyson.argtypes = [C.c_double, C.c_int, d_p, d_p, C.c_int, d_p, d_p, C.c_double, d_p]
        L.orc_backbone.argtypes = [i_p, C.c_int, C.c_int, d_p, C.c_int, C.c_int, i_p, i_p, i_p, i_p]
        L.orc_last_dropped.argtypes = [C.c_void_p, d_p]
        L.orc_num_correlator_backbones.restype = C.c_long
        L.orc_num_correlator_backbones.argtypes = [C.c_void_p, C.c_int]
        L.orc_correlator.argtypes = [C.c_void_p, d_p, C.c_int, i_p, i_p, C.c_long, C.c_int, i_p, d_p, C.c_int, i_p, d_p, d_p]
        _libs[path] = L
    return _libs[path]


class OracleError(RuntimeError):
    pass


def _chk(st, L=None):
    if st != 0:
        raise OracleError((L or lib()).orc_last_error().decode())


def _z(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _dp(a):
    return a.ctypes.data_as(d_p)


def _ip(a):
    return a.ctypes.data_as(i_p)


# ------------------------------------------------------------------ DLR unit operations

def vals2coefs(it, rf, vals):
    v = _z(vals).copy()
    m = v.size // len(it)
    _chk(lib().orc_vals2coefs(len(it), _dp(np.ascontiguousarray(it)), _dp(np.ascontiguousarray(rf)), _dp(v), m))
    return v


def convolve(beta, it, rf, f, g):
    f, g = _z(f), _z(g)
    r, a, b = f.shape
    _, b2, c = g.shape
    out = np.zeros((r, a, c), dtype=np.complex128)
    _chk(lib().orc_convolve(beta, r, _dp(np.ascontiguousarray(it)), _dp(np.ascontiguousarray(rf)), _dp(f), _dp(g), a, b, c, _dp(out)))
    return out


def reflect(it, rf, vals):
    v = _z(vals)
    out = np.zeros_like(v)
    _chk(lib().orc_reflect(len(it), _dp(np.ascontiguousarray(it)), _dp(np.ascontiguousarray(rf)), _dp(v), v.size // len(it), _dp(out)))
    return out


def interp(it, rf, vals, ts, extended=False):
    v = _z(vals)
    ts = np.ascontiguousarray(ts, dtype=np.float64)
    m = v.size // len(it)
    out = np.zeros((len(ts),) + v.shape[1:], dtype=np.complex128)
    L = lib(extended)
    _chk(L.orc_interp(len(it), _dp(np.ascontiguousarray(it)), _dp(np.ascontiguousarray(rf)), _dp(v), m, _dp(ts), len(ts), _dp(out)), L)
    return out


def dyson(beta, it, rf, g0, sig, eta, extended=False):
    g0, sig = _z(g0), _z(sig)
    r, b, _ = g0.shape
    out = np.zeros_like(g0)
    L = lib(extended)
    _chk(L.orc_dyson(beta, r, _dp(np.ascontiguousarray(it)), _dp(np.ascontiguousarray(rf)), b, _dp(g0), _dp(sig), eta, _dp(out)), L)
    return out


def backbone(topology, n, poles_dimless, f_ix):
    topo = np.ascontiguousarray(topology, dtype=np.int32)
    m = topo.shape[0]
    poles = np.ascontiguousarray(poles_dimless, dtype=np.float64)
    vert = np.zeros((2 * m, 5), dtype=np.int32)
    edges = np.zeros((2 * m - 1, max(m - 1, 1)), dtype=np.int32)
    pref = np.zeros((max(m - 1, 1), 2), dtype=np.int32)
    misc = np.zeros(2 * m + 1, dtype=np.int32)
    _chk(lib().orc_backbone(_ip(topo), m, n, _dp(poles), len(poles), f_ix, _ip(vert), _ip(edges), _ip(pref), _ip(misc)))
    return dict(bar=vert[:, 0], dag=vert[:, 1], hyb_ind=vert[:, 2], Ksign=vert[:, 3], orb=vert[:, 4],
                edges=edges[:, : m - 1], pref_Ksign=pref[: m - 1, 0], pref_Kexp=pref[: m - 1, 1],
                prefactor_sign=int(misc[0]), fb=misc[1:1 + m], pole_inds=misc[1 + m:2 * m])


# ------------------------------------------------------------------ evaluator

class OracleEvaluator:
    """CPU restatement of DiagramEvaluator for one ImpurityProblem (xca_amd.models)."""

    def __init__(self, prob, complex_arith=True, extended=False):
        """extended: compute in long double (liboracle_ld.so); inputs and outputs stay float64."""
        ad = prob.ad
        self.L = lib(extended)
        self.prob = prob
        self._keep = []

        def keep(a):
            self._keep.append(a)
            return a

        it = keep(np.ascontiguousarray(prob.grid.it, dtype=np.float64))
        rf = keep(np.ascontiguousarray(prob.grid.rf, dtype=np.float64))
        poles = keep(np.ascontiguousarray(prob.poles, dtype=np.float64))
        coeffs = keep(_z(prob.coeffs))
        dims = keep(np.ascontiguousarray(ad.dims, dtype=np.int32))
        cc = keep(np.ascontiguousarray(ad.c_conn, dtype=np.int32))
        dc = keep(np.ascontiguousarray(ad.cdag_conn, dtype=np.int32))
        cm, dm = ad.packed_ops()
        cm, dm = keep(_z(cm)), keep(_z(dm))
        d = OrcProblem(prob.beta, prob.grid.r, _dp(it), _dp(rf), len(poles), _dp(poles), _dp(coeffs), prob.n, ad.nblocks, _ip(dims),
                       _ip(cc), _ip(dc), _dp(cm), _dp(dm))
        h = self.L.orc_create(C.byref(d), 1 if complex_arith else 0)
        if not h:
            raise OracleError(self.L.orc_last_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def num_backbones(self, m):
        return self.L.orc_num_backbones(self.h, m)

    def num_sym_sets(self):
        return self.L.orc_num_sym_sets(self.h)

    def sigma(self, G, topology, fix=None):
        """compute_self_energy(G, topology[, f_ix_vec]); G: list of (r, b, b) blocks."""
        prob = self.prob
        g = _z(np.concatenate([np.asarray(b).ravel() for b in G]))
        topo = np.ascontiguousarray(topology, dtype=np.int32)
        out = np.zeros_like(g)
        if fix is None:
            st = self.L.orc_sigma(self.h, _dp(g), topo.shape[0], _ip(topo), None, 0, _dp(out))
        else:
            fx = np.ascontiguousarray(fix, dtype=np.int32)
            st = self.L.orc_sigma(self.h, _dp(g), topo.shape[0], _ip(topo), _ip(fx), len(fx), _dp(out))
        _chk(st, self.L)
        return prob.unpack(out)


    def last_dropped(self):
        """Per block, the sum of max |.| over the running sums the last sigma call discarded: with a
        list of flat indices the reference re-flags blocks below 1e-16 after every index, and the
        next contribution to such a block replaces it."""
        out = np.zeros(self.prob.ad.nblocks)
        self.L.orc_last_dropped(self.h, _dp(out))
        return out

    def num_correlator_backbones(self, m):
        return self.L.orc_num_correlator_backbones(self.h, m)

    def correlator(self, G, topology, fix=None, mu_ops=None, kap_ops=None):
        """compute_one_time_correlator(G, topology[, f_ix_vec]) for block operators given as
        (conn (nops, nblocks), packed matrices), or compute_single_ptcle_gf when mu_ops is None:
        (r, nmu, nkap) complex, unsigned (the caller applies the order and topology sign)."""
        prob = self.prob
        g = _z(np.concatenate([np.asarray(b).ravel() for b in G]))
        topo = np.ascontiguousarray(topology, dtype=np.int32)
        fx = None if fix is None else np.ascontiguousarray(fix, dtype=np.int32)
        if mu_ops is None:
            nmu = nkap = prob.n
            args = (0, None, None, 0, None, None)
        else:
            mc, mm = np.ascontiguousarray(mu_ops[0], dtype=np.int32), _z(mu_ops[1])
            kc, km = np.ascontiguousarray(kap_ops[0], dtype=np.int32), _z(kap_ops[1])
            self._keep += [mc, mm, kc, km]
            nmu, nkap = mc.shape[0], kc.shape[0]
            args = (nmu, _ip(mc), _dp(mm), nkap, _ip(kc), _dp(km))
        out = np.zeros((prob.grid.r, nmu, nkap), dtype=np.complex128)
        st = self.L.orc_correlator(self.h, _dp(g), topo.shape[0], _ip(topo), None if fx is None else _ip(fx),
                                  0 if fx is None else len(fx), *args, _dp(out))
        _chk(st, self.L)
        return out

    def spgf(self, G, topology, fix=None):
        """compute_single_ptcle_gf(G, topology[, f_ix_vec])."""
        return self.correlator(G, topology, fix)


def spgf_sum(orc, G, max_order):
    """eval_single_particle_greens_function (block_sparse_solver.py:657-677): the sum over orders
    and connected topologies of (-1)^order * s_topo * compute_single_ptcle_gf."""
    from xca_amd.diag import all_connected_pairings
    tot = 0
    for order in range(1, max_order + 1):
        for s, t in all_connected_pairings(order):
            tot = tot + (-1) ** order * s * orc.spgf(G, np.array(t, dtype=np.int32))
    return tot


# ------------------------------------------------------------------ self-consistency loop

def ppsc_solve(prob, max_order, tol=1e-10, maxiter=100, mix=1.0, complex_arith=False, threads=1, extended=False):
    """BlockSparseSolver.solve with classic normalization (block_sparse_solver.py:240-355, 427-449,
    572-590) on top of the oracle's Sigma and Dyson: returns (G blocks, eta, iterations, diff). The
    per-block Dyson solves are independent and may run on `threads` host threads. extended: every
    step in long double (the iterate is rounded to double between steps)."""
    from concurrent.futures import ThreadPoolExecutor
    from xca_amd.diag import all_connected_pairings
    pool = ThreadPoolExecutor(max_workers=threads) if threads > 1 else None
    orc = OracleEvaluator(prob, complex_arith=complex_arith, extended=extended)
    grid = prob.grid
    # complex_arith keeps complex blocks (hermitisation with the conjugate transpose, Z = -Re Tr G(beta),
    # block_sparse_solver.py:482-487, 586); otherwise the imaginary parts are dropped
    part = (lambda x: x) if complex_arith else (lambda x: x.real)
    G = [np.array(g, dtype=complex if complex_arith else float) for g in prob.G0]
    eta = 0.0
    topos = [(np.array(t, dtype=np.int32), (-1) ** o * s * (-1) ** (o + 1)) for o in range(1, max_order + 1)
             for s, t in all_connected_pairings(o)]
    diff = 0.0
    it = 0
    for it in range(1, maxiter + 1):
        S = None
        for t, fac in topos:
            s = [part(x) for x in orc.sigma(G, t)]
            S = [fac * x for x in s] if S is None else [a + fac * x for a, x in zip(S, s)]
        def solve(a):
            g = part(dyson(prob.beta, grid.it, grid.rf, a[0], a[1], eta, extended=extended))
            return 0.5 * (g + g.transpose(0, 2, 1).conj())
        pairs = list(zip(prob.G0, S))
        Gn = list(pool.map(solve, pairs)) if pool else [solve(a) for a in pairs]
        tr = 0.0
        for g in Gn:
            ev = interp(grid.it, grid.rf, g, np.array([1.0]), extended=extended)[0].real
            tr += np.trace(ev)
        Z = -tr
        deta = np.log(abs(Z)) / prob.beta
        Gn = [g * np.exp(-grid.tau * deta)[:, None, None] for g in Gn]
        eta += deta
        diff = max(np.abs(a - b).max() for a, b in zip(Gn, G))
        if diff < tol:
            break
        G = [mix * a + (1 - mix) * b for a, b in zip(Gn, G)]
    if pool:
        pool.shutdown()
    return G, eta, it, diff


# ------------------------------------------------------------------ threaded evaluation

def default_threads():
    import os
    return max(1, min(16, len(os.sched_getaffinity(0))))


def split_runs(fix, runs):
    """The contiguous runs sigma_threads evaluates `fix` in."""
    return [x for x in np.array_split(np.ascontiguousarray(fix, dtype=np.int32), min(runs, max(1, len(fix)))) if len(x)]


def sigma_threads(prob, G, topology, fix, threads=None, extended=False, complex_arith=False, runs=None, dropped=False):
    """compute_self_energy(G, topology, run) over contiguous runs of `fix` (`runs` of them, default one
    per thread), summed block by block in run order: the reference's scatter of flat indices over MPI
    ranks followed by the all-reduce (block_sparse_solver.py:629-640). Each run is one oracle call, so
    the reference's per-call zero-block flags apply per run. One evaluator per thread (the oracle
    releases the GIL). Returns real blocks unless complex_arith; with dropped, also the per-block sum
    of the discarded running sums over all runs (OracleEvaluator.last_dropped)."""
    from concurrent.futures import ThreadPoolExecutor
    import threading
    threads = threads or default_threads()
    parts_in = split_runs(fix, runs or threads)
    t = np.ascontiguousarray(topology, dtype=np.int32)
    local = threading.local()

    def one(run):
        if not hasattr(local, "orc"):
            local.orc = OracleEvaluator(prob, complex_arith=complex_arith, extended=extended)
        return local.orc.sigma(G, t, run), local.orc.last_dropped()

    with ThreadPoolExecutor(max_workers=min(threads, len(parts_in))) as ex:
        res = list(ex.map(one, parts_in))
    parts = [x[0] for x in res]
    drop = sum(x[1] for x in res)
    out = [sum(p[b] for p in parts) for b in range(len(parts[0]))]
    out = out if complex_arith else [x.real for x in out]
    return (out, drop) if dropped else out
