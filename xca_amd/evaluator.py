"""Python facade of the device evaluator with the reference's method names and semantics.

Mirrors triqs_xca.block_sparse.DiagramEvaluator (python bindings block_sparse.wrap.cxx:173-208, C++
c++/triqs_xca/block_sparse_backbone.hpp:23-165): compute_self_energy(G, topology[, f_ix]),
get_num_self_energy_backbones(topology), reset(). G and the result are lists of (r, b, b) blocks (the
data of the reference's BlockGf). The device computes in real FP64. A problem with complex
coefficients, operators or propagators (complex_arith, detected from the inputs unless given) runs
on the device as its real embedding (xca_amd.cplx: blocks of twice the dimension, a second copy of
every flavour carrying the factor i), and its results are complex128; in real mode an input with a
nonzero imaginary part raises.
"""
import ctypes as C

import numpy as np

from . import cplx
from ._lib import XcaError, XcaProblem, as_f64, as_i32, check, dptr, i64_p, iptr, last_error, lib
from .diag import all_connected_pairings
from .dlr import kernel


def _real(a, what):
    a = np.asarray(a)
    if np.iscomplexobj(a):
        if np.any(a.imag != 0):
            raise XcaError(f"{what}: the device path is real FP64; imaginary parts must be zero")
        a = a.real
    return np.ascontiguousarray(a, dtype=np.float64)


def problem_struct(hyb_poles, hyb_coeffs, grid, ad, keep):
    """xca_problem for the ABI; every array it points to is appended to `keep` (poles first)."""
    n = ad.fs.n
    poles = keep(as_f64(hyb_poles).ravel())
    coeffs = keep(_real(hyb_coeffs, "hyb_coeffs"))
    # the library checks the extents against (p, n, n) before it reads the array
    shape = keep(np.array((list(coeffs.shape) + [0, 0, 0])[:3] if coeffs.ndim <= 3 else [-1, -1, -1], dtype=np.int64))
    cm, dm = ad.packed_ops()
    return XcaProblem(grid.beta, grid.r, dptr(keep(as_f64(grid.it))), dptr(keep(as_f64(grid.rf))), len(poles), dptr(poles), dptr(coeffs),
                      n, ad.nblocks, iptr(keep(as_i32(ad.dims))), iptr(keep(as_i32(ad.c_conn))), iptr(keep(as_i32(ad.cdag_conn))),
                      dptr(keep(_real(cm, "c"))), dptr(keep(_real(dm, "cdag"))), shape.ctypes.data_as(i64_p))


class DiagramEvaluator:
    """Pseudo-particle self-energy evaluator on one GPU.

    Parameters mirror DiagramEvaluator(hyb_poles, hyb_coeffs, tau_mesh, atom_diag): poles in energy
    units (p,), coefficients (p, n, n), a DLR grid (xca_amd.dlr.DLRGrid) and the block-sparse operators
    of an xca_amd.atom.AtomDiag.
    """

    def __init__(self, hyb_poles, hyb_coeffs, grid, ad, device=0, complex_arith=None):
        self.grid = grid
        self.ad = ad
        self.beta = grid.beta
        self.r = grid.r
        self.n = ad.fs.n
        self.block_dims = np.asarray(ad.dims, dtype=np.int32)
        if complex_arith is None:
            complex_arith = cplx.is_complex(hyb_coeffs, *ad.packed_ops())
        self.complex_arith = bool(complex_arith)
        if self.complex_arith:
            ad = cplx.embedded_atom(ad)
            hyb_coeffs = cplx.embedded_coeffs(hyb_coeffs)
        self.dims = np.asarray(ad.dims, dtype=np.int32)   # block dimensions on the device
        self._keep = []

        def keep(a):
            self._keep.append(a)
            return a

        d = problem_struct(hyb_poles, hyb_coeffs, grid, ad, keep)
        poles = self._keep[0]
        h = lib().xca_ctx_create(C.byref(d), device)
        if not h:
            raise XcaError(last_error())
        self._h = h
        self.p = len(poles)
        self.q = lib().xca_num_sym_sets(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().xca_ctx_destroy(h)
            self._h = None

    # ---------------------------------------------------------------- helpers
    def pack(self, G):
        if len(G) != len(self.block_dims):
            raise XcaError(f"expected {len(self.block_dims)} blocks, got {len(G)}")
        out = []
        for b, g in zip(self.block_dims, G):
            g = np.asarray(g) if self.complex_arith else _real(g, "G")
            if g.shape != (self.r, b, b):
                raise XcaError(f"G block has shape {g.shape}, expected {(self.r, b, b)}")
            out.append((cplx.embed(g) if self.complex_arith else g).ravel())
        return np.ascontiguousarray(np.concatenate(out))

    def unpack(self, flat):
        out, off = [], 0
        for b in self.dims:
            sz = self.r * b * b
            blk = flat[off:off + sz].reshape(self.r, b, b).copy()
            out.append(cplx.extract(blk) if self.complex_arith else blk)
            off += sz
        return out

    def _real_only(self, what):
        if self.complex_arith:
            raise XcaError(f"{what} is available for real problems; this evaluator runs a complex problem as its real embedding")

    # ---------------------------------------------------------------- reference API
    def get_num_self_energy_backbones(self, topology):
        self._real_only("get_num_self_energy_backbones")
        return int(lib().xca_num_backbones(self._h, len(topology)))

    def compute_self_energy(self, G, topology, f_ix=None):
        g = self.pack(G)
        topo = as_i32(topology)
        out = np.zeros_like(g)
        if f_ix is not None:
            self._real_only("compute_self_energy with flat indices (f_ix)")
        if f_ix is None:
            st = lib().xca_sigma(self._h, dptr(g), topo.shape[0], iptr(topo), None, 0, dptr(out))
        else:
            fx = as_i32(np.atleast_1d(f_ix))
            st = lib().xca_sigma(self._h, dptr(g), topo.shape[0], iptr(topo), fx.ctypes.data_as(C.POINTER(C.c_int32)), len(fx), dptr(out))
        check(st)
        return self.unpack(out)

    def reset(self):
        check(lib().xca_sigma_zero(self._h))

    # ---------------------------------------------------------------- correlators
    def get_num_single_ptcle_gf_backbones(self, topology):
        """2^(m-1) n^(m-1) p^(m-1) (block_sparse_backbone.cpp:732-735)."""
        self._real_only("get_num_single_ptcle_gf_backbones")
        return int(lib().xca_num_correlator_backbones(self._h, len(topology)))

    def _fix(self, f_ix):
        if f_ix is None:
            return None, 0
        fx = as_i32(np.atleast_1d(f_ix))
        self._keep_fix = fx
        return fx.ctypes.data_as(C.POINTER(C.c_int32)), len(fx)

    def compute_single_ptcle_gf(self, G, topology, f_ix=None):
        """compute_single_ptcle_gf(G, topology[, f_ix | f_ix_vec]) (block_sparse_backbone.cpp:779-825):
        (r, n, n) real, indexed by orbital, without the order/topology sign the solver applies.
        G None uses the G already on the device (set_G or the previous call)."""
        g = None if G is None else dptr(self.pack(G))
        topo = as_i32(topology)
        if f_ix is not None:
            self._real_only("compute_single_ptcle_gf with flat indices (f_ix)")
        fx, nf = self._fix(f_ix)
        m = 2 * self.n if self.complex_arith else self.n
        out = np.zeros((self.r, m, m))
        check(lib().xca_spgf(self._h, g, topo.shape[0], iptr(topo), fx, nf, dptr(out)))
        if self.complex_arith:  # traces of the embedding: 2 Re X, and -2 Im X from the copies with i
            return cplx.trace_from_copies(out, self.n)[:, :, :self.n]
        return out

    def compute_one_time_correlator(self, G, ops_tau, ops_0, topology, f_ix=None):
        """compute_one_time_correlator(G, ops_tau, ops_0, topology, f_ix_vec)
        (block_sparse_backbone.cpp:850-876) for block operators given as (conn (nops, nblocks),
        packed matrices in (operator, source block) order): (r, len(ops_tau), len(ops_0)) real."""
        g = None if G is None else dptr(self.pack(G))
        topo = as_i32(topology)
        if f_ix is not None:
            self._real_only("compute_one_time_correlator with flat indices (f_ix)")
        fx, nf = self._fix(f_ix)
        mc, kc = as_i32(ops_tau[0]), as_i32(ops_0[0])
        if mc.ndim != 2 or kc.ndim != 2 or mc.shape[1] != len(self.dims) or kc.shape[1] != len(self.dims):
            raise XcaError("operator connections must have shape (nops, nblocks)")
        nmu = mc.shape[0]
        if self.complex_arith:
            try:
                mc, mm = cplx.embedded_ops(mc, ops_tau[1], self.block_dims)
                kc, km = cplx.embedded_ops(kc, ops_0[1], self.block_dims)
            except ValueError as e:
                raise XcaError(str(e)) from None
            kc, km = kc[:kc.shape[0] // 2], km[:km.size // 2]   # the factor i on one side suffices
            mc, mm, kc, km = as_i32(mc), as_f64(mm), as_i32(kc), as_f64(km)
        else:
            mm, km = _real(ops_tau[1], "ops_tau"), _real(ops_0[1], "ops_0")
        self._keep_ops = (mc, mm, kc, km)
        out = np.zeros((self.r, mc.shape[0], kc.shape[0]))
        check(lib().xca_correlator(self._h, g, topo.shape[0], iptr(topo), fx, nf, mc.shape[0], iptr(mc), dptr(mm), kc.shape[0],
                                   iptr(kc), dptr(km), dptr(out)))
        return cplx.trace_from_copies(out, nmu) if self.complex_arith else out

    def single_particle_greens_function(self, G, max_order):
        """eval_single_particle_greens_function (block_sparse_solver.py:657-677): the sum over orders
        1..max_order and connected topologies of (-1)^order * s_topo * compute_single_ptcle_gf."""
        tot = np.zeros((self.r, self.n, self.n), dtype=np.complex128 if self.complex_arith else np.float64)
        self.set_G(G)
        for order in range(1, max_order + 1):
            for sign, topo in all_connected_pairings(order):
                tot += (-1) ** order * sign * self.compute_single_ptcle_gf(None, np.array(topo, dtype=np.int32))
        return tot

    def one_time_correlator(self, G, max_order, ops_tau, ops_0):
        """eval_one_time_correlator (block_sparse_solver.py:700-737): <T O(tau) O(0)> summed over
        orders and topologies with (-1)^order * s_topo."""
        tot = None
        self.set_G(G)
        for order in range(1, max_order + 1):
            for sign, topo in all_connected_pairings(order):
                c = (-1) ** order * sign * self.compute_one_time_correlator(None, ops_tau, ops_0, np.array(topo, dtype=np.int32))
                tot = c if tot is None else tot + c
        return tot

    # ---------------------------------------------------------------- device-resident form
    def set_G(self, G):
        g = self.pack(G)
        check(lib().xca_set_G(self._h, dptr(g)))

    def accumulate(self, topology, scale):
        topo = as_i32(topology)
        check(lib().xca_sigma_accumulate(self._h, topo.shape[0], iptr(topo), float(scale)))

    def get_sigma(self):
        out = np.zeros(sum(self.r * int(b) * int(b) for b in self.dims))
        check(lib().xca_get_sigma(self._h, dptr(out)))
        return self.unpack(out)

    def synchronize(self):
        check(lib().xca_synchronize(self._h))

    def last_stats(self):
        items = C.c_int64(0)
        macs = C.c_double(0)
        ms = C.c_double(0)
        check(lib().xca_last_stats(self._h, C.byref(items), C.byref(macs), C.byref(ms)))
        return items.value, macs.value, ms.value

    def set_tile_variant(self, variant):
        """Force the chain kernel's tile width (0, 1, 2 = 128, 64, 32 columns) or -1 for the autotuner."""
        check(lib().xca_set_tile_variant(self._h, int(variant)))

    def set_launch_mode(self, mode):
        """0: one persistent chain launch per plan (default); 1: one chain launch per chunk, which with
        an explicit flat-index subset is one launch per (path, pole tuple) item. Same results."""
        check(lib().xca_set_launch_mode(self._h, int(mode)))

    def last_tile(self):
        v = C.c_int(0)
        cols = C.c_int(0)
        check(lib().xca_last_tile(self._h, C.byref(v), C.byref(cols)))
        return v.value, cols.value

    def sigma_order(self, max_order, timed=False):
        """Device Sigma <- Sigma of the uploaded G over orders 1..max_order with the solver's signs, in
        one graph-replayed launch sequence (asynchronous; get_sigma or synchronize waits)."""
        check(lib().xca_sigma_order(self._h, int(max_order), 1 if timed else 0))

    def pseudo_particle_self_energy(self, G, max_order):
        """Sigma summed over orders 1..max_order and connected topologies with the solver's signs
        (block_sparse_solver.py:597-638: net factor -s_topo per topology)."""
        self.set_G(G)
        self.reset()
        for order in range(1, max_order + 1):
            for sign, topo in all_connected_pairings(order):
                self.accumulate(np.array(topo, dtype=np.int32), (-1) ** order * sign * (-1) ** (order + 1))
        return self.get_sigma()

    # ---------------------------------------------------------------- Dyson and self-consistency
    def dyson_init(self, G0):
        """One dyson_it_ppsc per block built from the atomic propagator G0 (list of blocks)."""
        self._G0 = self.pack(G0)
        check(lib().xca_dyson_init(self._h, dptr(self._G0)))

    def dyson_solve(self, Sigma, eta):
        s = self.pack(Sigma)
        out = np.zeros_like(s)
        check(lib().xca_dyson_solve(self._h, dptr(s), float(eta), dptr(out)))
        return self.unpack(out)

    def ppsc_solve(self, max_order, tol=1e-10, maxiter=100, mix=1.0, eta0=0.0):
        """Self-consistent pseudo-particle G; returns (G blocks, eta, iterations, last max|G_new - G|).
        The loop runs on the device (xca_ppsc_solve): one launch sequence per iteration, or, for
        blocks above 12 rows, Sigma on the large-block engine and the device-wide Dyson path around
        the same update launches. Complex problems run the loop on the host around the device's
        Sigma and Dyson solve."""
        if self.complex_arith:
            return self._ppsc_solve_host(max_order, tol, maxiter, mix, eta0)
        out = np.zeros(sum(self.r * int(b) * int(b) for b in self.dims))
        eta = C.c_double(0)
        it = C.c_int(0)
        diff = C.c_double(0)
        check(lib().xca_ppsc_solve(self._h, float(eta0), int(max_order), float(tol), int(maxiter), float(mix), dptr(out), C.byref(eta),
                                   C.byref(it), C.byref(diff)))
        return self.unpack(out), eta.value, it.value, diff.value

    def _ppsc_solve_host(self, max_order, tol, maxiter, mix, eta0):
        """The loop of BlockSparseSolver.solve with classic normalisation (block_sparse_solver.py:
        289-346, 427-449, 482-487) driven from the host: Sigma and the Dyson solve (with its
        hermitisation) on the device (for a complex problem in the embedding); Z = -Re Tr G_new(beta),
        the energy shift, the convergence test on max |G_new - G| and the mixing on the host
        (O(r sum b^2) work per iteration)."""
        if not hasattr(self, "_G0"):
            raise XcaError("ppsc_solve: dyson_init(G0) first")
        grid = self.grid
        kb = kernel(np.ones(1), grid.rf[None, :])[0]   # K(tau = beta, w_k)
        G = self.unpack(self._G0)
        eta = float(eta0)
        hist = []
        it, diff = 0, np.inf
        for it in range(1, maxiter + 1):
            G_new = self.dyson_solve(self.pseudo_particle_self_energy(G, max_order), eta)
            Z = float(-sum(np.trace(np.tensordot(kb, grid.vals2coefs(g), axes=(0, 0))) for g in G_new).real)
            deta = np.log(abs(Z)) / self.beta
            G_new = [g * np.exp(-grid.tau * deta)[:, None, None] for g in G_new]
            eta += deta
            diff = max(float(np.abs(a - b).max()) for a, b in zip(G_new, G))
            hist.append((diff, Z, eta))
            if diff < tol:
                break
            G = [mix * a + (1.0 - mix) * b for a, b in zip(G_new, G)]
        self._history = hist
        return G, eta, it, diff

    def ppsc_history(self):
        """Per-iteration report of the last ppsc_solve: [(diff_G, Z before the normalisation, eta)]."""
        if self.complex_arith:
            return list(getattr(self, "_history", []))
        n = check(lib().xca_ppsc_history(self._h, None, 0))
        h = np.zeros(3 * max(n, 1))
        check(lib().xca_ppsc_history(self._h, dptr(h), n))
        return [tuple(h[3 * k:3 * k + 3]) for k in range(n)]
