"""Fock-space input builder: restates what triqs.atom_diag supplies to the reference's
DiagramEvaluator (c++/triqs_xca/atom_diag_utils.cpp:39-343) and the atomic pseudo-particle
propagator of python/triqs_xca/block_sparse_solver.py:1305-1328.

Conventions:
  * flavour o is bit o of the Fock-state index; c_o|s> = (-1)^{#occupied flavours below o} |s - o>
    (the sign convention pinned by test/c++/test_atom_diag_utils.cpp:104-121);
  * blocks are the sectors of the chosen conserved quantities ('NSz': particle number and S_z,
    'N': particle number, 'none': one block), ordered by their first Fock state; states inside a
    block are in increasing Fock order;
  * operator and propagator blocks are in the Fock basis of each block, as get_operators_impl
    (atom_diag_utils.cpp:133-148) and ad_to_atom_prop build them.
"""
import itertools

import numpy as np
import scipy.sparse as sp


class FockSpace:
    def __init__(self, n):
        self.n = n
        self.dim = 1 << n
        s = np.arange(self.dim)
        self._c = []
        for o in range(n):
            occ = (s >> o) & 1
            src = s[occ == 1]
            dst = src ^ (1 << o)
            below = src & ((1 << o) - 1)
            sign = np.where(np.array([bin(x).count("1") for x in below]) % 2 == 0, 1.0, -1.0)
            self._c.append(sp.csr_matrix((sign, (dst, src)), shape=(self.dim, self.dim)))

    def c(self, o):
        return self._c[o]

    def action(self, o, dagger):
        """c_o (or c^dagger_o) as a map of basis states: (target state or -1, sign) per source state;
        every column of these operators holds at most one nonzero."""
        s = np.arange(self.dim)
        occ = (s >> o) & 1
        ok = occ == (0 if dagger else 1)
        below = s & ((1 << o) - 1)
        par = np.array([bin(x).count("1") & 1 for x in below])
        return np.where(ok, s ^ (1 << o), -1), np.where(par == 0, 1.0, -1.0)

    def cdag(self, o):
        return self._c[o].T.tocsr()

    def nop(self, o):
        return (self.cdag(o) @ self.c(o)).tocsr()


def auto_partition(fs, H):
    """Finest partition of the Fock space into blocks that H leaves invariant and that every c_o and
    c^dagger_o maps, block by block, into single blocks: the automatic partition of TRIQS's atom_diag
    (used where the reference builds its AtomDiag, python/triqs_xca/block_sparse_solver.py), after
    Seth et al., Comput. Phys. Commun. 200, 274 (2016), appendix on the automatic partition.
    1. Blocks = connected components of the graph of the nonzero elements of H.
    2. Repeat until nothing changes: for every operator and every block, merge the blocks that the
       block's states are mapped into.
    Returns a list of blocks (sorted Fock-state arrays), ordered by their smallest state."""
    dim = fs.dim
    parent = np.arange(dim)

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    def union(a, b):
        ra, rb = find(a), find(b)
        if ra == rb:
            return False
        if ra < rb:
            parent[rb] = ra
        else:
            parent[ra] = rb
        return True

    Hc = sp.coo_matrix(H)
    for i, j, v in zip(Hc.row, Hc.col, Hc.data):
        if v != 0.0:
            union(int(i), int(j))
    ops = []
    for o in range(fs.n):
        for M in (fs.c(o), fs.cdag(o)):
            Mc = M.tocoo()
            ops.append((Mc.col.astype(np.int64), Mc.row.astype(np.int64)))  # (source, target) pairs
    changed = True
    while changed:
        changed = False
        for src, dst in ops:
            first_target = {}  # source block -> a target state
            for a, t in zip(src, dst):
                ra = find(int(a))
                if ra in first_target:
                    changed |= union(first_target[ra], int(t))
                else:
                    first_target[ra] = int(t)
    roots = np.array([find(x) for x in range(dim)])
    blocks = {}
    for x in range(dim):
        blocks.setdefault(int(roots[x]), []).append(x)
    return [np.array(v, dtype=np.int64) for _, v in sorted(blocks.items(), key=lambda kv: min(kv[1]))]


class AtomDiag:
    """Block-diagonalised local Hamiltonian with block-sparse c and c^dagger.

    H: (dim, dim) Hermitian matrix, real or complex (dense or sparse), in the Fock basis. qn: "NSz" (spin[o] = +1/-1 gives
    S_z), "N", "none" (one block), or "auto" (the finest partition compatible with H and the
    operators, auto_partition).
    """

    def __init__(self, fs, H, qn="NSz", spin=None):
        self.fs = fs
        n, dim = fs.n, fs.dim
        H = sp.csr_matrix(H)
        states = np.arange(dim)
        N = np.array([bin(s).count("1") for s in states])
        if qn == "auto":
            keys = [0] * dim
            for b, blk in enumerate(auto_partition(fs, H)):
                for x in blk:
                    keys[x] = b
        elif qn == "NSz":
            spin = np.asarray(spin)
            sz = np.array([sum(spin[o] for o in range(n) if (s >> o) & 1) for s in states])
            keys = list(zip(N, sz))
        elif qn == "N":
            keys = list(N)
        elif qn == "none":
            keys = [0] * dim
        else:
            raise ValueError(qn)
        order = {}
        for s, k in zip(states, keys):
            order.setdefault(k, []).append(s)
        self.fock_states = [np.array(v) for v in order.values()]
        self.nblocks = len(self.fock_states)
        self.dims = np.array([len(f) for f in self.fock_states], dtype=np.int32)
        self.block_of = np.empty(dim, dtype=np.int64)
        self.pos_in_block = np.empty(dim, dtype=np.int64)
        for b, f in enumerate(self.fock_states):
            self.block_of[f] = b
            self.pos_in_block[f] = np.arange(len(f))
        Hc = H.tocoo()
        if np.any(self.block_of[Hc.row] != self.block_of[Hc.col]):
            raise ValueError("Hamiltonian couples different blocks of the chosen quantum numbers")
        Hd = H.toarray()
        self.H_blocks = [Hd[np.ix_(f, f)] for f in self.fock_states]
        self.eigvals, self.eigvecs = [], []
        for hb in self.H_blocks:
            E, U = np.linalg.eigh(hb)
            self.eigvals.append(E)
            self.eigvecs.append(U)
        self.gs_energy = min(E.min() for E in self.eigvals)
        self.energies = [E - self.gs_energy for E in self.eigvals]
        # block-sparse operators
        self.c_conn = -np.ones((n, self.nblocks), dtype=np.int32)
        self.cdag_conn = -np.ones((n, self.nblocks), dtype=np.int32)
        self.c_blocks = {}
        self.cdag_blocks = {}
        for o in range(n):
            for dagger, conn, store in ((False, self.c_conn, self.c_blocks), (True, self.cdag_conn, self.cdag_blocks)):
                dst, sign = fs.action(o, dagger)
                for j, fj in enumerate(self.fock_states):
                    cols = np.nonzero(dst[fj] >= 0)[0]
                    if len(cols) == 0:
                        continue
                    rows = dst[fj[cols]]
                    tgt = np.unique(self.block_of[rows])
                    if len(tgt) != 1:
                        raise ValueError("operator maps one block into several blocks")
                    t = int(tgt[0])
                    conn[o, j] = t
                    M = np.zeros((len(self.fock_states[t]), len(fj)))
                    M[self.pos_in_block[rows], cols] = sign[fj[cols]]
                    store[(o, j)] = M

    def packed_ops(self):
        """Concatenated c and c^dagger blocks in (o, j) order, as xca_problem expects."""
        c = [self.c_blocks[(o, j)].ravel() for o in range(self.fs.n) for j in range(self.nblocks) if self.c_conn[o, j] >= 0]
        d = [self.cdag_blocks[(o, j)].ravel() for o in range(self.fs.n) for j in range(self.nblocks) if self.cdag_conn[o, j] >= 0]
        return np.concatenate(c) if c else np.zeros(0), np.concatenate(d) if d else np.zeros(0)

    def atomic_G0(self, beta, tau):
        """G0_b(tau) = -U exp(-tau (E - eta0)) U^dagger and eta0 = -ln Z0 / beta with ground-state-shifted
        energies (block_sparse_solver.py:1305-1328). tau: absolute times (r,). Returns (blocks, eta0)."""
        Z0 = sum(np.exp(-beta * E).sum() for E in self.energies)
        eta0 = -np.log(Z0) / beta
        G = []
        for E, U in zip(self.energies, self.eigvecs):
            d = -np.exp(-tau[:, None] * (E - eta0)[None, :])
            G.append(np.einsum("ia,ta,ja->tij", U, d, U.conj()))
        return G, eta0

    def dense_H(self):
        dim = self.fs.dim
        Hd = np.zeros((dim, dim), dtype=np.result_type(*self.H_blocks))
        for f, hb in zip(self.fock_states, self.H_blocks):
            Hd[np.ix_(f, f)] = hb
        return Hd


class _Flavours:
    def __init__(self, n):
        self.n = n


class AtomDiagArrays:
    """The block-sparse operator data an evaluator needs, without a Hamiltonian: flavour count, block
    dimensions, c / c^dagger block connections (n, nblocks) and the Fock-basis blocks
    {(o, j): (dims[conn[o, j]], dims[j])}. It is the array form of what get_operators builds from an
    atom_diag (c++/triqs_xca/atom_diag_utils.cpp:39-185) and stands wherever an AtomDiag is accepted
    for evaluation."""

    def __init__(self, n, dims, c_conn, cdag_conn, c_blocks, cdag_blocks, fock_states=None):
        self.fs = _Flavours(int(n))
        self.dims = np.asarray(dims, dtype=np.int32)
        self.fock_states = None if fock_states is None else [np.asarray(f, dtype=np.int64) for f in fock_states]
        self.nblocks = len(self.dims)
        self.c_conn = np.asarray(c_conn, dtype=np.int32).reshape(n, self.nblocks)
        self.cdag_conn = np.asarray(cdag_conn, dtype=np.int32).reshape(n, self.nblocks)
        self.c_blocks = dict(c_blocks)
        self.cdag_blocks = dict(cdag_blocks)
        for conn, blocks, what in ((self.c_conn, self.c_blocks, "c"), (self.cdag_conn, self.cdag_blocks, "c^dagger")):
            for o in range(n):
                for j in range(self.nblocks):
                    t = conn[o, j]
                    if t < 0:
                        continue
                    if t >= self.nblocks:
                        raise ValueError(f"AtomDiagArrays: {what}_{o} connects block {j} to block {t}, beyond {self.nblocks} blocks")
                    shp = np.shape(blocks.get((o, j)))
                    if shp != (self.dims[t], self.dims[j]):
                        raise ValueError(f"AtomDiagArrays: {what}_{o} block from {j} has shape {shp}, expected {(self.dims[t], self.dims[j])}")

    packed_ops = AtomDiag.packed_ops

    def __reduce_to_dict__(self):
        d = {"n": self.fs.n, "dims": self.dims, "c_conn": self.c_conn, "cdag_conn": self.cdag_conn,
             "c_blocks": {f"{o},{j}": m for (o, j), m in self.c_blocks.items()},
             "cdag_blocks": {f"{o},{j}": m for (o, j), m in self.cdag_blocks.items()}}
        if self.fock_states is not None:
            d["fock_states"] = list(self.fock_states)
        return d

    @classmethod
    def __factory_from_dict__(cls, name, d):
        key = lambda k: tuple(int(x) for x in k.split(","))
        return cls(int(d["n"]), d["dims"], d["c_conn"], d["cdag_conn"], {key(k): np.asarray(m) for k, m in d["c_blocks"].items()},
                   {key(k): np.asarray(m) for k, m in d["cdag_blocks"].items()}, d.get("fock_states"))

    def __eq__(self, other):
        if not isinstance(other, AtomDiagArrays):
            return False
        same = lambda x, y: x.keys() == y.keys() and all(np.array_equal(x[k], y[k]) for k in x)
        fa, fb = self.fock_states, other.fock_states
        return (self.fs.n == other.fs.n and np.array_equal(self.dims, other.dims) and np.array_equal(self.c_conn, other.c_conn)
                and np.array_equal(self.cdag_conn, other.cdag_conn) and same(self.c_blocks, other.c_blocks)
                and same(self.cdag_blocks, other.cdag_blocks)
                and ((fa is None and fb is None) or (fa is not None and fb is not None and len(fa) == len(fb)
                                                    and all(np.array_equal(x, y) for x, y in zip(fa, fb)))))

    @classmethod
    def from_atom_diag(cls, ad):
        return cls(ad.fs.n, ad.dims, ad.c_conn, ad.cdag_conn, ad.c_blocks, ad.cdag_blocks, getattr(ad, "fock_states", None))

    @classmethod
    def from_triqs(cls, ad):
        """From a TRIQS atom_diag as get_operators_impl reads it (atom_diag_utils.cpp:39-148): the
        fundamental operators (`fops`, flavour o = bit o of a Fock state), the subspaces' Fock states
        (`fock_states`), and the connections `c_connection` / `cdag_connection` (n, n_subspaces).
        Blocks are the Fock-basis operator matrices restricted to (final, initial) subspace states."""
        n = len(ad.fops)
        fock = [np.asarray(f, dtype=np.int64) for f in ad.fock_states]
        nb = len(fock)
        cc = np.asarray(ad.c_connection, dtype=np.int64).reshape(n, nb)
        dc = np.asarray(ad.cdag_connection, dtype=np.int64).reshape(n, nb)
        fs = FockSpace(n)
        cb, db = {}, {}
        for o in range(n):
            C, Cd = fs.c(o).tocsr(), fs.cdag(o).tocsr()
            for j in range(nb):
                if cc[o, j] >= 0:
                    cb[(o, j)] = C[fock[cc[o, j]], :][:, fock[j]].toarray()
                if dc[o, j] >= 0:
                    db[(o, j)] = Cd[fock[dc[o, j]], :][:, fock[j]].toarray()
        return cls(n, [len(f) for f in fock], cc, dc, cb, db, fock)


# --------------------------------------------------------------------------- Hamiltonians

def kanamori_hamiltonian(fs, norb, U, J, mu):
    """Kanamori interaction with spin flip and pair hopping, U' = U - 2J, flavours ordered
    (do, 0..norb-1), (up, 0..norb-1) (test/c++/block_sparse_utils.cpp:189-208)."""
    do = lambda o: o
    up = lambda o: norb + o
    nn = lambda a: fs.nop(a)
    cd, c = fs.cdag, fs.c
    Up = U - 2 * J
    H = sp.csr_matrix((fs.dim, fs.dim))
    for o in range(norb):
        H = H - mu * (nn(up(o)) + nn(do(o)))
        H = H + U * (nn(up(o)) @ nn(do(o)))
    for o1 in range(norb):
        for o2 in range(norb):
            if o1 == o2:
                continue
            H = H + 0.5 * Up * (nn(up(o1)) @ nn(do(o2)) + nn(do(o1)) @ nn(up(o2)))
            H = H + 0.5 * (Up - J) * (nn(up(o1)) @ nn(up(o2)) + nn(do(o1)) @ nn(do(o2)))
            H = H - J * (cd(up(o1)) @ c(do(o1)) @ cd(do(o2)) @ c(up(o2)))
            H = H - J * (cd(up(o1)) @ cd(do(o1)) @ c(up(o2)) @ c(do(o2)))
    return H.tocsr()


def kanamori_spins(norb):
    return np.array([-1] * norb + [1] * norb)


def wigner_3j(j1, j2, j3, m1, m2, m3):
    """Wigner 3j symbol for integer arguments (Racah's formula)."""
    from math import factorial, sqrt
    if m1 + m2 + m3 != 0 or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    f = factorial
    tri = f(j1 + j2 - j3) * f(j1 - j2 + j3) * f(-j1 + j2 + j3) / f(j1 + j2 + j3 + 1)
    pre = sqrt(tri * f(j1 + m1) * f(j1 - m1) * f(j2 + m2) * f(j2 - m2) * f(j3 + m3) * f(j3 - m3))
    s = 0.0
    for k in range(0, j1 + j2 + j3 + 1):
        d = [k, j1 + j2 - j3 - k, j1 - m1 - k, j2 + m2 - k, j3 - j2 + m1 + k, j3 - j1 - m2 + k]
        if min(d) < 0:
            continue
        s += (-1) ** k / (f(d[0]) * f(d[1]) * f(d[2]) * f(d[3]) * f(d[4]) * f(d[5]))
    return (-1) ** (j1 - j2 - m3) * pre * s


def slater_u_matrix(l, F):
    """Slater-Condon U_{m1 m2 m3 m4} = sum_k a_k(m1, m2, m3, m4) F^k in the complex spherical-harmonic
    basis (m = -l..l), with a_k = (2l+1)^2 (l k l; 0 0 0)^2 sum_q (-1)^(m1+m2+q)
    (l k l; -m1 q m3)(l k l; -m2 -q m4), k = 0, 2, .., 2l (the standard Gaunt-coefficient form; real)."""
    n = 2 * l + 1
    ms = range(-l, l + 1)
    U = np.zeros((n, n, n, n))
    for ik, k in enumerate(range(0, 2 * l + 1, 2)):
        ck = (2 * l + 1) ** 2 * wigner_3j(l, k, l, 0, 0, 0) ** 2
        for a, m1 in enumerate(ms):
            for b, m2 in enumerate(ms):
                for c, m3 in enumerate(ms):
                    for d, m4 in enumerate(ms):
                        s = 0.0
                        for q in range(-k, k + 1):
                            s += (-1) ** (m1 + m2 + q) * wigner_3j(l, k, l, -m1, q, m3) * wigner_3j(l, k, l, -m2, -q, m4)
                        U[a, b, c, d] += ck * s * F[ik]
    return U


def slater_hamiltonian(fs, l, F, mu):
    """H = 1/2 sum U_{m1 m2 m3 m4} c^dag_{m1 s} c^dag_{m2 s'} c_{m4 s'} c_{m3 s} - mu N for one l shell,
    flavours ordered (do, m = -l..l), (up, m = -l..l) (h_int_slater with off-diagonal terms, as used at
    benchmarks/slater_condon/slater_condon_interaction.py:50-59)."""
    n = 2 * l + 1
    U = slater_u_matrix(l, F)
    cd, c = fs.cdag, fs.c
    H = sp.csr_matrix((fs.dim, fs.dim))
    for s1 in range(2):
        for s2 in range(2):
            for a in range(n):
                for b in range(n):
                    for cc in range(n):
                        for d in range(n):
                            u = U[a, b, cc, d]
                            if abs(u) < 1e-14:
                                continue
                            A, B, C, D = s1 * n + a, s2 * n + b, s1 * n + cc, s2 * n + d
                            if A == B or C == D:
                                continue
                            H = H + 0.5 * u * (cd(A) @ cd(B) @ c(D) @ c(C))
    N = sum(fs.nop(o) for o in range(2 * n))
    return (H - mu * N).tocsr()


# --------------------------------------------------------------------------- hybridization

def semicircle_delta_tau(tau_abs, beta, D=1.0, t2=0.25, nquad=4000):
    """Delta(tau) = t2 * int rho(e) K(tau/beta, beta e) de for the semicircular DOS of half-bandwidth
    D, rho(e) = 2 sqrt(D^2 - e^2)/(pi D^2); Gauss-Chebyshev (second kind) quadrature."""
    k = np.arange(1, nquad + 1)
    th = k * np.pi / (nquad + 1)
    x = np.cos(th)
    w = np.pi / (nquad + 1) * np.sin(th) ** 2
    e = D * x
    w = w * (2.0 / np.pi)
    t = np.asarray(tau_abs)[:, None] / beta
    om = beta * e[None, :]
    # K(t, om) for t in [0, 1], om any sign, without overflow
    Kv = np.where(om >= 0, -np.exp(-t * om) / (1 + np.exp(-np.abs(om))), -np.exp((1 - t) * om) / (1 + np.exp(om)))
    return t2 * (Kv * w[None, :]).sum(axis=1)


def dlr_hybridization(grid, delta_vals):
    """Poles (energy units) and DLR coefficients (p, n, n) of a hybridization given by its values
    (r, n, n) at the DLR nodes, without pole compression (block_sparse_solver.py:181-185)."""
    coeffs = grid.vals2coefs(delta_vals)
    poles = grid.rf / grid.beta
    return poles, coeffs
