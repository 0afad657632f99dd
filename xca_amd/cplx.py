"""Complex128 problems on the real FP64 device path, by the real embedding of complex matrices.

The reference evaluates in complex128 (c++/triqs_xca/block_sparse_backbone.hpp:145-162). Every
quantity it multiplies is a matrix, with two kinds of exception: a hybridization coefficient, which
enters either through the barred operators (block_sparse.cpp:556-576, sums of coefficient times
operator) or as the scalar value Delta(t) on the zero-vertex line (block_sparse_backbone.cpp:204-225),
and the real DLR kernel values. The map

    e(X + iY) = [[X, -Y], [Y, X]]

is an injective algebra homomorphism from complex (m, k) matrices to real (2m, 2k) ones that commutes
with every real-linear operation on time arguments (DLR transforms, convolutions, kernel tables), and
e(M^dagger) = e(M)^T. So a complex problem is solved exactly by a real one in which:

- every block of G, Sigma and G0 is replaced by its embedding (block dimensions double);
- flavour o keeps its operators, embedded (e(c_o), e(c^dagger_o)), and gains a copy o + n whose
  operators carry a factor i (e(i c_o), e(i c^dagger_o)); the copy has the same block connections, so it
  is in the same symmetry set;
- the coefficients become real: C'[a, b] = Re C[a, b], C'[a + n, b] = Im C[a, b], C'[., b + n] = 0.
  A hybridization line of flavours (a, b) weighs the operator of its first flavour and the operator
  of its second flavour with C[a, b]; with the copies, the sum over (a, s) of C'[(a, s), b] times the
  line with i^s on the first operator is e((Re C + i Im C) * line), because e(i) commutes with every
  embedded matrix.

Traces do not survive the embedding as such: tr e(M) = 2 Re tr M. The single-particle Green's function
and the one-time correlators are traces; their imaginary parts are recovered from the copies with the
factor i (tr e(iM) = -2 Im tr M), and the partition function of the pseudo-particle loop is taken on
the complex blocks on the host.
"""
import numpy as np

from .atom import AtomDiagArrays


def embed(M):
    """(..., m, k) complex -> (..., 2m, 2k) real [[Re, -Im], [Im, Re]]."""
    M = np.asarray(M)
    X, Y = M.real, M.imag if np.iscomplexobj(M) else np.zeros(M.shape)
    top = np.concatenate([X, -Y], axis=-1)
    bot = np.concatenate([Y, X], axis=-1)
    return np.ascontiguousarray(np.concatenate([top, bot], axis=-2), dtype=np.float64)


def extract(E):
    """Inverse of embed on its image: (..., 2m, 2k) real -> (..., m, k) complex (left column blocks)."""
    E = np.asarray(E)
    m, k = E.shape[-2] // 2, E.shape[-1] // 2
    return E[..., :m, :k] + 1j * E[..., m:, :k]


def embedded_atom(ad):
    """The real block-sparse operators of the embedded problem: 2n flavours (copy o + n carries a factor
    i), block dimensions doubled, connections repeated for the copies."""
    n = ad.fs.n
    cb, db = {}, {}
    for src, dst in ((ad.c_blocks, cb), (ad.cdag_blocks, db)):
        for (o, j), M in src.items():
            M = np.asarray(M)
            dst[(o, j)] = embed(M)
            dst[(o + n, j)] = embed(1j * M)
    cc = np.concatenate([np.asarray(ad.c_conn), np.asarray(ad.c_conn)], axis=0)
    dc = np.concatenate([np.asarray(ad.cdag_conn), np.asarray(ad.cdag_conn)], axis=0)
    return AtomDiagArrays(2 * n, 2 * np.asarray(ad.dims), cc, dc, cb, db)


def embedded_coeffs(coeffs):
    """(p, n, n) complex -> (p, 2n, 2n) real: Re C in rows < n, Im C in rows >= n, columns < n."""
    C = np.asarray(coeffs)
    p, n = C.shape[0], C.shape[1]
    out = np.zeros((p, 2 * n, 2 * n))
    out[:, :n, :n] = C.real
    if np.iscomplexobj(C):
        out[:, n:, :n] = C.imag
    return out


def embedded_ops(conn, mats, dims):
    """Block operators (conn (nops, nblocks), packed matrices in (operator, source block) order) ->
    (conn', packed') with 2 nops operators: the embeddings, then the embeddings with a factor i."""
    conn = np.asarray(conn, dtype=np.int32)
    mats = np.asarray(mats).ravel()
    nops, nb = conn.shape
    blocks, off = [], 0
    for o in range(nops):
        row = []
        for j in range(nb):
            t = conn[o, j]
            if t < 0:
                row.append(None)
                continue
            sz = int(dims[t]) * int(dims[j])
            row.append(mats[off:off + sz].reshape(dims[t], dims[j]))
            off += sz
        blocks.append(row)
    if off != mats.size:
        raise ValueError(f"operator matrices: {mats.size} values given, the connections need {off}")
    out = []
    for phase in (1.0, 1j):
        for row in blocks:
            out += [embed(phase * M).ravel() for M in row if M is not None]
    packed = np.concatenate(out) if out else np.zeros(0)
    return np.concatenate([conn, conn], axis=0), packed


def trace_from_copies(out, nleft):
    """Complex traces from the embedded ones: out (..., 2 nleft, k) holds 2 Re X in rows < nleft and
    2 Re(i X) = -2 Im X in rows >= nleft."""
    out = np.asarray(out)
    return 0.5 * (out[..., :nleft, :] - 1j * out[..., nleft:, :])


def is_complex(*arrays):
    """True if any array (or list of arrays) has a nonzero imaginary part."""
    for a in arrays:
        items = a if isinstance(a, (list, tuple)) else [a]
        for x in items:
            x = np.asarray(x)
            if np.iscomplexobj(x) and np.any(x.imag != 0):
                return True
    return False
