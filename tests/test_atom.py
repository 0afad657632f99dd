"""Fock-space input builder against the reference's hard-coded matrices
(test/c++/test_atom_diag_utils.cpp:104-192)."""
import numpy as np

from xca_amd.atom import AtomDiag, FockSpace


def expected_fock_ops(nflav):
    N = 1 << nflav
    c = np.zeros((nflav, N, N))
    for o in range(nflav):
        for s in range(N):
            if s & (1 << o):
                sign = 1 if bin(s & ((1 << o) - 1)).count("1") % 2 == 0 else -1
                c[o, s ^ (1 << o), s] = sign
    return c, c.transpose(0, 2, 1)


def two_orbital_H(fs, norb=2, mu=0.25, U=1.0, V=0.1):
    do = lambda i: i
    up = lambda i: norb + i
    H = 0
    for i in range(norb):
        H = H + U * (fs.nop(up(i)) @ fs.nop(do(i))) + mu * (fs.nop(up(i)) + fs.nop(do(i)))
        H = H + V * (fs.cdag(up(i)) @ fs.c(do(i)) + fs.cdag(do(i)) @ fs.c(up(i)))
    return H


def test_fock_operators_and_hamiltonian():
    fs = FockSpace(4)
    c, cd = expected_fock_ops(4)
    for o in range(4):
        np.testing.assert_array_equal(fs.c(o).toarray(), c[o])
        np.testing.assert_array_equal(fs.cdag(o).toarray(), cd[o])
    H = two_orbital_H(fs).toarray()
    exp = np.zeros((16, 16))
    for (i, j), v in {(1, 1): .25, (1, 4): .1, (2, 2): .25, (2, 8): .1, (3, 3): .5, (3, 6): -.1, (3, 9): .1, (4, 1): .1,
                      (4, 4): .25, (5, 5): 1.5, (6, 3): -.1, (6, 6): .5, (6, 12): -.1, (7, 7): 1.75, (7, 13): -.1, (8, 2): .1,
                      (8, 8): .25, (9, 3): .1, (9, 9): .5, (9, 12): .1, (10, 10): 1.5, (11, 11): 1.75, (11, 14): -.1,
                      (12, 6): -.1, (12, 9): .1, (12, 12): .5, (13, 7): -.1, (13, 13): 1.75, (14, 11): -.1, (14, 14): 1.75,
                      (15, 15): 3.0}.items():
        exp[i, j] = v
    assert np.abs(H - exp).max() < 1e-13


def test_block_operators_consistent_with_dense():
    fs = FockSpace(4)
    H = two_orbital_H(fs)
    ad = AtomDiag(fs, H, qn="N")
    assert list(ad.dims) == [1, 4, 6, 4, 1]
    for o in range(4):
        dense = fs.c(o).toarray()
        for j in range(ad.nblocks):
            t = ad.c_conn[o, j]
            if t < 0:
                assert not dense[:, ad.fock_states[j]].any()
                continue
            np.testing.assert_array_equal(ad.c_blocks[(o, j)], dense[np.ix_(ad.fock_states[t], ad.fock_states[j])])
    np.testing.assert_allclose(ad.dense_H(), H.toarray(), atol=0)


def test_serving_request_stream_shape():
    """The serving stream: 1-3 orbitals, Lambda = 20 beta, beta in [10, 100], about 70 % NCA, and the
    same stream for the same seed."""
    from xca_amd import models
    a = models.serving_requests(40, seed=7)
    b = models.serving_requests(40, seed=7)
    assert [P.name for P, _ in a] == [P.name for P, _ in b]
    for P, order in a:
        assert P.n in (2, 4, 6) and order in (1, 2)
        assert 10.0 <= P.beta <= 100.0 and P.grid.lam == 20.0 * P.beta
    frac_nca = np.mean([o == 1 for _, o in a])
    assert 0.5 <= frac_nca <= 0.9
