"""Automatic partition of the Fock space (AtomDiag(qn="auto")): the finest partition that H leaves
invariant and that every c / c^dagger maps block by block into single blocks. Pinned against a
brute-force fixpoint on sets, the invariants that define it, refinement of (N, S_z), and the
oracle's Sigma, which must not depend on the partition once mapped back to the Fock basis."""
import numpy as np
import pytest

import pyoracle as po
from xca_amd import models
from xca_amd.atom import AtomDiag, FockSpace, auto_partition, kanamori_hamiltonian, kanamori_spins
from xca_amd.diag import all_connected_pairings


def _brute_force_partition(fs, H):
    """Singletons, then merge until the defining conditions hold (sets of frozensets, no union-find)."""
    dim = fs.dim
    Hd = np.asarray(H.todense()) if hasattr(H, "todense") else np.asarray(H)
    ops = [np.asarray(M.todense()) for o in range(fs.n) for M in (fs.c(o), fs.cdag(o))]
    blocks = [frozenset([x]) for x in range(dim)]

    def block_of(x):
        return next(b for b in blocks if x in b)

    changed = True
    while changed:
        changed = False
        pairs = [(i, j) for i in range(dim) for j in range(dim) if i != j and Hd[i, j] != 0.0]
        for M in ops:
            for b in blocks:
                targets = sorted({int(t) for s in b for t in np.nonzero(M[:, s])[0]})
                pairs += [(targets[0], t) for t in targets[1:]]
        for i, j in pairs:
            bi, bj = block_of(i), block_of(j)
            if bi != bj:
                blocks = [b for b in blocks if b not in (bi, bj)] + [bi | bj]
                changed = True
                break
    return sorted((sorted(b) for b in blocks), key=min)


def _kanamori_H(norb, U=4.0, J=0.5):
    fs = FockSpace(2 * norb)
    return fs, kanamori_hamiltonian(fs, norb, U, J, (3 * U - 5 * J) / 2)


@pytest.mark.parametrize("norb", [1, 2])
def test_auto_partition_equals_brute_force_fixpoint(norb):
    fs, H = _kanamori_H(norb)
    auto = [list(b) for b in auto_partition(fs, H)]
    assert auto == _brute_force_partition(fs, H)


@pytest.mark.parametrize("norb", [1, 2, 3])
def test_auto_partition_invariants_and_refines_n_sz(norb):
    fs, H = _kanamori_H(norb)
    blocks = auto_partition(fs, H)
    assert sorted(np.concatenate(blocks).tolist()) == list(range(fs.dim))
    lab = np.empty(fs.dim, dtype=int)
    for b, blk in enumerate(blocks):
        lab[blk] = b
    Hc = H.tocoo()
    assert np.all(lab[Hc.row[Hc.data != 0]] == lab[Hc.col[Hc.data != 0]])
    for o in range(fs.n):
        for M in (fs.c(o), fs.cdag(o)):
            Mc = M.tocoo()
            for b in range(len(blocks)):
                tg = set(lab[Mc.row[lab[Mc.col] == b]].tolist())
                assert len(tg) <= 1
    nsz = AtomDiag(fs, H, qn="NSz", spin=kanamori_spins(norb))
    for blk in blocks:  # every automatic block lies inside one (N, S_z) block
        assert len(set(nsz.block_of[blk].tolist())) == 1
    auto = AtomDiag(fs, H, qn="auto")
    assert auto.nblocks == len(blocks) and max(auto.dims) <= max(nsz.dims)


def _fock(prob, blocks):
    """Blocks (r, b, b) in the AtomDiag's block order -> (r, dim, dim) in the Fock basis."""
    ad = prob.ad
    r = prob.grid.r
    out = np.zeros((r, ad.fs.dim, ad.fs.dim))
    for f, blk in zip(ad.fock_states, blocks):
        out[:, f[:, None], f[None, :]] = blk
    return out


def test_oracle_sigma_does_not_depend_on_the_partition():
    """2-orbital Kanamori with a diagonal hybridization, orders 1-2: the Fock-basis Sigma of the
    (N, S_z) and of the automatic partition agree to rounding. The grid resolves the atomic
    energies (beta E_max ~ 40 < Lambda = 100); on an unresolved grid (beta E beyond Lambda) the two
    partitions' roundings already differ at 1e-6 relative, as (N, S_z) and N partitions do."""
    kw = dict(norb=2, U=4.0, J=0.5, mu=4.75, beta=2.0, lam=100.0, eps=1e-10)
    A = models.kanamori(**kw, qn="NSz")
    B = models.kanamori(**kw, qn="auto")
    assert B.ad.nblocks > A.ad.nblocks
    oa = po.OracleEvaluator(A, complex_arith=False)
    ob = po.OracleEvaluator(B, complex_arith=False)
    for order in (1, 2):
        for _, t in all_connected_pairings(order):
            topo = np.array(t, dtype=np.int32)
            sa = _fock(A, [x.real for x in oa.sigma(A.G0, topo)])
            sb = _fock(B, [x.real for x in ob.sigma(B.G0, topo)])
            assert np.abs(sa - sb).max() <= 1e-12 * np.abs(sa).max()


@pytest.mark.gpu
@pytest.mark.parametrize("norb,order", [(2, 2), (3, 3)])
def test_device_sigma_with_the_automatic_partition(norb, order):
    """Device Sigma (orders 1..order, solver signs) on the automatic partition against the same
    problem on (N, S_z) blocks, in the Fock basis (relative 1e-12), and against the oracle on the
    automatic partition for order 2."""
    from xca_amd.evaluator import DiagramEvaluator
    kw = dict(norb=norb, U=4.0, J=0.5, mu=(3 * 4.0 - 5 * 0.5) / 2, beta=2.0, lam=100.0, eps=1e-10)
    A = models.kanamori(**kw, qn="NSz")
    B = models.kanamori(**kw, qn="auto")
    ea = DiagramEvaluator(A.poles, A.coeffs, A.grid, A.ad)
    eb = DiagramEvaluator(B.poles, B.coeffs, B.grid, B.ad)
    sa = _fock(A, ea.pseudo_particle_self_energy(A.G0, order))
    sb = _fock(B, eb.pseudo_particle_self_energy(B.G0, order))
    err = np.abs(sa - sb).max() / np.abs(sa).max()
    print(f"[gpu] auto partition norb={norb} order={order} blocks {list(A.dims)} -> {list(B.dims)} rel={err:.2e}", flush=True)
    assert err <= 1e-12
    if order == 2:
        ob = po.OracleEvaluator(B, complex_arith=False)
        ref = 0
        for o in range(1, order + 1):
            for sgn, t in all_connected_pairings(o):
                ref = ref + (-1) ** o * sgn * (-1) ** (o + 1) * _fock(B, [x.real for x in ob.sigma(B.G0, np.array(t, dtype=np.int32))])
        assert np.abs(sb - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.gpu
def test_five_orbital_slater_automatic_partition():
    """5-orbital Slater-Condon d shell: the automatic partition (blocks of at most 16 states, which
    the tile engine or the large-block engine takes by size) gives the (N, S_z) partition's Sigma
    (blocks up to 100, large-block engine) in the Fock basis, orders 1-2, on a grid that resolves
    the atomic energies (beta = 2, Lambda = 800)."""
    from xca_amd.evaluator import DiagramEvaluator
    A = models.oca_5orb_slater(beta=2.0, lam=800.0, qn="NSz")
    B = models.oca_5orb_slater(beta=2.0, lam=800.0, qn="auto")
    assert max(A.dims) == 100 and max(B.dims) <= 16
    ea = DiagramEvaluator(A.poles, A.coeffs, A.grid, A.ad)
    eb = DiagramEvaluator(B.poles, B.coeffs, B.grid, B.ad)
    sa = _fock(A, ea.pseudo_particle_self_energy(A.G0, 2))
    sb = _fock(B, eb.pseudo_particle_self_energy(B.G0, 2))
    err = np.abs(sa - sb).max() / np.abs(sa).max()
    print(f"[gpu] 5-orbital automatic partition: {len(B.dims)} blocks <= {max(B.dims)}, rel={err:.2e}", flush=True)
    assert err <= 1e-12
