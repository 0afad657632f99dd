"""The interaction Hamiltonians of the benchmark workloads pinned by mathematics, independently of the
code that builds them: multiplet spectra from textbook closed forms, conservation laws, and the
particle-hole symmetry that fixes the half-filling chemical potentials.

A wrong sign of the spin-flip or pair-hopping term, U' != U - 2J, a wrong Slater-integral
normalisation (F2/49, F4/441) or a wrong half-filling mu fails one of these.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from xca_amd import models
from xca_amd.atom import FockSpace, kanamori_hamiltonian, slater_hamiltonian


def sector(fs, H, N):
    """Dense block of H on the Fock states with N particles."""
    states = [s for s in range(fs.dim) if bin(s).count("1") == N]
    return sp.csr_matrix(H)[states, :][:, states].toarray()


def spectrum(fs, H, N):
    return np.sort(np.linalg.eigvalsh(sector(fs, H, N)))


def levels(values, mults):
    return np.sort(np.repeat(np.asarray(values, dtype=float), mults))


def total_N(fs):
    return sum(fs.nop(o) for o in range(fs.n))


def spin_ops(fs, norb):
    """S_z and S^2 with flavours (down 0..norb-1), (up 0..norb-1)."""
    Sz = 0.5 * sum(fs.nop(norb + a) - fs.nop(a) for a in range(norb))
    Sp = sum(fs.cdag(norb + a) @ fs.c(a) for a in range(norb))
    Sm = Sp.T
    return Sz, (Sm @ Sp + Sz @ Sz + Sz)


def comm_norm(A, B):
    C = (A @ B - B @ A)
    return abs(C).max() if C.nnz else 0.0


# --------------------------------------------------------------------------------- Kanamori

@pytest.mark.parametrize("U,J", [(4.0, 0.5), (5.0, 0.7), (3.0, 1.1)])
def test_kanamori_two_orbital_two_particle_spectrum(U, J):
    """Two electrons in two orbitals (test/c++/block_sparse_utils.cpp:189-208 interaction): the
    inter-orbital triplet at U' - J = U - 3J (3 states), the inter-orbital singlet U' + J and the
    antisymmetric pair state U - J (2 states), the symmetric pair state U + J (1 state)."""
    fs = FockSpace(4)
    H = kanamori_hamiltonian(fs, 2, U, J, 0.0)
    assert np.allclose(spectrum(fs, H, 2), levels([U - 3 * J, U - J, U + J], [3, 2, 1]), rtol=0, atol=1e-12)
    assert np.allclose(spectrum(fs, H, 1), 0.0, atol=1e-14)


@pytest.mark.parametrize("U,J", [(5.0, 0.7), (4.0, 0.5)])
def test_kanamori_three_orbital_two_particle_spectrum(U, J):
    """t2g^2 with U' = U - 2J (Georges, de' Medici, Mravlje, Annu. Rev. Condens. Matter Phys. 4, 137
    (2013)): 3T1 at U - 3J (9 states), 1T2 + 1E at U - J (5 states), 1A1 at U + 2J (1 state). The pair
    hopping sign decides U + 2J against U - 2J."""
    fs = FockSpace(6)
    H = kanamori_hamiltonian(fs, 3, U, J, 0.0)
    assert np.allclose(spectrum(fs, H, 2), levels([U - 3 * J, U - J, U + 2 * J], [9, 5, 1]), rtol=0, atol=1e-12)
    # three electrons: the high-spin ground state 4A2 at 3(U - 3J) (4 states)
    e3 = spectrum(fs, H, 3)
    assert np.allclose(e3[:4], 3 * (U - 3 * J), atol=1e-12) and e3[4] > 3 * (U - 3 * J) + 1e-3
    # the same check fails with the pair-hopping sign flipped
    pair = sum(fs.cdag(3 + a) @ fs.cdag(a) @ fs.c(3 + b) @ fs.c(b) for a in range(3) for b in range(3) if a != b)
    Hw = H + 2 * J * pair
    assert not np.allclose(spectrum(fs, Hw, 2), levels([U - 3 * J, U - J, U + 2 * J], [9, 5, 1]), rtol=0, atol=1e-6)


@pytest.mark.parametrize("norb", [2, 3])
def test_kanamori_conservation_laws(norb):
    """[H, N] = [H, S_z] = [H, S^2] = 0, and for three orbitals with U' = U - 2J the orbital
    rotation invariance [H, L^2] = 0 with the t2g angular momentum (L_k)_{ab} = -i eps_{kab}."""
    fs = FockSpace(2 * norb)
    H = kanamori_hamiltonian(fs, norb, 4.3, 0.6, 1.7)
    Sz, S2 = spin_ops(fs, norb)
    assert comm_norm(H, total_N(fs)) < 1e-12
    assert comm_norm(H, Sz) < 1e-12
    assert comm_norm(H, S2) < 1e-12
    if norb == 3:
        eps = np.zeros((3, 3, 3))
        for (a, b, c), s in (((0, 1, 2), 1), ((1, 2, 0), 1), ((2, 0, 1), 1), ((0, 2, 1), -1), ((2, 1, 0), -1), ((1, 0, 2), -1)):
            eps[a, b, c] = s
        L2 = None
        for k in range(3):
            Lk = sum(-1j * eps[k, a, b] * (fs.cdag(sig * 3 + a) @ fs.c(sig * 3 + b)) for sig in range(2) for a in range(3) for b in range(3)
                     if eps[k, a, b] != 0)
            L2 = Lk @ Lk if L2 is None else L2 + Lk @ Lk
        assert comm_norm(sp.csr_matrix(H, dtype=complex), L2) < 1e-12
        # with a different U' the rotation invariance is lost: the check has teeth
        Hb = H + 0.3 * sum(fs.nop(sig * 3) @ fs.nop(sig2 * 3 + 1) for sig in range(2) for sig2 in range(2))
        assert comm_norm(sp.csr_matrix(Hb, dtype=complex), L2) > 1e-3


def test_kanamori_half_filling_mu_is_particle_hole_symmetric():
    """mu = (3U - 5J)/2 (two orbitals) and mu = [U + (M - 1)(2U' - J)]/2 (M orbitals, the serving
    stream): the spectra of sectors N and 2M - N coincide."""
    for norb, U, J, mu in ((2, 4.0, 0.5, (3 * 4.0 - 5 * 0.5) / 2), (3, 4.8, 0.6, None), (1, 2.0, 0.25, None)):
        if mu is None:
            Up = U - 2 * J
            mu = (U + (norb - 1) * (2 * Up - J)) / 2
        fs = FockSpace(2 * norb)
        H = kanamori_hamiltonian(fs, norb, U, J, mu)
        for N in range(norb):
            assert np.allclose(spectrum(fs, H, N), spectrum(fs, H, 2 * norb - N), atol=1e-11)


def test_tca_filling_two_window():
    """The three-orbital TCA workload's mu = 4.35 is the midpoint of the window in which N = 2 is the
    atomic ground-state sector, (E0(2) - E0(1), E0(3) - E0(2)) = (U - 3J, 2(U - 3J)) = (2.9, 5.8)."""
    U, J = 5.0, 0.7
    fs = FockSpace(6)
    H = kanamori_hamiltonian(fs, 3, U, J, 0.0)
    e0 = [spectrum(fs, H, N)[0] for N in range(4)]
    lo, hi = e0[2] - e0[1], e0[3] - e0[2]
    assert lo == pytest.approx(2.9, abs=1e-12) and hi == pytest.approx(5.8, abs=1e-12)
    P = models.tca_3orb()
    gs = [np.linalg.eigvalsh(h).min() for h in P.ad.H_blocks]
    b = int(np.argmin(gs))
    assert bin(int(P.ad.fock_states[b][0])).count("1") == 2


# --------------------------------------------------------------------------------- Slater-Condon

def racah(F0, F2, F4):
    """Racah parameters from the Slater integrals: F_2 = F2/49, F_4 = F4/441,
    A = F0 - 49 F_4, B = F_2 - 5 F_4, C = 35 F_4 (Condon & Shortley normalisation)."""
    f2, f4 = F2 / 49.0, F4 / 441.0
    return F0 - 49 * f4, f2 - 5 * f4, 35 * f4


@pytest.mark.parametrize("F", [(4.0, 6.0, 3.75), (2.0, 8.0, 5.0), (0.0, 1.0, 0.0)])
def test_slater_d2_multiplets(F):
    """d^2 terms (Griffith, The Theory of Transition-Metal Ions, table A6): 3F = A - 8B (21 states),
    3P = A + 7B (9), 1G = A + 4B + 2C (9), 1D = A - 3B + 2C (5), 1S = A + 14B + 7C (1)."""
    A, B, C = racah(*F)
    fs = FockSpace(10)
    H = slater_hamiltonian(fs, 2, F, 0.0)
    want = levels([A - 8 * B, A + 7 * B, A + 4 * B + 2 * C, A - 3 * B + 2 * C, A + 14 * B + 7 * C], [21, 9, 9, 5, 1])
    assert np.allclose(spectrum(fs, H, 2), want, rtol=0, atol=1e-12)


def test_slater_conservation_laws():
    """Rotational invariance of the Slater-Condon interaction: [H, L^2] = [H, L_z] = [H, S^2] =
    [H, N] = 0, with L in the complex spherical-harmonic basis (flavours (down, m = -2..2),
    (up, m = -2..2))."""
    l = 2
    fs = FockSpace(10)
    H = slater_hamiltonian(fs, l, (4.0, 6.0, 3.75), 1.3)
    ms = list(range(-l, l + 1))
    Lz = sum(m * fs.nop(s * 5 + i) for s in range(2) for i, m in enumerate(ms))
    Lp = sum(np.sqrt(l * (l + 1) - m * (m + 1)) * (fs.cdag(s * 5 + i + 1) @ fs.c(s * 5 + i)) for s in range(2) for i, m in enumerate(ms[:-1]))
    L2 = Lp.T @ Lp + Lz @ Lz + Lz
    Sz, S2 = spin_ops(fs, 5)
    for X in (L2, Lz, S2, Sz, total_N(fs)):
        assert comm_norm(H, X) < 1e-11


def test_slater_half_filling_mu_is_particle_hole_symmetric():
    """mu = (45 F0 - 70/49 F2 - 630/441 F4)/10 (benchmarks/slater_condon/slater_condon_interaction.py:29)
    makes the spectra of N and 10 - N equal; 0.01 off it does not."""
    F0, F2, F4 = 4.0, 6.0, 3.75
    mu = (45 * F0 - 70 / 49 * F2 - 630 / 441 * F4) / 10
    fs = FockSpace(10)
    H = slater_hamiltonian(fs, 2, (F0, F2, F4), mu)
    for N in (1, 3, 4):
        assert np.allclose(spectrum(fs, H, N), spectrum(fs, H, 10 - N), atol=1e-10)
    Hb = slater_hamiltonian(fs, 2, (F0, F2, F4), mu + 0.01)
    assert np.abs(spectrum(fs, Hb, 4) - spectrum(fs, Hb, 6)).max() > 1e-3
