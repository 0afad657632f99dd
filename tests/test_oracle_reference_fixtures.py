"""Oracle against data files the reference's own tests hold (decoded by tests/golden/make_reference_fixtures.py).

OCA_tpz_result (test/c++/h5/tpz100.ref.h5, checked at test_block_sparse_OCA_manual.cpp:237-282): the OCA
self-energy of the two-band Kanamori atom coupled to a two-level discrete bath, computed by trapezoidal
quadrature on 101 equispaced tau points, tolerance 4e-4 there. The fixture is in TRIQS atom_diag's
eigenbasis and block order; the oracle works in the Fock basis. A block-diagonal unitary maps one to the
other and is tau independent, so the eigenvalues of Sigma_b(tau) at every tau are compared, block by
block. The fixture's blocks hold N = 1, 2, 0, 3, 4 particles in that order. The OCA contribution is the
topology ((0,2),(1,3)) with the solver's net factor -s = +1 (block_sparse_solver.py:609-638).
"""
import json
from pathlib import Path

import numpy as np
import pytest

import pyoracle as po
from xca_amd import models

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURE_N = [1, 2, 0, 3, 4]   # particle number of each fixture block


def equispaced_rel(n):
    """n equispaced points on [0, beta] as relative times; tau = beta is passed as t = 1."""
    t = np.arange(n) / (n - 1)
    rel = np.where(t > 0.5, t - 1.0, t)
    rel[-1] = 1.0
    return rel


def block_particle_numbers(ad):
    """Fock-state index bit o is the occupation of flavour o."""
    return [bin(int(f[0])).count("1") for f in ad.fock_states]


@pytest.fixture(scope="module")
def tpz():
    return json.loads((GOLDEN / "oca_tpz100_blocks.json").read_text())


def oca_on_equispaced(P, S, n):
    rel = equispaced_rel(n)
    return [po.interp(P.grid.it, P.grid.rf, s, rel).real for s in S]


def test_fixture_is_block_diagonal_and_real(tpz):
    assert tpz["max_abs_imag"] == 0.0 and tpz["max_abs_offblock"] == 0.0
    assert [len(b[0]) for b in tpz["blocks_real"]] == [4, 6, 1, 4, 1]


def test_oca_two_band_discrete_bath_matches_tpz100(tpz):
    P = models.two_band_discrete_bath()
    Nb = block_particle_numbers(P.ad)
    O = po.OracleEvaluator(P, complex_arith=False)
    S = [x.real for x in O.sigma(P.G0, np.array([[0, 2], [1, 3]], dtype=np.int32))]
    Seq = oca_on_equispaced(P, S, tpz["ntau"])
    worst = 0.0
    for k, fb in enumerate(tpz["blocks_real"]):
        b = Nb.index(FIXTURE_N[k])
        ev = np.linalg.eigvalsh(Seq[b])
        evf = np.linalg.eigvalsh(np.array(fb))
        err = np.abs(ev - evf).max()
        worst = max(worst, err)
        # the opposite overall sign is far off
        assert np.abs(np.sort(-ev, axis=1) - evf).max() > 1e-2
    assert worst <= 4e-4, worst
