"""Complex128 problems through the real embedding (xca_amd.cplx), pinned on the CPU:

- the embedding is an algebra homomorphism that maps the conjugate transpose to the transpose;
- the oracle in real arithmetic on the embedded problem equals the oracle in complex arithmetic on
  the original problem (Sigma per topology, the single-particle Green's function, a one-time
  correlator), on the complex-rotation fixture of test/python/unitary_invariance_bs.py, dense and
  block-sparse; a conjugated hybridization does not;
- the fixture itself: the oracle's complex self-consistency in two bases related by a complex
  rotation gives covariant Green's functions (unitary_invariance_bs.py:72-101,118-136), and the
  rotation is not of the kind that cannot see an untransposed backward index;
- the pole fit of a complex Hermitian hybridization recovers its exact poles and residues.
"""
import numpy as np
import pytest

import pyoracle as po
from xca_amd import cplx, hybfit, models
from xca_amd.dlr import kernel
from xca_amd.models import ImpurityProblem

U_CPLX = np.array([[1.0, 1j], [1.0, -1j]]) / np.sqrt(2)
TH = 0.4
U_REAL = np.array([[np.cos(TH), -np.sin(TH)], [np.sin(TH), np.cos(TH)]], dtype=complex)
TOPOS = [[[0, 1]], [[0, 2], [1, 3]], [[0, 3], [1, 2]]]


def embedded(P, coeffs=None):
    return ImpurityProblem(grid=P.grid, ad=cplx.embedded_atom(P.ad), poles=P.poles,
                           coeffs=cplx.embedded_coeffs(P.coeffs if coeffs is None else coeffs),
                           G0=[cplx.embed(g) for g in P.G0], eta0=P.eta0)


def rel(a, b):
    return max(np.abs(x - y).max() for x, y in zip(a, b)) / max(np.abs(y).max() for y in b)


def test_embedding_is_a_homomorphism():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((3, 4)) + 1j * rng.standard_normal((3, 4))
    B = rng.standard_normal((4, 2)) + 1j * rng.standard_normal((4, 2))
    assert np.allclose(cplx.embed(A @ B), cplx.embed(A) @ cplx.embed(B), rtol=0, atol=1e-14)
    assert np.array_equal(cplx.embed(A.conj().T), cplx.embed(A).T)
    assert np.array_equal(cplx.extract(cplx.embed(A)), A)
    assert np.array_equal(cplx.embed(2.5 * A), 2.5 * cplx.embed(A))
    # traces: 2 Re tr M, and -2 Im tr M with the factor i
    M = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    t = np.array([[np.trace(cplx.embed(M))], [np.trace(cplx.embed(1j * M))]])
    assert np.isclose(cplx.trace_from_copies(t, 1)[0, 0], np.trace(M), rtol=0, atol=1e-14)


@pytest.mark.parametrize("rotation", ["none", "complex"])
@pytest.mark.parametrize("qn", ["auto", "none"])
def test_oracle_on_embedding_equals_complex_oracle(rotation, qn):
    P = models.two_flavour_hopping(None if rotation == "none" else U_CPLX, qn=qn)
    assert cplx.is_complex(P.G0, P.coeffs)
    E = embedded(P)
    oc = po.OracleEvaluator(P, complex_arith=True)
    orr = po.OracleEvaluator(E, complex_arith=False)
    G = P.perturbed_G(seed=3, amp=1e-2)
    Ge = [cplx.embed(g) for g in G]
    for t in TOPOS:
        t = np.array(t, dtype=np.int32)
        a = oc.sigma(G, t)
        b = [cplx.extract(x.real) for x in orr.sigma(Ge, t)]
        assert max(np.abs(x.imag).max() for x in a) > 1e-4 * max(np.abs(x).max() for x in a)  # really complex
        assert rel(b, a) <= 1e-13, t.tolist()
        g = oc.spgf(G, t)
        ge = cplx.trace_from_copies(orr.spgf(Ge, t).real, 2)[:, :, :2]
        assert np.abs(ge - g).max() <= 1e-13 * np.abs(g).max()
    # a one-time correlator <n_0(tau) n_1(0)>, the operators given as Fock-basis blocks
    fs_blocks = P.ad.fock_states
    dims = P.ad.dims
    from xca_amd.atom import FockSpace
    fs = FockSpace(2)
    ops = []
    for o in range(2):
        N = (fs.cdag(o) @ fs.c(o)).toarray()
        conn = np.arange(len(dims), dtype=np.int32)[None, :]
        ops.append((conn, np.concatenate([N[np.ix_(f, f)].ravel() for f in fs_blocks])))
    t = np.array(TOPOS[1], dtype=np.int32)
    c = oc.correlator(G, t, mu_ops=ops[0], kap_ops=ops[1])
    mc, mm = cplx.embedded_ops(*ops[0], dims)
    kc, km = cplx.embedded_ops(*ops[1], dims)
    ce = orr.correlator(Ge, t, mu_ops=(mc, mm), kap_ops=(kc[:1], km[:km.size // 2]))
    assert np.abs(cplx.trace_from_copies(ce.real, 1) - c).max() <= 1e-13 * np.abs(c).max()


def test_conjugated_hybridization_is_detected():
    """A plausible mistake, the conjugate coefficient (Im with the wrong sign), changes Sigma."""
    P = models.two_flavour_hopping(U_CPLX)
    oc = po.OracleEvaluator(P, complex_arith=True)
    orr = po.OracleEvaluator(embedded(P, coeffs=P.coeffs.conj()), complex_arith=False)
    t = np.array(TOPOS[1], dtype=np.int32)
    a = oc.sigma(P.G0, t)
    b = [cplx.extract(x.real) for x in orr.sigma([cplx.embed(g) for g in P.G0], t)]
    assert rel(b, a) > 1e-3


def _spgf_after_one_iteration(P, order):
    G, eta, it, diff = po.ppsc_solve(P, order, tol=1e-12, maxiter=1, complex_arith=True)
    return po.spgf_sum(po.OracleEvaluator(P, complex_arith=True), G, order)


@pytest.mark.parametrize("order", [1, 2])
def test_fixture_covariance_on_the_oracle(order):
    """unitary_invariance_bs.py:118-136 on the oracle: solve in two bases related by the complex
    rotation (one iteration, hybridization given directly), rotate back, compare; and the real
    rotation control (:98-115)."""
    M = U_CPLX.conj().T @ U_CPLX.conj()
    assert np.max(np.abs(M - np.trace(M) / 2 * np.eye(2))) > 0.1
    for U in (U_REAL, U_CPLX):
        for qn in ("none", "auto"):
            g1 = _spgf_after_one_iteration(models.two_flavour_hopping(None, qn=qn), order)
            g2 = _spgf_after_one_iteration(models.two_flavour_hopping(U, qn=qn), order)
            back = np.einsum("ab,tbc,cd->tad", U, g2, U.conj().T)
            assert np.abs(g1 - back).max() < 1e-10 * np.abs(g1).max()
    # a wrong direction of the rotation (U^dagger g2 U) is not covariant
    assert np.abs(g1 - np.einsum("ab,tbc,cd->tad", U_CPLX.conj().T, g2, U_CPLX)).max() > 1e-4 * np.abs(g1).max()


def test_complex_pole_fit_recovers_exact_poles():
    P = models.two_flavour_hopping(U_CPLX)
    vals = np.einsum("tl,lab->tab", kernel(P.grid.it[:, None], P.beta * P.poles[None, :]), P.coeffs)
    x, W, err = hybfit.fit_hybridization(P.grid, vals, tol=1e-10, compression=True)
    assert err <= 1e-10 and len(x) == 2
    assert np.abs(x - P.poles).max() < 1e-9
    assert np.abs(W - P.coeffs).max() < 1e-9 * np.abs(P.coeffs).max()
    assert np.abs(W - W.conj().transpose(0, 2, 1)).max() < 1e-12


@pytest.mark.parametrize("qn", ["auto", "none"])
def test_library_host_checks_accept_the_embedding(qn):
    """The evaluator detects the complex problem, embeds it, and the C-ABI's host checks (symmetry
    sets from the repeated connections, cross-set coefficients, operator shapes) accept it; without a
    GPU only device selection fails."""
    from xca_amd._lib import XcaError
    from xca_amd.evaluator import DiagramEvaluator
    P = models.two_flavour_hopping(U_CPLX, qn=qn)
    try:
        DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)
    except XcaError as e:
        assert "no GPU device available" in str(e), str(e)
    with pytest.raises(XcaError, match="imaginary parts"):
        DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad, complex_arith=False)
