"""xca_dyson_solve_noalloc (the Dyson solve that allocates nothing, its scratch in a static device array)
against xca_dyson_solve on the same inputs, called through the C-ABI: two calls in a row, the second
on a smaller problem, and a problem beyond the static maximum rejected with a clear error."""
import numpy as np
import pytest

from xca_amd import models
from xca_amd._lib import XcaError, dptr, last_error, lib

pytestmark = pytest.mark.gpu


def _noalloc():
    return lib().xca_dyson_solve_noalloc


def _three_orbital():
    U = 1.42
    J = U / 8
    Up = U - 2 * J
    return models.kanamori(3, U, J, (U + 2 * (2 * Up - J)) / 2, beta=90.6, lam=20 * 90.6)


def _compare(P, order, eta):
    from xca_amd.evaluator import DiagramEvaluator
    ev = DiagramEvaluator(P.poles, P.coeffs, P.grid, P.ad)
    S = ev.pseudo_particle_self_energy(P.G0, order)
    ev.dyson_init(P.G0)
    ref = ev.dyson_solve(S, eta)
    g0 = ev.pack(P.G0)
    s = ev.pack(S)
    out = np.full_like(g0, np.nan)
    st = _noalloc()(ev._h, dptr(g0), dptr(s), float(eta), dptr(out))
    if st != 0:
        raise XcaError(last_error())
    got = ev.unpack(out)
    e = max(np.abs(a - b).max() for a, b in zip(got, ref)) / max(np.abs(b).max() for b in ref)
    print(f"[gpu] no-alloc dyson {P.name} r={P.grid.r} max N={P.grid.r * max(P.dims)} rel diff {e:.2e}", flush=True)
    return got, ref, e


def test_noalloc_dyson_equals_allocating_entry_point():
    # first call: the 3-orbital problem, systems up to 513 x 513 (in the static scratch)
    got, ref, e = _compare(_three_orbital(), 2, 0.013)
    assert all(np.isfinite(a).all() for a in got)
    assert all(np.array_equal(a, b) for a, b in zip(got, ref)), e
    # second call, smaller problem (systems in LDS): the scratch keeps the first call's contents
    got, ref, e = _compare(models.anderson_nca(), 1, 0.2)
    assert all(np.array_equal(a, b) for a, b in zip(got, ref)), e


def test_noalloc_dyson_rejects_problems_beyond_the_static_maximum():
    """Eight blocks of 12 states at r = 64 need 8 systems of 768 x 768 doubles, above 32 MiB."""
    from xca_amd.atom import AtomDiagArrays
    from xca_amd.dlr import DLRGrid
    from xca_amd.evaluator import DiagramEvaluator
    lam = 4000.0
    grid = DLRGrid(100.0, lam, 1e-10)
    while grid.r < 64:
        lam *= 1.2
        grid = DLRGrid(100.0, lam, 1e-10)
    if grid.r > 64:
        pytest.skip("no grid of rank exactly 64 on this ladder")
    nb, d = 8, 12
    rng = np.random.default_rng(0)
    cb = {(0, j): rng.standard_normal((d, d)) for j in range(nb - 1)}
    db = {(0, j + 1): cb[(0, j)].T.copy() for j in range(nb - 1)}
    cc = np.array([[j + 1 if j < nb - 1 else -1 for j in range(nb)]])
    dc = np.array([[j - 1 if j > 0 else -1 for j in range(nb)]])
    ad = AtomDiagArrays(1, [d] * nb, cc, dc, cb, db)
    ev = DiagramEvaluator(np.array([0.0]), np.ones((1, 1, 1)), grid, ad)
    G0 = [-np.ones((grid.r, d, d)) * 0.1 for _ in range(nb)]
    g0 = ev.pack(G0)
    out = np.zeros_like(g0)
    assert _noalloc()(ev._h, dptr(g0), dptr(np.zeros_like(g0)), 0.0, dptr(out)) != 0
    msg = last_error()
    print(f"[gpu] no-alloc rejection: {msg}", flush=True)
    assert "static maximum" in msg
