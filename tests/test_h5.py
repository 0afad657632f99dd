"""HDF5 persistence of the solver (test/python/h5_io_bs.py:30-52): a solver written to an archive and
read back compares equal, field by field, and its problem and hybridization come back intact. The
archive itself is checked on the value types it stores (arrays of every rank, complex, strings, None,
nested dicts, lists and tuples)."""
import numpy as np
import pytest

from xca_amd import h5, models
from xca_amd.solver import BlockSparseSolver, Hybridization

try:
    h5.lib()
except h5.H5Error as e:  # pragma: no cover - the library is part of this image
    pytest.skip(f"HDF5 C library unavailable: {e}", allow_module_level=True)


def test_archive_value_types(tmp_path):
    rng = np.random.default_rng(3)
    vals = {"a0": 1.5, "i0": 7, "a1": rng.standard_normal(5), "a3": rng.standard_normal((2, 3, 4)),
            "c2": rng.standard_normal((3, 2)) + 1j * rng.standard_normal((3, 2)), "i2": np.arange(6).reshape(2, 3),
            "s": "classic", "none": None, "empty": np.zeros((0, 3)),
            "nested": {"x": [1.0, np.ones(2), ("t", 3)], "y": {"z": np.eye(2)}}}
    f = tmp_path / "v.h5"
    with h5.HDFArchive(f, "w") as A:
        for k, v in vals.items():
            A[k] = v
    with h5.HDFArchive(f, "r") as A:
        assert sorted(A.keys()) == sorted(vals)
        for k in ("a0", "i0", "s"):
            assert A[k] == vals[k] and type(A[k]) is type(vals[k])
        assert A["none"] is None
        for k in ("a1", "a3", "c2", "i2", "empty"):
            assert np.array_equal(A[k], vals[k]) and A[k].dtype == np.asarray(vals[k]).dtype
        n = A["nested"]
        assert isinstance(n["x"], list) and n["x"][0] == 1.0 and np.array_equal(n["x"][1], np.ones(2))
        assert n["x"][2] == ("t", 3) and np.array_equal(n["y"]["z"], np.eye(2))
        with pytest.raises(KeyError):
            A["missing"]


def test_solver_roundtrip(tmp_path):
    """A solver with a compressed hybridization and the state a solve leaves (G, eta, history, the
    settings) written and read back: equal, and the pieces that drive the device path identical."""
    P = models.oca_2orb()
    S = BlockSparseSolver(P)
    S.fit_hybridization(tol=1e-9, compression=True)
    S.G = P.perturbed_G(seed=5)
    S.eta, S.max_order, S.normalization, S.hyb_comp = 0.125, 2, "classic", True
    S.history = [(1e-2, 1.01, 0.1), (3e-5, 1.0001, 0.125)]
    f = tmp_path / "s.h5"
    with h5.HDFArchive(f, "w") as A:
        A["S"] = S
    with h5.HDFArchive(f, "r") as A:
        R = A["S"]
    assert isinstance(R, BlockSparseSolver) and R == S
    assert isinstance(R.hyb, Hybridization) and R.hyb == S.hyb and len(R.hyb.poles) < P.grid.r
    assert R.problem.grid == P.grid
    assert R.history == S.history and R.eta == S.eta and R.normalization == "classic"
    # the operator data the device evaluator packs is the same
    a, b = P.ad.packed_ops(), R.problem.ad.packed_ops()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    for g, h in zip(S.G, R.G):
        assert np.array_equal(g, h)
    # observables that need the Fock basis still work on the reloaded atom
    assert [np.array_equal(x, y) for x, y in zip(P.ad.fock_states, R.problem.ad.fock_states)] == [True] * len(P.ad.fock_states)
    # a changed field breaks equality
    R.eta += 1e-15
    assert R != S


def test_unregistered_format_raises(tmp_path):
    class Thing:
        def __reduce_to_dict__(self):
            return {"x": 1.0}
    f = tmp_path / "u.h5"
    with h5.HDFArchive(f, "w") as A:
        A["t"] = Thing()
    with h5.HDFArchive(f, "r") as A, pytest.raises(h5.H5Error, match="no class registered"):
        A["t"]
