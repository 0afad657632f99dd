"""DLR imaginary-time grid (Python side of xca_dlr_build / xca_dlr_kernel).

Replaces the cppdlr objects the reference builds in python/triqs_xca/block_sparse_solver.py:105,136
(MeshDLRImTime, ImTimeOps). Times are in relative format; frequencies are dimensionless (beta*w).
"""
import numpy as np

from ._lib import as_f64, check, dptr, lib


def kernel(t, om):
    """Fermionic kernel K(t, om) = -exp(-om t)/(1 + exp(-om)), relative-format t, broadcasting."""
    t = np.asarray(t, dtype=np.float64)
    om = np.asarray(om, dtype=np.float64)
    tb, ob = np.broadcast_arrays(t, om)
    f = lib().xca_dlr_kernel
    out = np.empty(tb.shape)
    for idx in np.ndindex(tb.shape):
        out[idx] = f(float(tb[idx]), float(ob[idx]))
    return out


def rel2abs(t):
    t = np.asarray(t, dtype=np.float64)
    return np.where(t < 0, 1.0 + t, t)


class DLRGrid:
    """DLR nodes and the basic (host, numpy) transforms on them for one (beta, lambda, eps)."""

    def __init__(self, beta, lam, eps, max_r=512):
        self.beta, self.lam, self.eps = float(beta), float(lam), float(eps)
        rf = np.zeros(max_r)
        it = np.zeros(max_r)
        r = check(lib().xca_dlr_build(self.lam, self.eps, max_r, dptr(rf), dptr(it)))
        self._set_nodes(rf[:r].copy(), it[:r].copy())

    def _set_nodes(self, rf, it):
        self.r = len(rf)
        self.rf = rf
        self.it = it
        self.tau = rel2abs(self.it) * self.beta           # absolute times in [0, beta]
        self.cf2it = kernel(self.it[:, None], self.rf[None, :])

    @classmethod
    def from_nodes(cls, beta, it, rf, lam=None, eps=None):
        """Grid on externally supplied nodes, e.g. those of the reference's MeshDLRImTime /
        ImTimeOps (block_sparse_solver.py:105,136): it in relative format (t in [0, 1/2] as t, t in
        (1/2, 1) as t - 1; absolute times in [0, beta] are accepted and converted), rf dimensionless
        (beta * w); times tau / beta in [0, 1] are accepted and converted. lam and eps are
        informational."""
        it = np.array(it, dtype=np.float64).ravel()
        rf = np.array(rf, dtype=np.float64).ravel()
        if it.shape != rf.shape or it.size == 0:
            raise ValueError(f"DLRGrid.from_nodes: it and rf must be nonempty and of equal length, got {it.size} and {rf.size}")
        if np.any(np.abs(it) > 1.0):
            raise ValueError("DLRGrid.from_nodes: times must be relative (tau / beta), not absolute")
        it = np.where(it > 0.5, it - 1.0, it)
        g = cls.__new__(cls)
        g.beta = float(beta)
        g.lam = float(lam) if lam is not None else float(np.abs(rf).max())
        g.eps = float(eps) if eps is not None else float("nan")
        g._set_nodes(rf, it)
        return g

    def __reduce_to_dict__(self):
        return {"beta": self.beta, "lam": self.lam, "eps": self.eps, "it": self.it, "rf": self.rf}

    @classmethod
    def __factory_from_dict__(cls, name, d):
        return cls.from_nodes(d["beta"], d["it"], d["rf"], lam=d["lam"], eps=d["eps"])

    def __eq__(self, other):
        return (isinstance(other, DLRGrid) and (self.beta, self.lam) == (other.beta, other.lam)
                and (self.eps == other.eps or (self.eps != self.eps and other.eps != other.eps))
                and np.array_equal(self.it, other.it) and np.array_equal(self.rf, other.rf))

    def vals2coefs(self, g):
        g = np.asarray(g)
        shp = g.shape
        c = np.linalg.solve(self.cf2it, g.reshape(self.r, -1))
        return c.reshape(shp)

    def coefs2vals(self, c):
        c = np.asarray(c)
        return (self.cf2it @ c.reshape(self.r, -1)).reshape(c.shape)

    def eval(self, c, t_rel):
        """Evaluate from coefficients at relative times t_rel (1-d array)."""
        c = np.asarray(c)
        K = kernel(np.asarray(t_rel)[:, None], self.rf[None, :])
        return (K @ c.reshape(self.r, -1)).reshape((len(t_rel),) + c.shape[1:])
