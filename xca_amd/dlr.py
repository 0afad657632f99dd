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
        self.r = r
        self.rf = rf[:r].copy()
        self.it = it[:r].copy()
        self.tau = rel2abs(self.it) * self.beta           # absolute times in [0, beta]
        self.cf2it = kernel(self.it[:, None], self.rf[None, :])

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
