"""Hybridization pole compression: Delta(tau) with fewer poles than the DLR rank (p < r).

The hot path's cost grows as p^(m-1) with the diagram order m, so a third-order evaluation with p
poles instead of r costs (p/r)^2 as much. This module replaces the pole fit the reference solver
calls before its first iteration, BlockSparseSolver.fit_hybridization
(python/triqs_xca/block_sparse_solver.py:146-185), whose compressed path is the third-party adapol
(absent here) and whose legacy path is python/triqs_xca/adapol_depr (fit_utils_xca.py:136-206,
aaa.py:22-106). The method is the published one those follow:

1. AAA rational approximation (Nakatsukasa, Sete, Trefethen, SIAM J. Sci. Comput. 40, A1494 (2018))
   of the matrix-valued Delta(i nu) on fermionic Matsubara frequencies, with support points in
   conjugate pairs and weights constrained to conjugate pairs, so that the rational function
   satisfies Delta(conj z) = Delta(z)^dagger (real coefficients for a real symmetric Delta); its real
   poles are the candidate pole energies. A complex (Hermitian) Delta is fitted entry by entry with
   entry (a, b) mirrored by (b, a), and gets complex residues.
2. Residues by linear least squares against the DLR representation of Delta(tau) on a dense time
   grid (clustered geometrically at both ends, where the DLR basis varies on the scale 1/Lambda).
3. Poles whose residues are negligible are dropped, and the remaining poles are refined by
   nonlinear least squares (variable projection: the residues are re-solved for every pole set).
   Once a fit meets tol, its lightest pole is dropped and the rest refitted for as long as the fit
   still meets tol.
4. The AAA degree grows until the fit error is below tol.

fit_hybridization keeps the reference's contract: without compression, or when the fit fails, the
DLR expansion itself is returned (p = r poles at the DLR frequencies).

The fit error is max over the dense grid and over the matrix entries of |Delta_fit - Delta|, divided
by max |Delta|.
"""
import numpy as np
import scipy.linalg
from scipy.optimize import least_squares
from scipy.special import expit


def kernel_rel(t, w):
    """K(t, w) = -exp(-w t) / (1 + exp(-w)) for relative time t in [0, 1] and dimensionless w, in the
    form that does not overflow for either sign of w; t (k,), w (m,) -> (k, m)."""
    t = np.asarray(t, dtype=np.float64)[:, None]
    w = np.asarray(w, dtype=np.float64)[None, :]
    wp, wm = np.maximum(w, 0.0), np.minimum(w, 0.0)
    return np.where(w >= 0, -np.exp(-wp * t) * expit(wp), -np.exp((1.0 - t) * wm) * expit(-wm))


def dense_grid(lam, per_decade=24):
    """Relative times in [0, 1]: geometric in the distance to the nearer end, from 1e-2 / lam to 1/2."""
    lo = 1e-2 / max(lam, 1.0)
    n = int(np.ceil(per_decade * np.log10(0.5 / lo))) + 1
    h = np.geomspace(lo, 0.5, n)
    return np.unique(np.concatenate([[0.0], h, 1.0 - h, [1.0]]))


def matsubara_values(w, coeffs, beta, k):
    """Delta(i nu_k), nu_k = (2k + 1) pi / beta, of Delta(tau) = sum_l K(tau/beta, w_l) c_l with w
    dimensionless: Delta(i nu) = sum_l c_l / (i nu - w_l / beta). k (K,) -> (K, n, n)."""
    nu = (2 * np.asarray(k, dtype=np.float64) + 1) * np.pi / beta
    den = 1j * nu[:, None] - np.asarray(w)[None, :] / beta
    return np.einsum("kl,lab->kab", 1.0 / den, coeffs)


def aaa_real_poles(z, F, degree, mirror=None):
    """Real poles of the AAA approximant of degree `degree` (even) of the matrix-valued samples
    F (K, e) at points z (K,) on the upper imaginary axis, mirrored to the lower half plane with
    F(conj z) = conj F(z) entry by entry, or, with mirror (a permutation of the e entries),
    F_e(conj z) = conj F_mirror[e](z) (a Hermitian hybridization: entry (a, b) mirrors (b, a)).
    Support points and barycentric weights come in conjugate pairs."""
    K, e = F.shape
    Z = np.concatenate([z, np.conj(z)])
    FF = np.concatenate([F, np.conj(F if mirror is None else F[:, mirror])], axis=0)
    N = 2 * K
    support = []
    fit = np.full_like(FF, FF.mean())
    rest = list(range(N))
    w = None
    for _ in range(degree // 2):
        j = int(np.argmax(np.sum(np.abs(FF[rest] - fit[rest]) ** 2, axis=1)))
        j = rest[j]
        jc = (j + K) % N
        support += [j, jc]
        rest = [i for i in rest if i != j and i != jc]
        s = Z[support]
        C = 1.0 / (Z[rest, None] - s[None, :])
        # Loewner matrix of every entry stacked: A[(x, entry), k] = (F(x) - F(s_k)) / (x - s_k)
        A = ((FF[rest, None, :] - FF[support][None, :, :]) * C[:, :, None]).transpose(0, 2, 1).reshape(-1, len(support))
        # weights (u + i v, u - i v) on each conjugate pair: a real least-squares problem in (u, v)
        Al, Ar = A[:, 0::2], A[:, 1::2]
        B = np.concatenate([Al + Ar, 1j * (Al - Ar)], axis=1)
        B = np.concatenate([B.real, B.imag], axis=0)
        _, _, vh = scipy.linalg.svd(B, full_matrices=False)
        uv = vh[-1].reshape(2, -1)
        w = np.empty(len(support), dtype=np.complex128)
        w[0::2] = uv[0] + 1j * uv[1]
        w[1::2] = uv[0] - 1j * uv[1]
        fit = FF.copy()
        num = C @ (w[:, None] * FF[support])
        den = C @ w
        fit[rest] = num / den[:, None]
    s = Z[support]
    m = len(w)
    # poles: finite eigenvalues of the arrowhead pencil of the barycentric denominator
    E = np.zeros((m + 1, m + 1), dtype=np.complex128)
    Bm = np.eye(m + 1, dtype=np.complex128)
    Bm[0, 0] = 0.0
    E[0, 1:] = w
    E[1:, 0] = 1.0
    E[1:, 1:] = np.diag(s)
    ev = scipy.linalg.eigvals(E, Bm)
    ev = ev[np.isfinite(ev)]
    return ev


def residues(t, x, target):
    """Least-squares residues (p, e) of poles x (dimensionless) against target (k, e) at times t."""
    M = kernel_rel(t, x)
    W, *_ = np.linalg.lstsq(M, target, rcond=None)
    return W, M @ W - target


def compress(w_dlr, coeffs, beta, tol, max_poles=None, lam=None, verbose=False):
    """Poles (dimensionless, beta * energy) and residues (p, n, n) of a hybridization given by its
    DLR expansion (w_dlr dimensionless, coeffs (r, n, n)) with relative fit error below tol, or the
    best fit found. Real coefficients: Delta symmetric, real residues, the upper triangle fitted.
    Complex coefficients: Delta Hermitian, complex residues, every nonzero entry fitted, the mirror
    pairing (a, b) with (b, a). Returns (x, W, error)."""
    coeffs = np.asarray(coeffs)
    cx = np.iscomplexobj(coeffs) and bool(np.any(coeffs.imag != 0))
    coeffs = coeffs.astype(np.complex128 if cx else np.float64)
    r, n, _ = coeffs.shape
    lam = lam or float(np.abs(w_dlr).max())
    t = dense_grid(lam)
    if cx:
        iu = [(a, b) for a in range(n) for b in range(n) if np.abs(coeffs[:, a, b]).max() > 0 or np.abs(coeffs[:, b, a]).max() > 0]
        pos = {ab: k for k, ab in enumerate(iu)}
        mirror = np.array([pos[(b, a)] for a, b in iu], dtype=np.int64)
    else:  # the independent nonzero entries (Delta is symmetric)
        iu = [(a, b) for a in range(n) for b in range(a, n) if np.abs(coeffs[:, a, b]).max() > 0]
        mirror = None
    if not iu:
        return np.zeros(0), np.zeros((0, n, n), dtype=coeffs.dtype), 0.0
    C = np.stack([coeffs[:, a, b] for a, b in iu], axis=1)   # (r, e)
    target = kernel_rel(t, w_dlr) @ C                          # (k, e)
    scale = float(np.abs(target).max())
    kk = np.unique(np.concatenate([np.arange(64), np.round(np.geomspace(64, max(128.0, 4 * lam / (2 * np.pi)), 96))]).astype(int))
    z = 1j * (2 * kk + 1) * np.pi / beta
    F = np.stack([matsubara_values(w_dlr, coeffs, beta, kk)[:, a, b] for a, b in iu], axis=1)
    max_poles = max_poles or r
    best = (np.inf, None, None)
    for degree in range(2, 2 * max_poles + 1, 2):
        pol = aaa_real_poles(z, F, degree, mirror)
        pol = pol[np.abs(pol.imag) <= 1e-6 * max(1.0, np.abs(pol).max())].real
        x = np.sort(beta * pol)
        if len(x) == 0 or len(x) > max_poles:
            continue
        W, res = residues(t, x, target)
        keep = np.abs(W).max(axis=1) > 1e-3 * tol * scale
        x = x[keep]
        if len(x) == 0:
            continue

        def fun(xx):
            res = residues(t, xx, target)[1].ravel() / scale
            return np.concatenate([res.real, res.imag]) if cx else res

        while True:  # refine; drop the poles whose residues fall below the tolerance, then refine again
            sol = least_squares(fun, x, method="lm", xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=200 * (len(x) + 1))
            x = np.sort(sol.x)
            W, res = residues(t, x, target)
            small = np.abs(W).max(axis=1) <= 0.1 * tol * scale
            if not small.any() or small.all():
                break
            x = x[~small]
        err = float(np.abs(res).max()) / scale
        while err <= tol and len(x) > 1:  # the lightest pole goes while the refitted rest still meets tol
            drop = int(np.argmin(np.abs(W).max(axis=1)))
            sol = least_squares(fun, np.delete(x, drop), method="lm", xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=200 * len(x))
            x2 = np.sort(sol.x)
            W2, res2 = residues(t, x2, target)
            err2 = float(np.abs(res2).max()) / scale
            if err2 > tol:
                break
            x, W, res, err = x2, W2, res2, err2
        if verbose:
            print(f"hybfit: degree {degree}: {len(x)} poles, error {err:.2e}")
        if err < best[0]:
            best = (err, x, W)
        if err <= tol:
            break
    err, x, W = best
    Wf = np.zeros((len(x), n, n), dtype=coeffs.dtype)
    for k, (a, b) in enumerate(iu):
        Wf[:, a, b] = W[:, k]
        if not cx:
            Wf[:, b, a] = W[:, k]
    return x, Wf, err


def fit_hybridization(grid, delta_vals, tol=None, compression=False, max_poles=None, verbose=False):
    """BlockSparseSolver.fit_hybridization (block_sparse_solver.py:146-185): poles (energy units),
    residues (p, n, n) and the fit error of the hybridization given by its values (r, n, n) at the
    grid's nodes. With compression, a pole fit to relative error tol (tol required and positive); if
    it does not reach tol, or without compression, the DLR expansion (p = r, error = the grid's
    eps)."""
    delta_vals = np.asarray(delta_vals)
    if not (np.iscomplexobj(delta_vals) and np.any(delta_vals.imag != 0)):
        delta_vals = np.asarray(delta_vals.real, dtype=np.float64)
    coeffs = grid.vals2coefs(delta_vals)
    if compression:
        if tol is None or not tol > 0:
            raise ValueError("fit_hybridization: tol must be provided and positive when compression is enabled")
        x, W, err = compress(grid.rf, coeffs, grid.beta, tol, max_poles=max_poles, lam=grid.lam, verbose=verbose)
        if len(x) > 0 and err <= tol:
            return x / grid.beta, W, err
        if verbose:
            print(f"hybfit: no fit reached tol = {tol:.1e} (best {err:.2e}); using the DLR expansion")
    return grid.rf / grid.beta, coeffs, grid.eps
