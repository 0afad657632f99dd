"""Parity helpers shared by the GPU tests.

The bar is max |S_gpu - S_ref| / max |S_ref| <= 1e-12 against the oracle on identical inputs. Where
the double-precision oracle itself is further than that from the exact result of those inputs (DLR
representations whose coefficients are much larger than their values), the exact result is taken
from the oracle in long double, and the device must be at most twice as far from it as the
double-precision oracle is; both distances are printed.
"""
import numpy as np

import pyoracle as po

TOL = 1e-12
EXT_FACTOR = 2.0


RUNS = 16  # flat-index lists are evaluated in this many contiguous runs on both sides


def device_runs(ev, G, t, fix, runs=RUNS):
    """The device Sigma of `fix` evaluated in the same contiguous runs as pyoracle.sigma_threads,
    summed in run order (each run is one compute_self_energy call, with its zero-block flags)."""
    parts = [ev.compute_self_energy(G, t, run) for run in po.split_runs(fix, runs)]
    return [sum(p[b] for p in parts) for b in range(len(parts[0]))]


def rel_err(a, b):
    num = max(np.abs(x - y).max() for x, y in zip(a, b))
    den = max(np.abs(y).max() for y in b)
    return num / den if den > 0 else num


def list_err(dev, ref, dropped):
    """Relative error of a flat-index-list evaluation beyond what the reference itself discarded:
    max over blocks of (max |dev_b - ref_b| - dropped_b)^+ / max |ref|. With a list, the reference
    re-flags blocks below 1e-16 after every flat index and replaces such a block at its next
    contribution, dropping that running sum (pyoracle.OracleEvaluator.last_dropped); the device keeps
    it. Blocks without such an event are held to the plain relative error."""
    den = max(float(np.abs(y).max()) for y in ref)
    ex = max(max(0.0, float(np.abs(x - y).max()) - float(d)) for x, y, d in zip(dev, ref, dropped))
    return ex / den if den > 0 else ex


def check_list(label, dev, P, G, t, fix, runs=RUNS, threads=None):
    """dev (device_runs of the same runs) against the double-precision oracle's list evaluation at
    TOL, beyond the reference's discarded running sums. Returns (plain error, error beyond drops)."""
    ref, drop = po.sigma_threads(P, G, t, fix, threads=threads, runs=runs, dropped=True)
    e, ex = rel_err(dev, ref), list_err(dev, ref, drop)
    print(f"[gpu] {label} rel_err={e:.2e}; beyond the reference's dropped sums ({drop.sum():.1e} absolute) {ex:.2e}", flush=True)
    assert max(np.abs(y).max() for y in ref) > 0
    assert ex <= TOL, f"{label}: {ex:.3e}"
    return e, ex


_EXT = {}


def _extended(P, G, t, allf, threads):
    """The long-double oracle's Sigma, memoised by the inputs (several tests revisit one case)."""
    import hashlib
    h = hashlib.sha1()
    for a in [P.grid.it, P.grid.rf, np.asarray(P.poles), np.asarray(P.coeffs), np.asarray(t), np.asarray(allf)] + list(G):
        h.update(np.ascontiguousarray(a).tobytes())
    key = h.hexdigest()
    if key not in _EXT:
        _EXT[key] = po.sigma_threads(P, G, t, allf, threads=threads, extended=True, runs=RUNS)
    return _EXT[key]


def assert_parity(label, dev, P, G, t, fix=None, orc=None, ref=None, threads=None):
    """dev against the double-precision oracle (ref, or computed), falling back to the extended-
    precision bar when the oracle's own rounding exceeds TOL. Returns (err vs double, err vs
    extended or None)."""
    if ref is None:
        if orc is not None and fix is None:
            ref = [x.real for x in orc.sigma(G, t)]
        else:
            allf = np.arange(orc.num_backbones(len(t)) if orc is not None else 0, dtype=np.int32) if fix is None else fix
            ref, drop = po.sigma_threads(P, G, t, allf, threads=threads, runs=RUNS, dropped=True)
            if fix is not None:
                e = list_err(dev, ref, drop)
                if e <= TOL:
                    print(f"[gpu] {label} rel_err={rel_err(dev, ref):.2e}, beyond the reference's dropped sums {e:.2e}", flush=True)
                    return e, None
    e = rel_err(dev, ref)
    if e <= TOL:
        print(f"[gpu] {label} rel_err={e:.2e}", flush=True)
        return e, None
    allf = fix if fix is not None else np.arange(po.OracleEvaluator(P, complex_arith=False).num_backbones(len(t)), dtype=np.int32)
    ext = _extended(P, G, t, allf, threads)
    e_dev, e_orc = rel_err(dev, ext), rel_err(ref, ext)
    print(f"[gpu] {label} rel_err={e:.2e} vs double oracle; against long double: device {e_dev:.2e}, oracle {e_orc:.2e}", flush=True)
    assert e_dev <= EXT_FACTOR * e_orc, f"{label}: device {e_dev:.3e} vs oracle {e_orc:.3e} from the extended-precision result"
    return e, e_dev
