"""Extended-precision reference values for the three-orbital TCA workload as specified.

At beta = 100, Lambda = 2000 the atomic propagator holds decay rates up to beta * E_max = 3370,
beyond the DLR cutoff. Every double-precision evaluation of Sigma from its DLR representation then
carries rounding error of order 1e-10 relative, the reference algorithm included. The reference
values here are the oracle's evaluation of the same inputs in long double (x87 extended precision,
64-bit mantissa; oracle/_build/liboracle_ld.so), which is accurate to ~1e-13 on them.

The fixture holds the inputs as float64 (DLR nodes, hybridization coefficients, the seed-perturbed
G), so that a test evaluates exactly these doubles on any host: recomputing them would go through
LAPACK and libm, whose rounding may differ between machines, and at this conditioning a 1e-16 change
of the inputs moves Sigma by ~1e-11. Entries topo0..topo3 are the third-order topologies, each with a
seeded sample of flat backbone indices; topo4 and topo5 are orders 1 and 2 with every flat index.
sigmaK is the extended-precision Sigma summed over fixK (unsigned, per topology, as
compute_self_energy(G, topology, f_ix_vec) returns it), evaluated in `runs` contiguous runs of fixK,
one call each, as the reference scatters flat indices over ranks.

Run: python tests/golden/make_extended_fixtures.py [--per-topology N] [--threads T]
Writes tests/golden/tca3_extended.npz.
"""
import argparse
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

OUT = Path(__file__).resolve().parent / "tca3_extended.npz"


def tca3_atom():
    from xca_amd.atom import AtomDiag, FockSpace, kanamori_hamiltonian, kanamori_spins
    fs = FockSpace(6)
    return AtomDiag(fs, kanamori_hamiltonian(fs, 3, 5.0, 0.7, 4.35), qn="NSz", spin=kanamori_spins(3))


def problem_from_fixture(fx):
    """The TCA workload rebuilt from the stored float64 inputs (operators are exact)."""
    from xca_amd import models
    from xca_amd.dlr import DLRGrid
    beta = float(fx["beta"])
    grid = DLRGrid.from_nodes(beta, fx["it"], fx["rf"], lam=float(fx["lam"]), eps=float(fx["eps"]))
    ad = tca3_atom()
    G0, eta0 = ad.atomic_G0(beta, grid.tau)
    P = models.ImpurityProblem(grid=grid, ad=ad, poles=grid.rf / beta, coeffs=np.array(fx["coeffs"]), G0=G0, eta0=eta0,
                               name="tca3_extended")
    G = P.unpack(np.array(fx["G"]))
    return P, G


def main():
    import pyoracle as po
    from xca_amd import models
    from xca_amd.diag import all_connected_pairings
    ap = argparse.ArgumentParser()
    ap.add_argument("--per-topology", type=int, default=384)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--runs", type=int, default=16, help="contiguous runs of each flat-index list (one oracle call each)")
    a = ap.parse_args()
    P0 = models.tca_3orb()
    G = P0.perturbed_G()
    data = dict(beta=P0.beta, lam=P0.grid.lam, eps=P0.grid.eps, it=P0.grid.it, rf=P0.grid.rf, coeffs=P0.coeffs, G=P0.pack(G),
                dims=np.asarray(P0.dims, dtype=np.int32), runs=a.runs)
    fx = dict(data)
    P, G = problem_from_fixture(fx)
    rng = np.random.default_rng(a.seed)
    jobs = [(3, topo) for _, topo in all_connected_pairings(3)] + [(1, [(0, 1)]), (2, [(0, 2), (1, 3)])]
    for k, (order, topo) in enumerate(jobs):
        t = np.array(topo, dtype=np.int32)
        nmax = po.OracleEvaluator(P, complex_arith=False).num_backbones(order)
        if order == 3:
            fix = np.sort(rng.choice(nmax, size=a.per_topology, replace=False)).astype(np.int32)
        else:  # orders 1 and 2 in full
            fix = np.arange(nmax, dtype=np.int32)
        t0 = time.perf_counter()
        s = po.sigma_threads(P, G, t, fix, threads=a.threads or None, extended=True, runs=a.runs)
        print(f"topology {topo}: {len(fix)} flat indices in {time.perf_counter() - t0:.1f} s", flush=True)
        data[f"topo{k}"] = t
        data[f"fix{k}"] = fix
        data[f"sigma{k}"] = P.pack(s)
    np.savez_compressed(OUT, **data)
    print(f"wrote {OUT} ({OUT.stat().st_size / 1e6:.2f} MB)")


if __name__ == "__main__":
    main()
