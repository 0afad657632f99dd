"""Decode the reference's HDF5 test fixtures into small JSON files (data only).

  test/c++/h5/tpz100.ref.h5            OCA_tpz_result (101, 16, 16) complex: trapezoidal-quadrature OCA
                                       self-energy of the two-band discrete-bath model on 101 equispaced
                                       tau points (used by test_block_sparse_OCA_manual.cpp:237-282)
  test/python/nessi_ppsc_one_spinful_fermion.ref.h5
                                       NESSi-ppsc atom/NCA/OCA single-particle g(tau) and Delta(tau) on
                                       801 points (used by one_spinful_fermion.py:122-244)

Only the diagonal (N-conserving) blocks of OCA_tpz_result are kept, real parts; the largest imaginary
part is recorded. Run from the repository root with the reference checked out at /root/reference and
h5dump on PATH or at /opt/conda/bin/h5dump.
"""
import json
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def read_dataset(path, name):
    txt = subprocess.run([H5DUMP, "-d", name, "-m", "%.17g", "-y", "-w", "0", str(path)], check=True, capture_output=True,
                         text=True).stdout
    shape = tuple(int(x) for x in re.search(r"DATASPACE\s+SIMPLE\s+\{\s*\(([^)]*)\)", txt).group(1).split(","))
    body = txt[txt.index("DATA {") + len("DATA {"):]
    body = body[:body.rindex("}")]
    body = body[:body.rindex("}")] if body.count("}") else body
    vals = np.array([float(x) for x in re.findall(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?|nan|inf", body)])
    return vals[: int(np.prod(shape))].reshape(shape)


def tpz():
    a = read_dataset(REF / "test/c++/h5/tpz100.ref.h5", "/OCA_tpz_result")
    z = a[..., 0] + 1j * a[..., 1]
    ranges = [(0, 4), (4, 10), (10, 11), (11, 15), (15, 16)]  # test_block_sparse_OCA_manual.cpp:264-281
    blocks = [z[:, lo:hi, lo:hi] for lo, hi in ranges]
    off = z.copy()
    for lo, hi in ranges:
        off[:, lo:hi, lo:hi] = 0
    out = {"source": "test/c++/h5/tpz100.ref.h5:/OCA_tpz_result", "ntau": z.shape[0], "block_ranges": ranges,
           "max_abs_imag": float(np.abs(z.imag).max()), "max_abs_offblock": float(np.abs(off).max()),
           "blocks_real": [b.real.tolist() for b in blocks]}
    (OUT / "oca_tpz100_blocks.json").write_text(json.dumps(out))


def nessi():
    f = REF / "test/python/nessi_ppsc_one_spinful_fermion.ref.h5"
    out = {"source": "test/python/nessi_ppsc_one_spinful_fermion.ref.h5"}
    for grp in ("atom", "nca", "oca"):
        out[grp] = {k: read_dataset(f, f"/{grp}/{k}").tolist() for k in ("tau", "g_tau", "delta_tau")}
    (OUT / "nessi_one_spinful_fermion.json").write_text(json.dumps(out))


if __name__ == "__main__":
    tpz()
    nessi()
