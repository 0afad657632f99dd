"""Build the native library xca_amd/_lib/libxca_amd.so (HIP kernels for gfx950 + host C++).

Run as ``python -m xca_amd.build`` or through ``__graft_entry__.build()``. hipcc cross-compiles for
gfx950 without a GPU present.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "csrc"
OUT = HERE / "_lib"
LIB = OUT / "libxca_amd.so"

ARCH = os.environ.get("XCA_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function"]
# packed FP32 instructions are kept out of every device code object
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-fno-gpu-rdc"]


def _sources():
    hip = sorted(SRC.glob("*.hip"))
    cpp = sorted(SRC.glob("*.cpp")) + sorted((SRC / "host").glob("*.cpp"))
    return hip, cpp


def _compile(src: Path, obj: Path, device: bool):
    cmd = [HIPCC] + COMMON + (DEVICE + ["-x", "hip"] if device else []) + ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OUT.mkdir(exist_ok=True)
    objdir = OUT / "obj"
    objdir.mkdir(exist_ok=True)
    hip, cpp = _sources()
    jobs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for s in hip:
            jobs.append(ex.submit(_compile, s, objdir / (s.stem + ".hip.o"), True))
        for s in cpp:
            jobs.append(ex.submit(_compile, s, objdir / (s.parent.name + "_" + s.stem + ".o"), False))
        objs = [j.result() for j in jobs]
    tmp = LIB.with_suffix(".so.tmp")
    # rocSOLVER / rocBLAS: LU factorisations and GEMMs of the Dyson systems too large for one workgroup
    cmd = ([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-Wl,-z,defs", "-o", str(tmp)] + [str(o) for o in objs]
           + ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
