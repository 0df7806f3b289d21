"""Builds libpgn.so (HIP kernels for gfx950 + the C++ host driver + C ABI).

Every source is compiled with hipcc for gfx950 only, with packed FP32 vector
ops disabled, then linked into physim_amd/libpgn.so.  Run as
`python -m physim_amd.build` or through `__graft_entry__.build()`.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libpgn.so")
OBJ = os.path.join(HERE, "_build")

SOURCES = ["rod_kernels.hip", "tet_kernels.hip", "shell_kernels.hip", "assembly_kernels.hip",
           "pcg_kernels.hip", "batch_kernels.hip", "model.cpp", "solver.cpp", "api.cpp",
           "rod_model.cpp", "rod_api.cpp", "batch.cpp", "server.cpp", "ldlt_kernels.hip",
           "ldlt.cpp", "csr_solve.hip", "matfree_kernels.hip", "scene.cpp", "coarse_kernels.hip", "coarse.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(OBJ, src.rsplit(".", 1)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    return src, obj, p.returncode, msgs


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        results = list(ex.map(_compile, SOURCES))
    failed = False
    for src, _, rc, msgs in results:
        if msgs and (verbose or rc):
            print(f"[{src}]\n{msgs}", file=sys.stderr)
        failed |= rc != 0
    if failed:
        raise RuntimeError("libpgn compilation failed")
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + \
        [r[1] for r in results]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("libpgn link failed:\n" + p.stdout + p.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
