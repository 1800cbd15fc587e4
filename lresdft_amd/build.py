"""Build liblres_hip.so in-tree with hipcc for gfx950 (no torch extension machinery).

Every translation unit is compiled to an object in parallel, then linked.  The
MFMA pole kernel (csrc/lres_mfma.hip) is compiled once per (precision, tiles
per dimension) part.  Usage: python -m lresdft_amd.build   (or
__graft_entry__.build())."""
import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/lres_kernels.hip", "csrc/lres_plan.hip", "csrc/lres_forces.hip"]
MFMA_SOURCE = "csrc/lres_mfma.hip"
MFMA_PARTS = [("float", nt) for nt in range(2, 17)] + [("double", nt) for nt in range(2, 9)]
LIB = os.path.join(HERE, "liblres_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         # packed FP32 VALU forms are kept out of the device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def build(verbose=False, jobs=None):
    """Compile every HIP source of the package into liblres_hip.so (always rebuilds)."""
    inc = ["-I", os.path.join(os.path.dirname(HERE), "include")]
    tmpdir = tempfile.mkdtemp(prefix="lres_build_")
    try:
        units = [(os.path.join(HERE, s), os.path.join(tmpdir, os.path.basename(s) + ".o"), [])
                 for s in SOURCES]
        units += [(os.path.join(HERE, MFMA_SOURCE), os.path.join(tmpdir, f"lres_mfma_{r}_{nt}.o"),
                   # explicit fma only: every variant of a part rounds identically
                   [f"-DLRES_MFMA_R={r}", f"-DLRES_MFMA_NT={nt}", "-ffp-contract=off"]) for r, nt in MFMA_PARTS]
        cmds = [[HIPCC, *FLAGS, *inc, *defs, "-c", src, "-o", obj] for src, obj, defs in units]
        if verbose:
            for c in cmds:
                print(" ".join(c), flush=True)
        jobs = jobs or max(1, min(16, os.cpu_count() or 1))
        with ThreadPoolExecutor(jobs) as ex:
            list(ex.map(_run, cmds))
        tmp = LIB + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[u[1] for u in units]])
        os.replace(tmp, LIB)
    finally:
        shutil.rmtree(tmpdir, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
