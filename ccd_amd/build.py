"""Build libccdhip.so (all HIP/C++ sources of the engine) for gfx950, in-tree.

    python -m ccd_amd.build          # or ccd_amd.build.build()
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libccdhip.so")
BUILD_DIR = os.path.join(ROOT, "_build")

SOURCES = [
    "ccd_engine.hip",
    "ccd_batch.hip",
    "ccd_ingest.hip",
    "ccd_tables.hip",
    "ccd_cv.cpp",
    "ccd_serve.cpp",
    "engine/HipModelSpecifics.cpp",
    "engine/HostCyclicCoordinateDescent.cpp",
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CCDH_ARCH", "gfx950")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-pthread", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-function", "-Wno-unused-result",
          f"-I{os.path.join(ROOT, 'include')}"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
    return err


def build(verbose=False, jobs=8):
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD_DIR, src.replace("/", "_") + ".o")
        cmd = [HIPCC] + COMMON + (["-x", "hip"] if src.endswith(".cpp") else []) + \
            ["-c", path, "-o", obj]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                            text=True)))
        objs.append(obj)
    for cmd, p in procs:
        out, err = p.communicate()
        err = "\n".join(ln for ln in err.splitlines() if "packed-fp32-ops" not in ln)
        if p.returncode != 0:
            raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
        if verbose and err.strip():
            print(err)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
