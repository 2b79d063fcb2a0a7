"""Build recipe: compiles the HIP library and the C++ host for gfx950.

    python -m thesis_amd.build        # liblbm.so + lbmsolver

Every .hip/.cpp file under csrc/ goes into one shared library with a plain C
ABI (include/lbm.h); host/lbmsolver.cpp links against it.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblbm.so")
CLI = os.path.join(HERE, "lbmsolver")
LATTICEGEN = os.path.join(HERE, "latticegen")
ARCH = os.environ.get("LBM_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          # packed FP32 instructions are kept out of every kernel
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-Wno-unused-command-line-argument"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib(jobs=None):
    objdir = os.path.join(ROOT, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", src, "-o", obj]
        print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    bad = [p.wait() for p in procs]
    if any(bad):
        raise SystemExit("compilation failed")
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)


def build_cli():
    for name, out in (("lbmsolver.cpp", CLI), ("latticegen.cpp", LATTICEGEN)):
        src = os.path.join(HERE, "host", name)
        if os.path.exists(src):
            _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-o", out,
                  "-L", HERE, "-llbm", "-Wl,-rpath,$ORIGIN"])


def build_calib():
    """Copy-bandwidth calibration program (counter scale factors)."""
    src = os.path.join(HERE, "calib", "copy_bandwidth.hip")
    if os.path.exists(src):
        _run([HIPCC, "-O3", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              src, "-o", os.path.join(HERE, "calib", "copy_bandwidth")])


def build():
    build_lib()
    build_cli()
    build_calib()


if __name__ == "__main__":
    build()
    sys.exit(0)
