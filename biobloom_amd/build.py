"""Builds libbbt_hip.so (gfx950 kernels + C ABI) in-tree with hipcc, and the
C++ front end.  Run as `python -m biobloom_amd.build`."""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "libbbt_hip.so")
CLI = os.path.join(PKG, "biobloomcategorizer")
MAKER = os.path.join(PKG, "biobloommaker")
MIMAKER = os.path.join(PKG, "biobloommimaker")
MICAT = os.path.join(PKG, "biobloommicategorizer")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ARCH = os.environ.get("BBT_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib():
    srcs = [os.path.join(CSRC, "bbt_hip.hip"), os.path.join(CSRC, "bbt_mibf.hip"),
            os.path.join(CSRC, "bbt_serve.cpp"), os.path.join(CSRC, "bbt_host.cpp")]
    tmp = LIB + ".tmp"
    _run([HIPCC, *DEVICE_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, LIB)


def build_cli():
    """The restated biobloomcategorizer / biobloommaker / biobloommimaker /
    biobloommicategorizer front ends."""
    for name, out in [("biobloomcategorizer.cpp", CLI), ("biobloommaker.cpp", MAKER),
                      ("biobloommimaker.cpp", MIMAKER),
                      ("biobloommicategorizer.cpp", MICAT)]:
        src = os.path.join(HOST, name)
        if not os.path.exists(src):
            continue
        _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-Wall",
              "-o", out, src, os.path.join(HOST, "front_end.cpp"),
              os.path.join(HOST, "seq_ingest.cpp"), "-pthread",
              "-Wl,-rpath,$ORIGIN", "-L", PKG, "-lbbt_hip", "-lz", "-ldl"])


def build():
    build_lib()
    build_cli()


if __name__ == "__main__":
    build()
    sys.exit(0)
