"""Build the gfx950 tracking library in-tree.

    python -m elegant_amd.build

Produces elegant_amd/libelegant_hip.so from elegant_amd/csrc/eh_lib.hip.  All
HIP code is compiled without floating-point contraction (-ffp-contract=off):
the maps must reproduce the CPU routines' IEEE-754 double results exactly.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libelegant_hip.so")
ARCH = os.environ.get("EH_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + CSRC,
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; the HIP toolchain is required to build elegant_amd")
    return exe


def generated_headers():
    """Regenerate the constant tables from their generators."""
    for gen, out in (("gen_math_consts.py", "eh_math_consts.h"), ("gen_csbend_terms.py", "eh_csbend_terms.h")):
        text = subprocess.run([sys.executable, os.path.join(CSRC, gen)], check=True, capture_output=True,
                              text=True).stdout
        path = os.path.join(CSRC, out)
        old = open(path).read() if os.path.exists(path) else None
        if old != text:
            with open(path, "w") as f:
                f.write(text)


def build(verbose=False):
    generated_headers()
    src = os.path.join(CSRC, "eh_lib.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building %s" % src)
    if verbose:
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
