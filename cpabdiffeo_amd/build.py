"""Build libcpab_hip.so in-tree for gfx950 with hipcc.

    python -m cpabdiffeo_amd.build

Flags: -ffp-contract=off keeps every a*b+c of the kernels as the separate
multiply and add the reference's source writes (the parity contract with the
CPU oracle); packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
SO = os.path.join(LIBDIR, "libcpab_hip.so")
SOURCES = [os.path.join(CSRC, "cpab_api.hip")]
HEADERS = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")] + [
    os.path.join(os.path.dirname(HERE), "include", "cpab_hip.h")]
ARCH = os.environ.get("CPAB_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def command(out=SO):
    return [hipcc(), "-O3", "-std=c++17", "--offload-arch=" + ARCH, "-fPIC", "-shared",
            "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-o", out] + SOURCES


def build(verbose=False):
    """Compile unconditionally (the build check must exercise the compiler)."""
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = SO + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed (%d):\n%s" % (r.returncode, out))
    os.replace(tmp, SO)
    return SO


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
