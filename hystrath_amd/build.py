"""Builds libhystrath_dsmc.so (gfx950) in-tree with hipcc.

Usage: python -m hystrath_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhystrath_dsmc.so")
SOURCES = ["dsmc_engine.hip", "dsmc_kernels.hip", "ws_kernels.hip"]

# -ffp-contract=off: every a*b+c is two IEEE roundings (bit-reproducible FP64)
# -packed-fp32-ops off: no v_pk_*_f32 generated
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


HOST = os.path.join(HERE, "host")
BIN = os.path.join(HERE, "bin")
HOST_LIB_SOURCES = ["dictionary.C", "lagrangianIO.C", "polyMesh.C", "dsmcCloud.C"]
APPS = ["dsmcFoamPlus", "dsmcInitialisePlus"]


def build_host(verbose=False):
    """The C++ host mirror of dsmcCloud and its applications (g++, linked
    against the engine library)."""
    os.makedirs(BIN, exist_ok=True)
    common = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(HERE, "..", "include")]
    srcs = [os.path.join(HOST, s) for s in HOST_LIB_SOURCES]
    outs = []
    for app in APPS:
        out = os.path.join(BIN, app)
        cmd = common + srcs + [os.path.join(HOST, app + ".C"), "-L", HERE, "-lhystrath_dsmc",
                               "-Wl,-rpath,$ORIGIN/..", "-o", out]
        outs.append((cmd, out))
    # the tests' own tools for the host's readers (no device needed)
    tools = os.path.join(HERE, "..", "tests", "tools")
    tbin = os.path.join(tools, "_bin")
    os.makedirs(tbin, exist_ok=True)
    inc = ["-I", HOST]
    outs.append((common + inc + [os.path.join(HOST, "dictionary.C"), os.path.join(tools, "dictCheck.C"), "-o",
                                 os.path.join(tbin, "dictCheck")], os.path.join(tbin, "dictCheck")))
    outs.append((common + inc + [os.path.join(HOST, "dictionary.C"), os.path.join(HOST, "lagrangianIO.C"),
                                 os.path.join(tools, "cloudCheck.C"), "-o", os.path.join(tbin, "cloudCheck")],
                 os.path.join(tbin, "cloudCheck")))
    outs.append((common + inc + [os.path.join(HOST, "dictionary.C"), os.path.join(HOST, "polyMesh.C"),
                                 os.path.join(tools, "meshCheck.C"), "-o", os.path.join(tbin, "meshCheck")],
                 os.path.join(tbin, "meshCheck")))
    for cmd, out in outs:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {out}:\n{r.stdout}{r.stderr}")
        if verbose and (r.stdout or r.stderr):
            print(r.stdout + r.stderr)
    return [o for _, o in outs]


def build(verbose=False):
    cmd = [hipcc()] + HIPCC_FLAGS + [os.path.join(CSRC, s) for s in SOURCES] + ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l and l.strip()]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(LIB + ".tmp", LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
