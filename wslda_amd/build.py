"""Build libwslda_td.so (HIP, gfx950) in-tree with hipcc.

Each .hip translation unit is compiled to an object in parallel, then linked
into wslda_amd/libwslda_td.so.  No CMake, no JIT cache: the library sits next
to this file so it travels with the source tree.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwslda_td.so")
ARCH = os.environ.get("TDW_OFFLOAD_ARCH", "gfx950")
UNITS = ["tdw_api.hip", "tdw_fft.hip", "tdw_physics.hip", "tdw_fused.hip", "tdw_adjoint.hip", "tdw_aslda.hip", "tdw_grid.hip", "tdw_uniform.cpp", "tdw_checkpoint.cpp", "tdw_wdata.cpp", "tdw_stio.cpp"]

FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-fno-gpu-flush-denormals-to-zero",
         "-I" + SRC, "-I" + os.path.join(ROOT, "include"),
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result"]


def hipcc():
    for p in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(p) or p == "hipcc":
            return p


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    procs = []
    objs = []
    for u in UNITS:
        src = os.path.join(SRC, u)
        obj = os.path.join(OBJ, os.path.splitext(u)[0] + ".o")
        objs.append(obj)
        cmd = [hipcc(), "-c", src, "-o", obj] + FLAGS
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((u, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for u, p in procs:
        out, _ = p.communicate()
        txt = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((u, txt))
        elif verbose and txt.strip():
            print(txt)
    if failed:
        for u, txt in failed:
            sys.stderr.write("=== %s ===\n%s\n" % (u, txt))
        raise RuntimeError("hipcc failed for " + ", ".join(u for u, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-o", tmp, "--offload-arch=" + ARCH] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
