"""Build libwslda_td.so (HIP, gfx950) in-tree with hipcc.

Each .hip translation unit is compiled to an object in parallel, then linked
into wslda_amd/libwslda_td.so.  No CMake, no JIT cache: the library sits next
to this file so it travels with the source tree.

A project build (``python -m wslda_amd.build --project DIR --codedim D``)
compiles a problem's predefines.h and problem-definition.h (in DIR) into
csrc/project/tdw_project.hip and links it with the same engine objects into
DIR/libwslda_td.so: the problem's v_ext, delta_ext, velocity_ext,
scattering_length, modify_* callbacks then drive the engine, as the
reference compiles them into its kernels.
"""
import os
import subprocess
import sys
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwslda_td.so")
ARCH = os.environ.get("TDW_OFFLOAD_ARCH", "gfx950")
UNITS = ["tdw_api.hip", "tdw_fft.hip", "tdw_physics.hip", "tdw_fused.hip", "tdw_adjoint.hip", "tdw_aslda.hip", "tdw_grid.hip", "tdw_hooks.hip", "tdw_noalloc.hip", "tdw_uniform.cpp", "tdw_checkpoint.cpp", "tdw_wdata.cpp", "tdw_stio.cpp"]

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


def _objects():
    return [os.path.join(OBJ, os.path.splitext(u)[0] + ".o") for u in UNITS]


def build_project(project_dir, codedim, out=None, verbose=False):
    """Compile the problem in project_dir (predefines.h, problem-definition.h)
    for the td code of dimension codedim and link it with the engine objects
    (built by build() first) into out (default project_dir/libwslda_td.so)."""
    project_dir = os.path.abspath(project_dir)
    if int(codedim) not in (1, 2, 3):
        raise ValueError("codedim must be 1, 2 or 3")
    for f in ("predefines.h", "problem-definition.h"):
        if not os.path.exists(os.path.join(project_dir, f)):
            raise FileNotFoundError(os.path.join(project_dir, f))
    objs = _objects()
    if not all(os.path.exists(o) for o in objs):
        build(verbose=verbose)
    out = out or os.path.join(project_dir, "libwslda_td.so")
    pobj = os.path.join(OBJ, "tdw_project_%d_%08x.o" % (int(codedim), zlib.crc32(project_dir.encode())))
    cmd = [hipcc(), "-c", os.path.join(SRC, "project", "tdw_project.hip"), "-o", pobj] + FLAGS + \
        ["-I" + project_dir, "-I" + os.path.join(SRC, "project"), "-DCODEDIM=%d" % int(codedim),
         "-Wno-unused-but-set-variable", "-Wno-unused-parameter"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("hipcc failed for the project in " + project_dir)
    tmp = out + ".tmp"
    subprocess.run([hipcc(), "-shared", "-o", tmp, "--offload-arch=" + ARCH] + objs + [pobj], check=True)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--project", help="directory with predefines.h and problem-definition.h")
    ap.add_argument("--codedim", type=int, default=3)
    ap.add_argument("--out")
    a = ap.parse_args()
    if a.project:
        print(build_project(a.project, a.codedim, a.out, verbose=a.v))
    else:
        print(build(verbose=a.v))
