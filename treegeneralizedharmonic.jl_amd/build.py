"""Build libtgh_mi355x.so for gfx950 with hipcc (in-tree, next to this file).

Used by __graft_entry__.build(); runnable directly: python build.py.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libtgh_mi355x.so")
# The API and pointwise kernels, the right-hand side at each order and the
# adjoint, each in a translation unit of its own: the six compile in parallel.
SOURCES = [os.path.join(HERE, "csrc", f) for f in
           ("tgh_api.hip", "tgh_q2.hip", "tgh_q4.hip", "tgh_q6.hip", "tgh_q8.hip", "tgh_adjoint.hip")]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def build(verbose=False, defines=(), out=None):
    """Compile every translation unit (in parallel) and link the library.
    defines: extra -D tuning macros; out: another library path (a variant)."""
    lib = out or LIB
    objdir = os.path.join(os.path.dirname(lib), "_build", os.path.basename(lib))
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [hipcc()] + HIPCC_FLAGS + [f"-D{d}" for d in defines] + ["-c", "-o", obj, src]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = []
    for src, pr in procs:
        out, _ = pr.communicate()
        if pr.returncode != 0:
            sys.stderr.write(out)
            failed.append(os.path.basename(src))
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)} building {lib}")
    tmp = lib + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"linking {lib} failed ({res.returncode})")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
