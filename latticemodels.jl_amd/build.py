"""Build liblmhip.so (HIP C++ for gfx950) in-tree.

    python latticemodels.jl_amd/build.py

Compiles every csrc/*.hip translation unit with hipcc for gfx950 and links
them into latticemodels.jl_amd/lmhip/liblmhip.so, next to the Python host
package that loads it.  rocSOLVER / rocBLAS are linked for the dense
eigensolver behind the Fermi projector.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lmhip")
LIB = os.path.join(OUT_DIR, "liblmhip.so")
OBJ_DIR = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# -packed-fp32-ops: no v_pk_*_f32 instructions are generated.
COMMON = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(HERE, "..", "include"),
]
LINK_LIBS = ["-L" + os.path.join(ROCM, "lib"), "-lrocsolver", "-lrocblas",
             "-Wl,-rpath," + os.path.join(ROCM, "lib")]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=True, jobs=None):
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), 8)
    objs = []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for src, obj, rc, out in ex.map(_compile, srcs):
            if rc != 0:
                raise RuntimeError("hipcc failed for %s:\n%s" % (src, out))
            if verbose and out.strip():
                sys.stderr.write(out)
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + LINK_LIBS
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % (r.stdout + r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
