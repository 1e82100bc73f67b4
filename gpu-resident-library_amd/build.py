"""Builds libgpufhe_amd.so for gfx950 from csrc/ with hipcc (in-tree, so the
library travels with the source tree).  Usage: python build.py [-j N] [--incremental]"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpufhe_amd.so")
ARCH = os.environ.get("GPUFHE_ARCH", "gfx950")
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         # packed FP32 instructions stay off in every code object of this library
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["kernels_ntt_f.hip", "kernels_ntt_i.hip", "kernels_ksi.hip", "kernels_modup.hip", "kernels_ntt_ref.hip", "kernels.hip", "engine.hip", "capi.hip", "serve.hip", "params.cpp", "api.cpp", "keys.cpp", "wire.cpp"]


def _stale(src, obj):
    """object missing or older than its source or any header of csrc/ and include/"""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [os.path.join(SRC, src)] + [os.path.join(d, f) for d in (SRC, os.path.join(HERE, "..", "include"))
                                       for f in os.listdir(d) if f.endswith((".hpp", ".h"))]
    return any(os.path.getmtime(p) > t for p in deps)


def _compile(src, incremental=False):
    obj = os.path.join(OUT, src + ".o")
    if incremental and not _stale(src, obj):
        return obj, ""
    cmd = ["hipcc"] + FLAGS + (["-x", "hip"] if src.endswith(".hip") else []) + \
          ["-c", os.path.join(SRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, msg))
    return obj, msg


def build(jobs=8, verbose=False, incremental=False):
    """Compiles every source (incremental: only those whose object is stale)
    and links the library."""
    os.makedirs(OUT, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(SRC, s))]
    with cf.ThreadPoolExecutor(jobs) as ex:
        res = list(ex.map(lambda s: _compile(s, incremental), srcs))
    for _, msg in res:
        if msg.strip() and verbose:
            print(msg)
    tmp = LIB + ".tmp"
    cmd = ["hipcc", "--offload-arch=" + ARCH, "-shared", "-o", tmp] + [o for o, _ in res]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    j = int(sys.argv[sys.argv.index("-j") + 1]) if "-j" in sys.argv else 8
    print(build(j, verbose=True, incremental="--incremental" in sys.argv))
