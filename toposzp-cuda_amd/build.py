"""Builds libtoposzp_hip.so (gfx950) in place with hipcc.

    python toposzp-cuda_amd/build.py [--jobs N]

Every HIP source is compiled with -ffp-contract=off (the byte stream depends on
exact float/double rounding) and with packed FP32 instructions disabled.
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtoposzp_hip.so")
ARCH = os.environ.get("TSZP_OFFLOAD_ARCH", "gfx950")

SOURCES = ["tszp_codec.hip", "tszp_sort.hip", "tszp_topo.hip", "tszp_util.hip", "tszp_gen.hip",
           "tszp_post.hip", "tszp_binrank.hip", "tszp_parse.hip", "tszp_quality.hip", "tszp_elem.hip",
           "tszp_saddle.hip", "tszp_noalloc.hip", "tszp_topo_batch.hip", "tszp_platmc.hip", "tszp_api.cpp", "tszp_serve.cpp",
           "tszp_batch_api.cpp", "tszp_topo_serve.cpp"]
FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    cand = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return cand if os.path.exists(cand) else "hipcc"


def compile_one(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), f"--offload-arch={ARCH}", *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".cpp"):  # host code only
        cmd = [hipcc(), f"--offload-arch={ARCH}", "--cuda-host-only", *FLAGS[:5], "-c",
               os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{msgs}")
    return out, msgs


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    if os.path.exists(LIB):  # a failed build must not leave an older library behind
        os.remove(LIB)
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    for _, msgs in results:
        if verbose and msgs.strip():
            print(msgs)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
