"""Builds libalsvinn_hip.so (the hip platform: gfx950 kernels + C ABI) in-tree.

Every HIP translation unit is compiled with -ffp-contract=off so that FP64
results reproduce the reference CPU path's roundings exactly, and with the
packed-FP32 target feature disabled.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libalsvinn_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# (source, object, extra defines): the equation units are compiled once per
# state precision, which halves the longest compile and runs the two halves
# in parallel; the longest jobs are listed first.
SOURCES = []
for _eq in ("euler3", "euler2", "burgers", "euler1"):
    SOURCES.append(("kernels_%s.hip" % _eq, "kernels_%s_f64" % _eq, ["-DALSVINN_HIP_FP64_ONLY"]))
    SOURCES.append(("kernels_%s.hip" % _eq, "kernels_%s_f32" % _eq, ["-DALSVINN_HIP_FP32_ONLY"]))
SOURCES += [("ensemble.hip", "ensemble", []), ("operators.hip", "operators", []),
            ("stats.hip", "stats", [])]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(job):
    src, name, defines = job
    obj = os.path.join(BUILD, name + ".o")
    cmd = [HIPCC] + FLAGS + defines + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(jobs=None, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, len(SOURCES))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
