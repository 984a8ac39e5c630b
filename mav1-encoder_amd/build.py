"""Builds libmav1hip.so in-tree: the gfx950 kernels and the C-ABI layer
(hipcc), plus the host-side parameter code (gcc). Sources are compiled in
parallel; every call rebuilds everything."""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libmav1hip.so")

HIP_SOURCES = ["api.hip", "me.hip", "analysis.hip", "recon.hip", "filters.hip", "lrsearch.hip", "noalloc.hip", "seq.hip", "pruner.hip"]
# sources whose floating-point results must round exactly as written (no FMA contraction)
STRICT_FP = {"lrsearch.hip"}
C_SOURCES = ["params.c", "fwdmats.c", "y4m.c"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off (the kernels are integer; this also
    # keeps the compiler from forming v_pk_* FP32 ops anywhere)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + p.stdout)
    return p.stdout


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(BUILD, s.replace(".hip", ".o"))
        extra = ["-ffp-contract=off"] if s in STRICT_FP else []
        jobs.append(([HIPCC] + HIP_FLAGS + extra + ["-c", os.path.join(CSRC, s), "-o", o], o))
    for s in C_SOURCES:
        o = os.path.join(BUILD, s.replace(".c", ".o"))
        jobs.append((["gcc", "-O2", "-fPIC", "-std=gnu11", "-Wall", "-ffp-contract=off", "-c", os.path.join(CSRC, s),
                      "-o", o], o))
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        outs = list(ex.map(lambda j: _run(j[0]), jobs))
    if verbose:
        for o in outs:
            if o.strip():
                sys.stdout.write(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for _, o in jobs] + ["-lm"])
    os.replace(tmp, LIB)
    return LIB


def build_oracle(with_ref=True):
    """The oracle's C restatement, and (when the reference sources are
    present) the reference's kernels into oracle/_ref."""
    root = os.path.dirname(PKG)
    _run(["make", "-s", "-B", "-C", os.path.join(root, "oracle"), "liboracle.so"])
    if with_ref and os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-B", "-C", os.path.join(root, "oracle"), "ref"])


if __name__ == "__main__":
    print(build(verbose=True))
    build_oracle()
