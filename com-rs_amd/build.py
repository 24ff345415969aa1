"""Build the HIP library com-rs_amd/_build/libcomfec_gpu.so for gfx950 (in-tree).

hipcc cross-compiles without a GPU. Packed FP32 instructions are disabled for every HIP
translation unit, and FMA contraction is off so that the f64 kernels reproduce the
reference's f64 arithmetic bit for bit.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libcomfec_gpu.so")
# the microbenchmarks of the HBM, LDS and VALU peaks (bench.py only; not part of the library)
PEAKS_LIB = os.path.join(OUT_DIR, "libcomfec_peaks.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SOURCES = ["nr_code.cpp", "capi.cpp", "sim.cpp", "sim_report.cpp", "campaign.cpp", "hooks_api.cpp", "learned_api.cpp", "batch.cpp", "decode_generic.hip", "decode_layered.hip", "decode_layered_f32.hip", "decode_layered_f16.hip", "trials.hip", "learned.hip", "static_scratch.hip"]
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def commit():
    """The commit the library is built from, as crates/com-sim/build.rs records it: the hash,
    with "-dirty" when the library's tracked sources have uncommitted changes, or "unknown"
    outside a Git checkout."""
    def git(*args):
        r = subprocess.run(["git", "-C", HERE, *args], capture_output=True, text=True)
        return r.stdout.strip() if r.returncode == 0 else None
    top = git("rev-parse", "--show-toplevel")
    head = top and git("rev-parse", "HEAD")
    if not head:
        return "unknown"
    status = git("--no-optional-locks", "status", "--porcelain", "--untracked-files=no", "--",
                 os.path.relpath(HERE, top), "include")
    return head + "-dirty" if status else head


def _compile(src, defines=()):
    obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, *defines, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed on " + src)
    noise = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                      if "packed-fp32-ops" not in l and l.strip())
    if noise:
        sys.stderr.write(noise + "\n")
    return obj


def build(verbose=True):
    os.makedirs(OUT_DIR, exist_ok=True)
    defines = {"sim_report.cpp": ['-DCOMFEC_COMMIT="%s"' % commit()]}
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(lambda s: _compile(s, defines.get(s, ())), SOURCES + ["peaks.hip"]))
    for lib, parts in ((LIB, objs[:-1]), (PEAKS_LIB, objs[-1:])):
        tmp = lib + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *parts, "-lpthread"]
        subprocess.run(cmd, check=True)
        os.replace(tmp, lib)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
