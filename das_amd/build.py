"""Builds libtswrtm_hip.so (the HIP engine) in-tree with hipcc for gfx950.

    python -m das_amd.build             # or das_amd.build.build()
    python -m das_amd.build --sanitize  # or das_amd.build.build_sanitized()

The second form builds, for testing only, the same library with its host
code instrumented by AddressSanitizer and UndefinedBehaviorSanitizer (the
device code is built as usual) into das_amd/_san/, together with the driver
tests/cpp/host_san.cpp linked against it.  The sanitizer runtime is linked
into that executable; the library resolves its hooks from there.
"""
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtswrtm_hip.so")
SAN = os.path.join(PKG, "_san")
SAN_LIB = os.path.join(SAN, "libtswrtm_hip_san.so")
SAN_EXE = os.path.join(SAN, "host_san")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["acoustic.hip", "elastic.hip", "post.hip", "engine.hip", "batch.hip", "elabatch.hip", "serve.hip", "fwi.hip", "vgather.hip", "models.hip", "output.hip", "dsp.hip", "ts_a.hip", "ts_b.hip"]

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # the kernels follow the reference's FP32 association; no FMA contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


# host-only instrumentation: each option reaches the host compiler and linker
SAN_FLAGS = [f for opt in ("-fsanitize=address", "-fsanitize=undefined",
                           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g")
             for f in ("-Xarch_host", opt)]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _build_lib(obj_dir, lib, extra, verbose):
    os.makedirs(obj_dir, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        for r in ex.map(_run, cmds):
            if verbose and r.stderr.strip():
                print(r.stderr, file=sys.stderr)
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lhipfft"])
    os.replace(tmp, lib)
    return lib


def build(verbose=False):
    return _build_lib(OBJ, LIB, [], verbose)


def build_sanitized(verbose=False):
    """The sanitizer-instrumented library and its driver (test use only)."""
    _build_lib(os.path.join(SAN, "obj"), SAN_LIB, SAN_FLAGS, verbose)
    tmp = SAN_EXE + ".tmp"
    _run([HIPCC, *FLAGS, *SAN_FLAGS, os.path.join(ROOT, "tests", "cpp", "host_san.cpp"),
          "-o", tmp, "-L", SAN, "-ltswrtm_hip_san", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, SAN_EXE)
    return SAN_EXE


if __name__ == "__main__":
    if "--sanitize" in sys.argv[1:]:
        print(build_sanitized(verbose=True))
    else:
        print(build(verbose=True))
