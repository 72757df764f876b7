"""Builds libshipfem.so for gfx950 (hipcc, in-tree) and the CPU oracle.

    python sinking-simulator_amd/build.py            # library + oracle (+ oracle/_ref
                                                     # when the reference tree exists)
Every HIP translation unit is compiled with packed-FP32 instructions disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "lib", "obj")
LIB = os.path.join(LIBDIR, "libshipfem.so")

SOURCES = ["kernels.hip", "mesh.cpp", "prep.cpp", "engine.cpp", "api.cpp", "server.cpp", "handoff.cpp", "cpu_step.cpp", "timestep.hip",
           "casefile.cpp", "hex.cpp", "fem_gpu.cpp", "solid_shell.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=on: a multiply-add fuses only within one source expression, so every
# template instantiation of a kernel rounds identically (results do not depend on the
# launch shape or on which other cases share a launch).
COMMON = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC", "-Wall", "-ffp-contract=on",
          "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    # the host pass does not know the packed-fp32 target feature; that warning is noise
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
    return err


def build_library(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    jobs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        jobs.append((obj, [HIPCC] + COMMON + lang + ["-c", path, "-o", obj]))
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for out in ex.map(lambda j: _run(j[1]), jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [j[0] for j in jobs]
         + ["-lpthread"])
    os.replace(tmp, LIB)
    # the fem_spike driver, linked against the library (rpath to lib/)
    tool = os.path.join(LIBDIR, "fem_spike")
    _run([HIPCC] + COMMON + ["-x", "hip", os.path.join(HERE, "tools", "fem_spike.cpp"), "-o",
                             tool + ".tmp", "-L", LIBDIR, "-lshipfem", "-Wl,-rpath,$ORIGIN"])
    os.replace(tool + ".tmp", tool)
    return LIB


def build_oracle():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def write_build_info():
    """lib/build_info.json: the commit the library was built from (the tree travels
    without its git history), whether the tree had changes, and when."""
    import datetime
    import json

    def git(*args):
        r = subprocess.run(["git", "-C", ROOT] + list(args), capture_output=True, text=True)
        return r.stdout.strip() if r.returncode == 0 else None

    info = dict(commit=git("rev-parse", "HEAD") or "unknown",
                dirty=bool(git("status", "--porcelain", "--untracked-files=no")),
                built_utc=datetime.datetime.now(datetime.timezone.utc).isoformat(timespec="seconds"))
    with open(os.path.join(LIBDIR, "build_info.json"), "w") as f:
        json.dump(info, f, indent=1)


def build_spike():
    """The reference's own fem_spike driver switched over to this library (oracle/Makefile
    target `spike`; only where the reference tree is present)."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "spike"], capture_output=True,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError("fem_spike switch-over build failed:\n" + r.stdout + r.stderr)


def build_all(verbose=False):
    build_oracle()
    lib = build_library(verbose)
    write_build_info()
    build_spike()
    return lib


if __name__ == "__main__":
    print(build_all(verbose="-v" in sys.argv))
