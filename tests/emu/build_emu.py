"""Build the CPU emulation of libphrike_hip (tests only): the same kernel and
host sources compiled by g++ against tests/emu/hip/hip_runtime.h.

    python tests/emu/build_emu.py [--asan]
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "phrike_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libphrike_emu.so")
SOURCES = ["capi.hip", "inst_f64.hip", "inst_f32.hip", "hybrid.hip", "mhd.hip", "emu_stubs.cpp"]


def build(asan: bool = False, defines=(), tag: str = "") -> str:
    """The emulation library; `defines` (kernel compile-time options) go to
    a library of its own, named by `tag`."""
    # objects in a per-process folder and an atomic rename of the library, so
    # parallel test workers building at once never load a half-written file
    work = os.path.join(OUT, f"p{os.getpid()}")
    os.makedirs(work, exist_ok=True)
    flags = ["-std=c++20", "-O1", "-g", "-fPIC", "-I", HERE, "-Wno-unused-result"]
    # extra defines for tuning variants of the kernels, e.g. EMU_DEFINES="-DPH_XPADB=4"
    flags += os.environ.get("EMU_DEFINES", "").split()
    flags += list(defines)
    if asan:
        flags += ["-fsanitize=address", "-fno-omit-frame-pointer"]
    objs = []
    procs = []
    for s in SOURCES:
        o = os.path.join(work, os.path.splitext(s)[0] + ".o")
        objs.append(o)
        src = os.path.join(HERE, s) if s.endswith(".cpp") else os.path.join(CSRC, s)
        procs.append(subprocess.Popen(["g++", *flags, "-x", "c++", "-c", src, "-o", o],
                                      stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(out)
    lib = LIB.replace(".so", "_asan.so") if asan else LIB
    if tag:
        lib = lib.replace(".so", f"_{tag}.so")
    tmp = os.path.join(work, os.path.basename(lib))
    r = subprocess.run(["g++", *flags, "-shared", "-o", tmp, *objs, "-lpthread"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)
    os.replace(tmp, lib)
    shutil.rmtree(work, ignore_errors=True)
    return lib


if __name__ == "__main__":
    print(build("--asan" in sys.argv))
