"""Build the MI355X NUFHT libraries in-tree.

Produces, next to this file:
  libnufht_amd.so                       the C-ABI library (include/*.h)
  _pynufht.cpython-*.so                 the drop-in CPython extension

Device code is compiled for gfx950 only, with packed-FP32 instructions
disabled and hardware FP64 atomics enabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-std=c++17", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             # matrix-core accumulators in ordinary VGPRs: without it the
             # compiler copies them through AGPRs around every MFMA of a loop
             "-mllvm", "-amdgpu-mfma-vgpr-form",
             "-Wno-unused-result", "-Wno-unused-value"]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function"]


def _run(cmd, quiet):
    if not quiet:
        print(" ".join(cmd))
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)


def lib_path() -> str:
    return os.path.join(HERE, "libnufht_amd.so")


def ext_path() -> str:
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "_pynufht" + suffix)


def _jobs() -> int:
    n = os.environ.get("MAX_JOBS")
    n = int(n) if n and n.isdigit() else (os.cpu_count() or 1)
    return max(1, min(16, n))


# device translation units: (source, object, extra defines); the asymptotic
# stage is compiled once per grid precision and exponent bound so that its
# kernel instantiations build in parallel
def _hip_units():
    units = [("engine.hip", "engine.o", []), ("scratch.hip", "scratch.o", [])]
    for rname, tag in (("double", "f64"), ("float", "f32")):
        units.append(("fft_stage.hip", f"fft_{tag}.o", [f"-DNUFHT_ASY_R={rname}"]))
        for em in (6, 10, 14, 22):
            units.append(("asy_stage.hip", f"asy_{tag}_e{em}.o",
                          [f"-DNUFHT_ASY_R={rname}", f"-DNUFHT_ASY_EM={em}"]))
    return units


def build(quiet: bool = True) -> None:
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(BUILD, exist_ok=True)
    hipcc = os.path.join(ROCM, "bin", "hipcc")
    incs = ["-I", INC, "-I", CSRC]
    cmds, objs = [], []
    for src in ("plan.c", "tables.c"):
        obj = os.path.join(BUILD, src.replace(".c", ".o"))
        cmds.append(["gcc", *C_FLAGS, *incs, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src, obj, defs in _hip_units():
        obj = os.path.join(BUILD, obj)
        cmds.append([hipcc, *HIP_FLAGS, *defs, *incs, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=_jobs()) as ex:
        for f in [ex.submit(_run, c, quiet) for c in cmds]:
            f.result()
    tmp = lib_path() + ".tmp"
    _run([hipcc, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs,
          "-L", os.path.join(ROCM, "lib"), "-lrocfft", "-lamdhip64",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")], quiet)
    os.replace(tmp, lib_path())
    import numpy
    pyinc = sysconfig.get_paths()["include"]
    tmp = ext_path() + ".tmp"
    _run(["gcc", *C_FLAGS, "-shared", "-I", pyinc, "-I", numpy.get_include(), *incs,
          os.path.join(CSRC, "pynufht.c"), "-o", tmp, "-L", HERE, "-lnufht_amd",
          "-Wl,-rpath,$ORIGIN"], quiet)
    os.replace(tmp, ext_path())


if __name__ == "__main__":
    build(quiet="-v" not in sys.argv)
    print("built", lib_path(), ext_path())
