"""Build libq1tsim.so (HIP for gfx950 + the C ABI) in-tree with hipcc.

``python -m q1tsim_amd.build`` compiles every source under csrc/ into
q1tsim_amd/_build/ and links q1tsim_amd/libq1tsim.so.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libq1tsim.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
          "-I", os.path.join(os.path.dirname(HERE), "include")]
# Packed FP32 (v_pk_*_f32) stays off for all device code.  Sinking of
# common code out of switch cases is off: it merges the per-register-slot
# cases of the sweep kernels into one body with a run-time index into the
# register array, which puts that array in scratch memory.  Uniform regions
# are left unstructurized: the sweep kernels dispatch on wave-uniform op
# fields, and structurizing those branches copies the whole register array
# at every join.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics", "-mllvm", "-simplifycfg-sink-common=false",
          "-mllvm", "-structurizecfg-skip-uniform-regions"]

SOURCES = ["gvs.hip", "items.hip", "sweep.hip", "costmodel.hip", "gates.cpp", "plan.cpp", "circuit.cpp", "batch.cpp",
           "capi.cpp", "jit.cpp"]

# Device headers the run-time specialised sweep kernels include (jit.cpp
# hands them to hipRTC as text).
JIT_HEADERS = ["kernels.h", "sweep_types.h", "sweep_dev.h"]


def write_jit_text():
    """_build/jit_text.h: the JIT_HEADERS as raw string literals."""
    out = ["// Written by build.py: device headers for the run-time sweep kernels.", "#pragma once", ""]
    names, texts = [], []
    for h in JIT_HEADERS:
        text = open(os.path.join(CSRC, h)).read()
        if ")q1tjit\"" in text:
            raise RuntimeError("raw-string delimiter found in " + h)
        names.append('"%s"' % h)
        texts.append('R"q1tjit(' + text + ')q1tjit"')
    out.append("static const char* const kJitHeaderNames[] = {" + ", ".join(names) + "};")
    out.append("static const char* const kJitHeaderTexts[] = {\n" + ",\n".join(texts) + "};")
    path = os.path.join(BUILD, "jit_text.h")
    new = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != new:
        with open(path, "w") as f:
            f.write(new)


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    if src.endswith(".hip"):
        # the sweep kernels spell out their fused multiply-adds and contract
        # nothing else, like the kernels specialised at run time (jit.cpp)
        extra = ["-ffp-contract=off"] if src == "sweep.hip" else []
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + extra + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC, "-x", "c++"] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-I", "/opt/rocm/include", "-I", BUILD,
                                              "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    write_jit_text()
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
