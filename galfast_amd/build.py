"""Build the gfx950 shared library ``galfast_amd/libgalfast_hip.so``.

Every HIP source is compiled for gfx950 with hipcc; objects are linked into
one shared library that exports the C ABI declared in
``include/galfast_hip.h``.  FP contraction is disabled (the kernels spell out
every multiply-add they want, so the device arithmetic matches a CPU
evaluation of the same steps) and packed FP32 instructions are turned off.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgalfast_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GALFAST_ARCH", "gfx950")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
    "-I" + os.path.join(HERE, "..", "include"), "-I" + CSRC,
]
DEVICE = [
    "--offload-arch=" + ARCH,
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fgpu-rdc" if False else "-fno-gpu-rdc",
]


def sources():
    out = []
    for root, _, files in os.walk(CSRC):
        for f in sorted(files):
            if f.endswith((".hip", ".cpp")):
                out.append(os.path.join(root, f))
    return sorted(out)


def _compile(src):
    rel = os.path.relpath(src, CSRC).replace(os.sep, "_")
    obj = os.path.join(OBJ, rel + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    objs = []
    failed = False
    for src, obj, r in results:
        if r.returncode != 0:
            failed = True
            sys.stderr.write("compile failed: %s\n%s\n" % (src, r.stderr))
        elif verbose and r.stderr.strip():
            warn = [l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()]
            if warn:
                sys.stderr.write("%s:\n%s\n" % (os.path.basename(src), "\n".join(warn[:40])))
        objs.append(obj)
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
