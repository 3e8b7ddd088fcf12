"""AddressSanitizer + UndefinedBehaviorSanitizer run of the host code on the
CPU: the library's host side (configuration parsing, table construction,
transforms, pixelization, stream seeding and jump-ahead: everything that
runs without a device) and the oracle, built with clang's sanitizers and
exercised by the CPU tests in one process with the sanitizer runtime
preloaded.

    python -m galfast_amd.sanitize [pytest args]

builds galfast_amd/_build_asan/libgalfast_hip_asan.so (device code as usual,
host code instrumented) and oracle/_build/libgforacle_asan.so, then runs
the CPU tests against them.  Leak checking is off: the interpreter's own
allocations are not this code's.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

from . import build as B

OBJ = os.path.join(B.HERE, "_build_asan")
LIB = os.path.join(OBJ, "libgalfast_hip_asan.so")
ROOT = os.path.dirname(B.HERE)
# host-side instrumentation only: each sanitizer flag applies to the host pass
HOST_SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer", "-g"]


def runtime():
    """clang's ASan runtime (it carries the UBSan handlers too)."""
    c = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    if not c:
        raise RuntimeError("clang ASan runtime not found under /opt/rocm/llvm")
    return c[0]


def _compile(src):
    rel = os.path.relpath(src, B.CSRC).replace(os.sep, "_")
    obj = os.path.join(OBJ, rel + ".o")
    cmd = [B.HIPCC] + B.COMMON + B.DEVICE + HOST_SAN + ["-c", src, "-o", obj]
    return src, obj, subprocess.run(cmd, capture_output=True, text=True)


def build():
    os.makedirs(OBJ, exist_ok=True)
    srcs = B.sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        res = list(ex.map(_compile, srcs))
    bad = [(s, r.stderr) for s, _, r in res if r.returncode != 0]
    if bad:
        raise RuntimeError("sanitizer build failed: %s\n%s" % bad[0])
    cmd = [B.HIPCC, "-shared", "-fPIC", "--offload-arch=" + B.ARCH, "-Xarch_host", "-fsanitize=address",
           "-Xarch_host", "-fsanitize=undefined", "-shared-libsan", "-o", LIB] + [o for _, o, _ in res]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("sanitizer link failed:\n" + r.stderr)
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle sanitizer build failed:\n" + r.stderr)
    return LIB


def run(pytest_args):
    build()
    env = dict(os.environ)
    env["LD_PRELOAD"] = runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["GALFAST_LIB"] = LIB
    env["GFO_LIB"] = os.path.join(ROOT, "oracle", "_build", "libgforacle_asan.so")
    args = pytest_args or ["tests", "-q", "-m", "not gpu", "-p", "no:cacheprovider"]
    return subprocess.run([sys.executable, "-m", "pytest"] + args, cwd=ROOT, env=env).returncode


if __name__ == "__main__":
    sys.exit(run(sys.argv[1:]))
