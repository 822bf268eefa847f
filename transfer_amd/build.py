"""Builds libtransfer_amd.so in-tree: host C++ with g++, HIP with hipcc for
gfx950, linked by hipcc.  Run as ``python transfer_amd/build.py`` (it does not import the package)."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtransfer_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TR_OFFLOAD_ARCH", "gfx950")

CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
# Packed FP32 instructions are disabled for every HIP translation unit.
HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build(verbose=False, incremental=False):
    """Compiles every source (incremental: only those whose object is older
    than the source or than any header) and links the library."""
    os.makedirs(OBJ, exist_ok=True)
    headers = [os.path.join(CSRC, n) for n in os.listdir(CSRC) if n.endswith(".h")]
    headers.append(os.path.join(ROOT, "include", "transfer.h"))
    newest_header = max(os.path.getmtime(h) for h in headers)
    jobs, objs = [], []
    for name in sorted(os.listdir(CSRC)):
        src = os.path.join(CSRC, name)
        base, ext = os.path.splitext(name)
        obj = os.path.join(OBJ, base + ".o")
        if ext not in (".cpp", ".hip"):
            continue
        objs.append(obj)
        if incremental and os.path.exists(obj) and os.path.getmtime(obj) > max(os.path.getmtime(src), newest_header):
            continue
        if ext == ".cpp":
            jobs.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
        else:
            jobs.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
    if jobs:
        with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
            for r in ex.map(_run, jobs):
                if verbose and (r.stdout or r.stderr):
                    sys.stderr.write(r.stdout + r.stderr)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))
