"""Build recipe of libchameleon_amd.so (gfx950).

Every source under csrc/ is compiled with hipcc for gfx950 and linked into
one shared library next to this file.  Packed-FP32 instruction selection is
turned off for all device code (-packed-fp32-ops removed from the target
features).  Sources are compiled in parallel; the link always runs.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libchameleon_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", f"-I{os.path.join(ROOT, 'include')}", f"-I{SRC}",
          "-Wall", "-Wno-unused-function", "-Wno-unused-result", "-Wno-unused-variable"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def sources():
    out = []
    for dirpath, _, files in os.walk(SRC):
        for f in sorted(files):
            if f.endswith((".cpp", ".hip")):
                out.append(os.path.join(dirpath, f))
    return sorted(out)


def compile_one(src):
    rel = os.path.relpath(src, SRC).replace(os.sep, "_")
    obj = os.path.join(OBJ, rel + ".o")
    cmd = [HIPCC] + COMMON + DEVICE + ["-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l and "unknown target feature" not in l and l.strip())
    return src, obj, r.returncode, msgs


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(compile_one, srcs))
    failed = False
    for src, obj, rc, msgs in results:
        if rc != 0:
            failed = True
            print(f"--- {os.path.relpath(src, ROOT)} failed:\n{msgs}", file=sys.stderr)
        elif verbose and msgs:
            print(f"--- {os.path.relpath(src, ROOT)}:\n{msgs}", file=sys.stderr)
    if failed:
        raise RuntimeError("libchameleon_amd: compilation failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + [o for _, o, _, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr, file=sys.stderr)
        raise RuntimeError("libchameleon_amd: link failed")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
