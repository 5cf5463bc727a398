"""Builds libgpucad.so (HIP kernels + C-ABI) for gfx950 in-tree.

Every source under csrc/ is compiled by hipcc to an object file (in parallel),
then linked into gpucadforam_amd/libgpucad.so.  Flags:
  --offload-arch=gfx950          MI355X only
  -ffp-contract=off              no implicit a*b+c fusion: every fused
                                 multiply-add in the kernels is an explicit fmaf,
                                 which keeps results reproducible by a host
                                 build of the same arithmetic
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                 no packed FP32 (v_pk_*_f32) instructions
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpucad.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(HERE, "..", "include"),
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src, extra):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, "--offload-arch=" + ARCH] + COMMON + extra + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr, cmd


def build(verbose=False, extra=None):
    extra = list(extra or [])
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        for src, obj, rc, log, cmd in ex.map(lambda s: _compile(s, extra), srcs):
            if rc != 0:
                sys.stderr.write("compile failed: %s\n%s\n%s\n" % (src, " ".join(cmd), log))
                raise RuntimeError("hipcc failed for %s" % src)
            if verbose and log.strip():
                sys.stderr.write(log)
            objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
