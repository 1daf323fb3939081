"""Builds the library's sources for the CPU emulation (tests/simt/hip/hip_runtime.h) with g++,
into tests/simt/_build/libndhip_simt.so.

Two edits are applied to copies of the sources: the AMDGPU occupancy attributes are removed
(g++ does not parse them), and the cap on active hybrid mix-ins can be set (--max-mixins), so
that cycles the GPU build rejects can be run here. UBSan checks array bounds, float-to-int
overflow and the other undefined behaviour g++ instruments, and aborts on the first finding.
"""
import argparse
import concurrent.futures as cf
import os
import re
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "numericdream_amd", "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libndhip_simt.so")
ATTR = re.compile(r"__attribute__\(\(amdgpu_waves_per_eu\((?:[^()]|\([^()]*\))*\)\)\)")


def build(max_mixins=3, jobs=8):
    src = os.path.join(OUT_DIR, "src")
    obj = os.path.join(OUT_DIR, "obj")
    shutil.rmtree(OUT_DIR, ignore_errors=True)
    os.makedirs(src)
    os.makedirs(obj)
    for f in os.listdir(CSRC):
        if not f.endswith((".hip", ".cpp", ".h", ".hpp", ".inc")):
            continue
        text = open(os.path.join(CSRC, f)).read()
        text = ATTR.sub("", text)
        if f == "nd_dev.h":
            text, n = re.subn(r"#define ND_HIP_MAX_ACTIVE_MIXINS \d+", "#define ND_HIP_MAX_ACTIVE_MIXINS %d" % max_mixins, text)
            assert n == 1, "mix-in cap not found in nd_dev.h"
        open(os.path.join(src, f), "w").write(text)
    flags = ["-std=c++17", "-O1", "-g", "-fPIC", "-w", "-I" + HERE, "-I" + os.path.join(ROOT, "include"), "-I" + src,
             "-fsanitize=undefined,float-cast-overflow,bounds", "-fno-sanitize-recover=all"]
    units = [os.path.join(src, f) for f in sorted(os.listdir(src)) if f.endswith((".hip", ".cpp"))]
    units.append(os.path.join(HERE, "emu.cpp"))

    def cc(u):
        o = os.path.join(obj, os.path.basename(u) + ".o")
        r = subprocess.run(["g++"] + flags + ["-x", "c++", "-c", u, "-o", o], capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError("g++ failed on %s:\n%s" % (u, r.stderr[-4000:]))
        return o

    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(cc, units))
    r = subprocess.run(["g++", "-shared"] + flags + ["-Wl,-Bsymbolic"] + objs + ["-o", LIB], capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--max-mixins", type=int, default=3)
    print(build(ap.parse_args().max_mixins))
