"""Builds libmvr_hip.so for gfx950 with hipcc.

The FFT passes are templates on the transform length; this script writes one
small translation unit per group of lengths (so hipcc runs in parallel), a
dispatch table that maps a runtime length to its launchers, and links
everything with the host code into ``multiview-reconstruction_amd/lib/libmvr_hip.so``.
"""
from __future__ import annotations

import concurrent.futures as cf
import hashlib
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libmvr_hip.so")
BUILD = os.path.join(HERE, "_build")
GEN = os.path.join(BUILD, "gen")

# FFT lengths with compiled kernels.  Strided (y/z) lengths L; the x pass of a
# length-P axis runs a length-P/2 complex transform, so P/2 is compiled for every P.
STRIDED = [16, 20, 24, 32, 40, 48, 60, 64, 72, 80, 96, 100, 120, 128, 144, 160, 180, 192, 200, 216, 240, 256, 270,
           288, 300, 320, 360, 384, 400, 432, 450, 480, 512, 540, 576, 600, 640, 720, 768, 800, 864, 900, 960, 1024,
           1080, 1152, 1280]
HALF = sorted({p // 2 for p in STRIDED})

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function", "-Wno-unused-result"]

X_FUNCS = """template void launch_x_fwd<{M}>(const Dom&, const GatherParams&, float2*, const float2*, const float2*, hipStream_t);
template void launch_x_ratio<{M}>(const Dom&, const RatioParams&, const float2*, float2*, const float2*, const float2*, hipStream_t);
template int launch_x_update<{M}>(const Dom&, const UpdateParams&, const float2*, const float2*, const float2*, hipStream_t);
template void launch_x_store<{M}>(const Dom&, const StoreParams&, const float2*, const float2*, const float2*, hipStream_t);
"""
S_FUNCS = """template void launch_strided<{L}>(const StridedGeom&, int, const float2*, float2*, const float2*, float, const float2*, hipStream_t);
template int strided_span<{L}>();
"""


def _groups(seq, n):
    k = max(1, (len(seq) + n - 1) // n)
    return [seq[i:i + k] for i in range(0, len(seq), k)]


def generate():
    os.makedirs(GEN, exist_ok=True)
    files = []
    for gi, grp in enumerate(_groups(HALF, 10)):
        body = "".join(X_FUNCS.format(M=m) for m in grp)
        files.append((f"x_{gi}.hip", f'#include "xpasses.hpp"\nnamespace mvr {{\n{body}}}\n'))
    for gi, grp in enumerate(_groups(STRIDED, 10)):
        body = "".join(S_FUNCS.format(L=l) for l in grp)
        files.append((f"s_{gi}.hip", f'#include "passes.hpp"\nnamespace mvr {{\n{body}}}\n'))
    # dispatch table (host only: declarations of the explicit instantiations)
    d = ['#include "engine.hpp"', "namespace mvr {",
         "template <int M> void launch_x_fwd(const Dom&, const GatherParams&, float2*, const float2*, const float2*, hipStream_t);",
         "template <int M> void launch_x_ratio(const Dom&, const RatioParams&, const float2*, float2*, const float2*, const float2*, hipStream_t);",
         "template <int M> int launch_x_update(const Dom&, const UpdateParams&, const float2*, const float2*, const float2*, hipStream_t);",
         "template <int M> void launch_x_store(const Dom&, const StoreParams&, const float2*, const float2*, const float2*, hipStream_t);",
         "template <int L> void launch_strided(const StridedGeom&, int, const float2*, float2*, const float2*, float, const float2*, hipStream_t);",
         "template <int L> int strided_span();"]
    for name, fn, typ in [("get_x_fwd", "launch_x_fwd", "XFwdFn"), ("get_x_ratio", "launch_x_ratio", "XRatioFn"),
                          ("get_x_update", "launch_x_update", "XUpdateFn"), ("get_x_store", "launch_x_store", "XStoreFn")]:
        cases = " ".join(f"case {m}: return &{fn}<{m}>;" for m in HALF)
        d.append(f"{typ} {name}(int M) {{ switch (M) {{ {cases} default: return nullptr; }} }}")
    cases = " ".join(f"case {l}: return &launch_strided<{l}>;" for l in STRIDED)
    d.append(f"StridedFn get_strided(int L) {{ switch (L) {{ {cases} default: return nullptr; }} }}")
    cases = " ".join(f"case {l}: return strided_span<{l}>();" for l in STRIDED)
    d.append(f"int get_strided_span(int L) {{ switch (L) {{ {cases} default: return 0; }} }}")
    d.append("static const int kStrided[] = {" + ", ".join(map(str, STRIDED)) + ", 0};")
    d.append("static const int kHalf[] = {" + ", ".join(map(str, HALF)) + ", 0};")
    d.append("const int* strided_lengths() { return kStrided; }")
    d.append("const int* half_lengths() { return kHalf; }")
    d.append("}  // namespace mvr")
    files.append(("dispatch.cpp", "\n".join(d) + "\n"))
    paths = []
    for name, text in files:
        p = os.path.join(GEN, name)
        old = open(p).read() if os.path.exists(p) else None
        if old != text:
            with open(p, "w") as f:
                f.write(text)
        paths.append(p)
    return paths


def _deps_hash(src):
    h = hashlib.sha1()
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".hpp", ".h")):
            h.update(open(os.path.join(CSRC, f), "rb").read())
    h.update(open(os.path.join(HERE, "..", "include", "mvr_hip.h"), "rb").read())
    h.update(open(src, "rb").read())
    h.update(" ".join(FLAGS).encode())
    return h.hexdigest()[:16]


def _compile(src, force):
    obj = os.path.join(BUILD, "obj", os.path.basename(src) + ".o")
    stamp = obj + ".hash"
    key = _deps_hash(src)
    if not force and os.path.exists(obj) and os.path.exists(stamp) and open(stamp).read() == key:
        return obj, None
    os.makedirs(os.path.dirname(obj), exist_ok=True)
    cmd = [HIPCC] + FLAGS + ["-I", CSRC, "-c", src, "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, f"{' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    os.replace(obj + ".tmp", obj)
    with open(stamp, "w") as f:
        f.write(key)
    return obj, None


def build(force: bool = False, jobs: int | None = None, verbose: bool = True) -> str:
    srcs = generate()
    srcs += [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith((".hip", ".cpp"))]
    jobs = jobs or min(16, os.cpu_count() or 4)
    objs, errs = [], []
    with cf.ThreadPoolExecutor(jobs) as ex:
        for obj, err in ex.map(lambda s: _compile(s, force), srcs):
            if err:
                errs.append(err)
            else:
                objs.append(obj)
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n\n".join(errs[:3]))
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} ({len(objs)} objects)")
    return LIB


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build(force="--force" in sys.argv)
