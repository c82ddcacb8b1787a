"""Build the native engine libm2.so for gfx950 (in-tree, no JIT cache).

Steps: run the ERI code generator (csrc/gen_eri.py), compile every HIP/C++
translation unit with hipcc in parallel, link molecular2_amd/lib/libm2.so.

    python -m molecular2_amd.build [--force] [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
GEN = os.path.join(CSRC, "generated")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIBM2 = os.path.join(LIBDIR, "libm2.so")
LIBM2_TORCH = os.path.join(LIBDIR, "libm2_torch.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: keep the compiler from forming packed FP32 VALU ops.
# -munsafe-fp-atomics: FP64 atomicAdd lowers to global_atomic_add_f64.
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + CSRC, "-I" + os.path.join(CSRC, "host"), "-I" + INCLUDE, "-Wno-unused-result"]


def sources():
    srcs = sorted(glob.glob(os.path.join(GEN, "*.hip")))
    srcs += sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    srcs += sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    srcs += sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))
    return srcs


def _headers():
    hs = glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(GEN, "*.inc")) + \
        glob.glob(os.path.join(CSRC, "host", "*.hpp")) + \
        glob.glob(os.path.join(INCLUDE, "*.h"))
    return max((os.path.getmtime(h) for h in hs), default=0.0)


def _obj(src):
    base = os.path.relpath(src, CSRC).replace(os.sep, "_")
    return os.path.join(OBJDIR, base + ".o")


def _compile(src, force, hdr_time):
    obj = _obj(src)
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(src), hdr_time):
        return obj, 0.0, ""
    t = time.time()
    lang = ["-x", "hip"] if src.endswith(".hip") or src.endswith(".cpp") else []
    cmd = [HIPCC] + HIPFLAGS + lang + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    return obj, time.time() - t, ""


def build(force: bool = False, jobs: int = 0, verbose: bool = True) -> str:
    sys.path.insert(0, CSRC)
    try:
        import gen_eri
    finally:
        sys.path.pop(0)
    gen_eri.main(GEN)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sources()
    hdr = _headers()
    jobs = jobs or min(16, os.cpu_count() or 4)
    objs = []
    t0 = time.time()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, force, hdr): s for s in srcs}
        for f in cf.as_completed(futs):
            obj, dt, _ = f.result()
            objs.append(obj)
            if verbose and dt > 0:
                print("  compiled %-28s %6.1fs" % (os.path.basename(futs[f]), dt), flush=True)
    objs.sort()
    newest = max(os.path.getmtime(o) for o in objs)
    if force or not os.path.exists(LIBM2) or os.path.getmtime(LIBM2) < newest:
        cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIBM2 + ".tmp"] + objs + \
            ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr[-6000:])
        os.replace(LIBM2 + ".tmp", LIBM2)
    if verbose:
        print("libm2.so ready (%.1fs)" % (time.time() - t0), flush=True)
    build_torch_op(force=force, verbose=verbose)
    return LIBM2


def build_torch_op(force: bool = False, verbose: bool = True) -> str:
    """libm2_torch.so: the libtorch autograd operator torch.ops.m2.rhf_energy
    (csrc/autograd/rhf_energy.cpp), host C++ over libm2.so's C ABI."""
    import torch
    import torch.utils.cpp_extension as E
    src = os.path.join(CSRC, "autograd", "rhf_energy.cpp")
    if not force and os.path.exists(LIBM2_TORCH) and \
            os.path.getmtime(LIBM2_TORCH) >= max(os.path.getmtime(src), os.path.getmtime(LIBM2), _headers()):
        return LIBM2_TORCH
    t = time.time()
    tlib = E.library_paths()[0]
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
           "-D_GLIBCXX_USE_CXX11_ABI=%d" % int(torch._C._GLIBCXX_USE_CXX11_ABI), "-I" + INCLUDE] + \
        ["-I" + p for p in E.include_paths()] + \
        [src, "-o", LIBM2_TORCH + ".tmp", "-L" + tlib, "-ltorch", "-ltorch_cpu", "-lc10", "-L" + LIBDIR, "-lm2",
         "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + tlib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("g++ failed for %s:\n%s" % (src, r.stderr[-6000:]))
    os.replace(LIBM2_TORCH + ".tmp", LIBM2_TORCH)
    if verbose:
        print("  compiled %-28s %6.1fs" % ("rhf_energy.cpp", time.time() - t), flush=True)
    return LIBM2_TORCH


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--jobs", type=int, default=0)
    a = ap.parse_args()
    build(force=a.force, jobs=a.jobs)
