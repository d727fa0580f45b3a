"""Builds the native parts in-tree.

  pt_amd/lib/libpthost.so  host scene surface (g++)
  pt_amd/lib/libpthip.so   MI355X back end: HIP kernels + C ABI (hipcc, gfx950)
  pt_amd/bin/pt            headless CLI (the reference's `pt --headless ...`)
  oracle/liboracle.so      CPU oracle (test infrastructure; make -C oracle)

Usage: python -m pt_amd.build [--no-hip] [--no-oracle] [-j N]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "pt_amd")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
INC = os.path.join(ROOT, "include")

HOST_SRCS = ["mathx.cpp", "json.cpp", "import_steps.cpp", "scene.cpp", "hdrio.cpp", "png.cpp", "jpeg.cpp", "api.cpp"]
HIP_SRCS = ["bvh.hip", "trace.hip", "cache.hip", "lights.hip", "post.hip", "api.hip"]
HIP_HDRS = ["common.h", "device.h", "bvh.h", "shading.h", "cache.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-I" + INC,
    "-I" + os.path.join(PKG, "csrc"),
    # packed fp32 math stays off: hipcc -O3 would otherwise emit v_pk_* fp32 ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


# Per-file additions. The path tracer's float divisions and square roots (shading, traversal,
# accumulation) use the hardware's approximations (a few ulp) instead of the correctly rounded
# instruction sequences (about ten instructions each); its parity tests carry tolerances for
# that. The light tables (lights.hip) keep correctly rounded division: they must equal the
# host packing bit for bit.
HIP_FILE_FLAGS = {"trace.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt"]}


def run(cmd, cwd=ROOT):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def build_host(force=True):
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(PKG, "host", s) for s in HOST_SRCS]
    out = os.path.join(LIB, "libpthost.so")
    deps = srcs + [os.path.join(PKG, "host", h) for h in os.listdir(os.path.join(PKG, "host")) if h.endswith(".h")]
    if force or _newer(out, deps):
        run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-I" + INC, "-o", out + ".tmp"] + srcs)
        os.replace(out + ".tmp", out)
    return out


def build_hip(jobs=4, force=True):
    os.makedirs(LIB, exist_ok=True)
    obj_dir = os.path.join(PKG, "lib", "obj")
    os.makedirs(obj_dir, exist_ok=True)
    csrc = os.path.join(PKG, "csrc")
    hdrs = [os.path.join(csrc, h) for h in HIP_HDRS] + [os.path.join(INC, "pthip.h")]
    objs, cmds = [], []
    for s in HIP_SRCS:
        src = os.path.join(csrc, s)
        obj = os.path.join(obj_dir, s.replace(".hip", ".o"))
        objs.append(obj)
        if force or _newer(obj, [src] + hdrs):
            cmds.append([HIPCC] + HIP_FLAGS + HIP_FILE_FLAGS.get(s, []) + ["-c", src, "-o", obj])
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for f in [ex.submit(run, c) for c in cmds]:
            f.result()
    out = os.path.join(LIB, "libpthip.so")
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)
    return out


def build_cli(with_hip=True):
    os.makedirs(BIN, exist_ok=True)
    out = os.path.join(BIN, "pt")
    src = os.path.join(PKG, "host", "main.cpp")
    libs = ["-L" + LIB, "-lpthost"] + (["-lpthip"] if with_hip else [])
    run(["g++", "-O2", "-std=c++17", "-I" + INC, "-o", out, src] + libs + ["-Wl,-rpath,$ORIGIN/../lib"])
    return out


def build_oracle():
    run(["make", "-C", os.path.join(ROOT, "oracle")])
    if os.path.isdir("/root/reference/src"):
        run(["make", "-C", os.path.join(ROOT, "oracle"), "ref"])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-hip", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("-j", type=int, default=4)
    a = ap.parse_args(argv)
    build_host(force=not a.incremental)
    if not a.no_hip:
        build_hip(a.j, force=not a.incremental)
    build_cli(with_hip=not a.no_hip)
    if not a.no_oracle:
        build_oracle()


if __name__ == "__main__":
    sys.exit(main())
