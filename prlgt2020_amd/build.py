"""Build libnspt.so (HIP, gfx950) in-tree.

    python -m prlgt2020_amd.build [--force] [-j N]

Each kernel instance is one translation unit (nspt_inst.hip compiled with
-DINST_NC/-DINST_N/-DINST_QT/-DINST_LPW/-DINST_WPE, once more with
-DINST_BIG=1 for fields past 2 GiB)
so they compile in parallel; the host side
(nspt_capi.hip, nspt_rng.cpp) is linked with them into
prlgt2020_amd/libnspt.so.  The instance list mirrors NSPT_INSTANCES of
csrc/nspt_internal.hpp.
"""
import argparse
import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnspt.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NSPT_ARCH", "gfx950")

# packed-FP32 VALU forms are disabled for all device code (FP64 path anyway).
HIPFLAGS = ["-O3", "--offload-arch=" + ARCH, "-fPIC", "-std=c++17",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

HEADERS = ["nspt_device.hpp", "nspt_kernels.hpp", "nspt_internal.hpp"]
HOST = os.path.join(HERE, "host")
BIN = os.path.join(HERE, "bin")
# Quenched drivers built by default: (NC, allocORD); any other pair with
# python -m prlgt2020_amd.build --driver NC ALLOCORD
DRIVERS = [(3, 6), (3, 8), (3, 20), (2, 6)]


def instances():
    text = open(os.path.join(CSRC, "nspt_internal.hpp")).read()
    # the full list is the last definition (the first is a one-instance
    # variant for quick experiments)
    block = text[text.rindex("#define NSPT_INSTANCES"):]
    block = block[:block.index("\n\n")]
    return [tuple(int(v) for v in m)
            for m in re.findall(r"X\((\d+),\s*(\d+),\s*(\d+),\s*(\d+),\s*(\d+)\)", block)]


def _inst_obj(nc, n, q, lpw, wpe, big):
    return os.path.join(OBJ, "inst_%d_%d_q%d_l%d_w%d%s.o" % (nc, n, q, lpw, wpe, "_big" if big else ""))


def _newer(target, sources):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(s) > t for s in sources)


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(force=False, jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    hdrs = [os.path.join(CSRC, h) for h in HEADERS] + [os.path.join(HERE, "..", "include", "nspt.h")]
    jobs = jobs or min(8, os.cpu_count() or 4)
    tasks = []
    for nc, n, q, lpw, wpe in instances():
        src = os.path.join(CSRC, "nspt_inst.hip")
        for big in (0, 1):
            obj = _inst_obj(nc, n, q, lpw, wpe, big)
            if force or _newer(obj, [src] + hdrs):
                tasks.append([HIPCC] + HIPFLAGS + ["-DINST_NC=%d" % nc, "-DINST_N=%d" % n, "-DINST_QT=%d" % q,
                                                   "-DINST_LPW=%d" % lpw, "-DINST_WPE=%d" % wpe,
                                                   "-DINST_BIG=%d" % big, "-c", src, "-o", obj])
    capi_obj = os.path.join(OBJ, "nspt_capi.o")
    if force or _newer(capi_obj, [os.path.join(CSRC, "nspt_capi.hip")] + hdrs):
        tasks.append([HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, "nspt_capi.hip"), "-o", capi_obj])
    adj_obj = os.path.join(OBJ, "nspt_adjoint.o")
    if force or _newer(adj_obj, [os.path.join(CSRC, "nspt_adjoint.hip")] + hdrs):
        tasks.append([HIPCC] + HIPFLAGS + ["-c", os.path.join(CSRC, "nspt_adjoint.hip"), "-o", adj_obj])
    rng_obj = os.path.join(OBJ, "nspt_rng.o")
    if force or _newer(rng_obj, [os.path.join(CSRC, "nspt_rng.cpp")] + hdrs):
        tasks.append(["g++", "-O2", "-fPIC", "-std=c++17", "-c", os.path.join(CSRC, "nspt_rng.cpp"), "-o", rng_obj])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for r in ex.map(_run, tasks):
            if verbose and r.stderr:
                print(r.stderr, file=sys.stderr)
    objs = [_inst_obj(*inst, big) for inst in instances() for big in (0, 1)] + [capi_obj, adj_obj, rng_obj]
    relinked = False
    if force or tasks or _newer(LIB, objs):
        _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs)
        os.replace(LIB + ".tmp", LIB)
        relinked = True
    for nc, ao in DRIVERS:
        build_driver(nc, ao, force=force or relinked)
    return LIB


def driver_path(nc, alloc_ord):
    return os.path.join(BIN, "quenched_su%d_a%d" % (nc, alloc_ord))


def build_driver(nc, alloc_ord, force=False):
    """The Quenched driver (host/quenched.cc + host/nspt_host.cc) for one
    gauge group and allocORD, linked against libnspt.so."""
    os.makedirs(BIN, exist_ok=True)
    out = driver_path(nc, alloc_ord)
    srcs = [os.path.join(HOST, "quenched.cc"), os.path.join(HOST, "nspt_host.cc")]
    deps = srcs + [os.path.join(HOST, "nspt_host.h"), os.path.join(HERE, "..", "include", "nspt.h")]
    if force or _newer(out, deps):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-DNC=%d" % nc, "-DallocORD=%d" % alloc_ord] + srcs +
             ["-L" + HERE, "-lnspt", "-Wl,-rpath,$ORIGIN/..", "-o", out])
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--driver", nargs=2, type=int, metavar=("NC", "ALLOCORD"), help="build one more driver")
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.j, verbose=True))
    if a.driver:
        print(build_driver(a.driver[0], a.driver[1], force=True))
