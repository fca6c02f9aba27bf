"""Calibration of the oracle's CPU speed against the reference's own code
(test infrastructure; writes oracle/calibration.json).

    OMP_NUM_THREADS=1 python oracle/calibrate.py

The reference's step cannot be built in this image (nspt.cc and Quenched.cc
include FFTW3, QCDenvNODEpt_test.h:33, which is absent), and its sources are
not where the GPU runs, so the CPU baseline of a GPU run times the oracle
and states this ratio beside it.  What can be compared is the work that
dominates both steps: the series algebra.  oracle/_ref is the reference's
own MyMath.cc / QCDpt.cc compiled from its sources (oracle/Makefile, the
flags of its Makefile:2); ref_su3_algebra_seconds and orc_algebra_seconds
time the same rounds of SU(3) series products, exponentials and logarithms
(the mix of one RK2 + zero-mode + FA link-step: 30 products, 3 exp, 1 log)
at the reference's compiled order count (allocORD = 6, choices.h:30) on one
core.  ratio = oracle rate / reference rate.
"""
import ctypes
import json
import os
import platform
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the script's own folder would make "oracle" name oracle.py, not the package
sys.path[:] = [os.path.dirname(HERE)] + [p for p in sys.path if os.path.abspath(p or ".") != HERE]
from oracle import oracle as O  # noqa: E402

MIX = {"products": (30, 0, 0), "exp": (0, 1, 0), "log": (0, 0, 1), "step mix": (30, 3, 1)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def series_pair(N, seed=3):
    """Two SU(3) link series of N orders (order 0 = 1) from the oracle's
    random field, as flat double arrays of 1 + 9N complex."""
    U = O.random_field((2, 2, 2, 2), 3, N, seed=seed, scale=0.3)
    return (np.ascontiguousarray(U[0], np.complex128).view(np.float64),
            np.ascontiguousarray(U[1], np.complex128).view(np.float64))


def main():
    O.build()
    ref = O.Ref()
    N = ref.A  # the reference's allocORD
    a, b = series_pair(N)
    dp = ctypes.POINTER(ctypes.c_double)
    rl = ref.l
    rl.ref_su3_algebra_seconds.restype = ctypes.c_double
    rl.ref_su3_algebra_seconds.argtypes = [dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
    ol = ctypes.CDLL(O.LIB_PATH)
    ol.orc_algebra_seconds.restype = ctypes.c_double
    ol.orc_algebra_seconds.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, dp]
    rl.ref_set_ptord(N)
    pa, pb = a.ctypes.data_as(dp), b.ctypes.data_as(dp)
    cs_r, cs_o = ctypes.c_double(), ctypes.c_double()
    rows = []
    for name, (m, e, lg) in MIX.items():
        rounds = max(4000, 400000 // (m + 10 * (e + lg)))
        # alternating runs, the fastest of each side (the host's other load
        # varies over the runs)
        best_r = best_o = float("inf")
        for _ in range(5):
            best_r = min(best_r, rl.ref_su3_algebra_seconds(pa, pb, rounds, m, e, lg, ctypes.byref(cs_r)))
            best_o = min(best_o, ol.orc_algebra_seconds(3, N, pa, pb, rounds, m, e, lg, ctypes.byref(cs_o)))
        rows.append({"ops": name, "per_round": {"products": m, "exp": e, "log": lg}, "rounds": rounds,
                     "reference_s": best_r, "oracle_s": best_o, "ratio": best_r / best_o,
                     "checksums_agree": abs(cs_r.value - cs_o.value) <= 1e-9 * max(1.0, abs(cs_r.value))})
        print(rows[-1], flush=True)
    out = {"cpu": cpu_model(), "cores": 1, "orders": N, "how": __doc__.strip().split("\n\n")[1].replace("\n", " "),
           "rows": rows, "ratio": [r["ratio"] for r in rows if r["ops"] == "step mix"][0]}
    with open(os.path.join(HERE, "calibration.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("cpu", "orders", "ratio")}))


if __name__ == "__main__":
    main()
