"""Calibration of the oracle's CPU throughput against the reference's own
build (test infrastructure; writes oracle/calibration.json).

    OMP_NUM_THREADS=1 python oracle/calibrate.py

The reference cannot run where the GPU is (its sources stay here), so the
CPU baseline of a GPU run times the oracle and states this ratio beside it.
The reference's rates below were measured on this container's host (Intel
Xeon @ 2.70 GHz, one core, g++ 11.4 -O3 with the flags of its Makefile:2,
allocORD rebuilt to 20 for the 20-order case) in its default mode: update
on the fly (nspt.cc:365-441), the sequential gauge fixing (nspt.cc:660-862),
zero modes removed, Euler, RAND55 noise; per-step time differenced between
two run lengths.  The oracle is timed here in the same mode (update=1,
gauge_fix=2) on the same lattices; ratio = oracle rate / reference rate.
"""
import json
import os
import platform
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
# the script's own folder would make "oracle" name oracle.py, not the package
sys.path[:] = [os.path.dirname(HERE)] + [p for p in sys.path if os.path.abspath(p or ".") != HERE]
from oracle import oracle as O  # noqa: E402

# (lattice, orders, reference link-order updates/s on one core)
REFERENCE = [((8, 8, 8, 8), 6, 3.16e5), ((4, 4, 4, 4), 20, 7.09e4)]


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def oracle_rate(L, N, steps):
    nc = 3
    V = L[0] * L[1] * L[2] * L[3]
    U = O.cold_field(L, nc, N)
    rng = O.Rand55(1)
    for s in range(2):  # warm up, leave the cold start
        O.step(L, nc, N, U, 0.05, -0.05, integrator=0, update=1, gauge_fix=2, zero_modes=1, eta=rng.noise(nc, 4 * V))
    t = 0.0
    for s in range(steps):
        eta = rng.noise(nc, 4 * V)
        t0 = time.perf_counter()
        O.step(L, nc, N, U, 0.05, -0.05, integrator=0, update=1, gauge_fix=2, zero_modes=1, eta=eta)
        t += time.perf_counter() - t0
    return 4 * V * N * steps / t


def main():
    rows = []
    for L, N, ref in REFERENCE:
        r = oracle_rate(L, N, 6)
        rows.append({"lattice": list(L), "orders": N, "mode": "Euler, on the fly, sequential gauge fixing, zero modes",
                     "oracle": r, "reference": ref, "ratio": r / ref})
        print(rows[-1], flush=True)
    out = {"cpu": cpu_model(), "cores": 1, "rows": rows,
           "ratio": sum(r["ratio"] for r in rows) / len(rows)}
    with open(os.path.join(HERE, "calibration.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
