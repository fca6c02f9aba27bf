"""Long oracle runs for the statistical tests (writes tests/golden/long_runs_*.npz).

    python tests/golden/make_long_runs.py rk2      # ~20 min on 7 cores
    python tests/golden/make_long_runs.py config1  # ~15 min on 7 cores

Both use the oracle's Jacobi step (oracle/nspt_oracle.cpp) with its ranlxd
noise, each run from the cold start, thermalised, then binned.  The runs are
independent streams (ranlxd seeds listed in the file).

rk2: the RK2 step + zero modes + FA (alpha = 0.05) on a 2^4 lattice, N = 4,
  SU(3) and SU(2), eps in EPS_RK2, with the corrector's nc*eps/6 coefficient
  as restated (factor 1) and planted wrong (factors 0 and 2).  For each
  (nc, factor, eps): RUNS runs of 800/eps steps after 4/eps thermalisation
  steps, NBIN bins each.  The test fits the order-4 plaquette against eps:
  the published scheme is weak order 2, so its linear term must vanish, and
  the wrong coefficients must show one.
config1: the Euler step + zero modes + FA (alpha = 0.05) of SU(3) 4^4, N = 6,
  eps = 0.05 (the CPU-runnable workload's shape with the Jacobi update), RUNS
  runs of 20000 steps after 400, NBIN bins each: the CPU side of the per-order
  comparison with long GPU runs.

A prefix of run 0 of each group is stored step by step ("stamp"): the CPU
test recomputes it, so a change of the oracle that would change these runs
shows up without rerunning them.
"""
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
EPS_RK2 = (0.01, 0.02, 0.04, 0.06, 0.08)
FAC_RK2 = (1.0, 0.0, 2.0)
NC_RK2 = (3, 2)
L_RK2, N_RK2 = (2, 2, 2, 2), 4
RUNS = 8
NBIN = 40
STAMP = 20
L_C1, N_C1, EPS_C1, STEPS_C1, THERM_C1 = (4, 4, 4, 4), 6, 0.05, 20000, 400
ALPHA = 0.05


def cold(L, nc, N):
    V = int(np.prod(L))
    U = np.zeros((4 * V, 1 + nc * nc * N), np.complex128)
    U[:, 0] = 1.0
    return U


def rk2_seed(nc, ie, r):
    return 10000 * nc + 100 * ie + r + 1


def c1_seed(r):
    return 70000 + r + 1


def rk2_steps(eps):
    # per run; thermalisation 4/eps
    n = int(round(800 / eps))
    return n - n % NBIN, int(round(4 / eps))


def _rk2_job(args):
    nc, fac, ie, r = args
    eps = EPS_RK2[ie]
    n, nth = rk2_steps(eps)
    U = cold(L_RK2, nc, N_RK2)
    seed = rk2_seed(nc, ie, r)
    with O.planted(rk2=fac):
        O.run_stats(L_RK2, nc, N_RK2, U, eps, ALPHA, 1, 1, 1, seed, nth, 1)
        b = O.run_stats(L_RK2, nc, N_RK2, U, eps, ALPHA, 1, 1, 1, seed + 500000, n, NBIN)
    return args, b


def _c1_job(r):
    U = cold(L_C1, 3, N_C1)
    seed = c1_seed(r)
    O.run_stats(L_C1, 3, N_C1, U, EPS_C1, ALPHA, 0, 1, 1, seed, THERM_C1, 1)
    return r, O.run_stats(L_C1, 3, N_C1, U, EPS_C1, ALPHA, 0, 1, 1, seed + 500000, STEPS_C1, NBIN)


def stamp(L, nc, N, eps, integrator, seed, fac=1.0):
    """Plaquette (real, orders 0..N) of the first STAMP steps from the cold
    start with ranlxd(seed) noise, one step per bin."""
    U = cold(L, nc, N)
    with O.planted(rk2=fac):
        return O.run_stats(L, nc, N, U, eps, ALPHA, integrator, 1, 1, seed, STAMP, STAMP)


def make_rk2(procs):
    jobs = [(nc, fac, ie, r) for nc in NC_RK2 for fac in FAC_RK2 for ie in range(len(EPS_RK2)) for r in range(RUNS)]
    # longest first
    jobs.sort(key=lambda j: -rk2_steps(EPS_RK2[j[2]])[0] * (3 if j[0] == 3 else 1))
    bins = np.zeros((len(NC_RK2), len(FAC_RK2), len(EPS_RK2), RUNS, NBIN, N_RK2 + 1))
    t = time.time()
    with Pool(procs) as p:
        for (nc, fac, ie, r), b in p.imap_unordered(_rk2_job, jobs):
            bins[NC_RK2.index(nc), FAC_RK2.index(fac), ie, r] = b
            print("rk2 nc %d fac %.0f eps %.3f run %d  %.0f s" % (nc, fac, EPS_RK2[ie], r, time.time() - t),
                  flush=True)
    stamps = np.array([stamp(L_RK2, nc, N_RK2, EPS_RK2[0], 1, rk2_seed(nc, 0, 0)) for nc in NC_RK2])
    np.savez_compressed(os.path.join(HERE, "long_runs_rk2.npz"), bins=bins, eps=np.array(EPS_RK2),
                        fac=np.array(FAC_RK2), nc=np.array(NC_RK2), L=np.array(L_RK2), N=N_RK2, alpha=ALPHA,
                        steps=np.array([rk2_steps(e)[0] for e in EPS_RK2]),
                        therm=np.array([rk2_steps(e)[1] for e in EPS_RK2]),
                        seeds=np.array([[[rk2_seed(nc, ie, r) for r in range(RUNS)] for ie in range(len(EPS_RK2))]
                                        for nc in NC_RK2]),
                        stamps=stamps)


def make_config1(procs):
    bins = np.zeros((RUNS, NBIN, N_C1 + 1))
    t = time.time()
    with Pool(procs) as p:
        for r, b in p.imap_unordered(_c1_job, range(RUNS)):
            bins[r] = b
            print("config1 run %d  %.0f s" % (r, time.time() - t), flush=True)
    np.savez_compressed(os.path.join(HERE, "long_runs_config1.npz"), bins=bins, L=np.array(L_C1), N=N_C1, nc=3,
                        eps=EPS_C1, alpha=ALPHA, steps=STEPS_C1, therm=THERM_C1,
                        seeds=np.array([c1_seed(r) for r in range(RUNS)]),
                        stamps=stamp(L_C1, 3, N_C1, EPS_C1, 0, c1_seed(0)))


if __name__ == "__main__":
    procs = int(os.environ.get("PROCS", "7"))
    what = sys.argv[1:] or ["rk2", "config1"]
    if "rk2" in what:
        make_rk2(procs)
    if "config1" in what:
        make_config1(procs)
