"""Write the golden vectors of tests/golden/ from oracle/_ref, the reference's
own series algebra and random sources (MyMath.cc, MyRand.cc, QCDpt.cc,
ranlxd/ranlxd.cc compiled from /root/reference by oracle/Makefile).

    python tests/golden/make_golden.py

Outputs (numpy .npz, no pickles):
  algebra.npz  inputs and outputs of ptSU3/ptSU2 product, exp, log, reH,
               dag, Tr at PTORD 1..6 (allocORD = 6 of choices.h:30)
  rng.npz      ranlxd streams (levels 1, 2), SU3rand(ranlxd) matrices,
               RAND55 uniforms and Box-Muller gaussians, SU3rand/SU2rand(RAND55)
The ranlxd known-answer values of ranlxd/test/test.cc:47-94 are kept as data
in ranlxd_kat.json (seed 32767, level 1, draws 39..86 of the 20th batch of
99, times 2^48).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402


def main():
    O.build()
    R = O.Ref()
    rng = np.random.default_rng(20201)
    alg = {}
    for nc in (3, 2):
        for N in (1, 2, 4, 6):
            S = O.ser_len(nc, N)
            for case, (fa, fb) in enumerate([(0, 0), (1, 0), (0, 1), (1, 1), (0.5 + 0.25j, -1.5 + 2j)]):
                a = rng.normal(size=S) + 1j * rng.normal(size=S)
                b = rng.normal(size=S) + 1j * rng.normal(size=S)
                a[0], b[0] = fa, fb
                key = "nc%d_N%d_c%d" % (nc, N, case)
                alg[key + "_a"] = a
                alg[key + "_b"] = b
                alg[key + "_mul"] = R.op("mul", nc, N, a, b)
                alg[key + "_exp"] = R.op("exp", nc, N, a)
                alg[key + "_log"] = R.op("log", nc, N, a)
                alg[key + "_reh"] = R.op("reh", nc, N, a)
                alg[key + "_dag"] = R.op("dag", nc, N, a)
                alg[key + "_tr"] = R.tr(nc, N, a)
    np.savez_compressed(os.path.join(HERE, "algebra.npz"), **alg)

    r = {}
    r["ranlxd_l1_seed12345"] = R.ranlxd(12345, 3000, 1)
    r["ranlxd_l2_seed12345"] = R.ranlxd(12345, 3000, 2)
    r["ranlxd_l1_seed1804289383"] = R.ranlxd(1804289383, 3000, 1)
    r["su3rand_ranlxd_seed1804289383"] = R.su3rand_ranlxd(1804289383, 200)
    r["rand55_ranf_seed1"] = R.rand55_ranf(1, 3000)
    r["rand55_gauss_seed2682328493"] = R.rand55_gauss(2682328493, 3001)
    r["su3rand_rand55_seed1"] = R.su3rand_rand55(1, 200)
    r["su2rand_rand55_seed1"] = R.su2rand_rand55(1, 200)
    np.savez_compressed(os.path.join(HERE, "rng.npz"), **r)
    print("wrote", os.path.join(HERE, "algebra.npz"), os.path.join(HERE, "rng.npz"))


if __name__ == "__main__":
    main()
