"""Pin the oracle's series algebra and random sources against golden vectors
written by the reference's own code (tests/golden/make_golden.py, from
QCDpt.h/QCDpt.cc, MyMath.cc, MyRand.cc, ranlxd/ranlxd.cc) and against the
ranlxd known-answer test of ranlxd/test/test.cc:47-102."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

G = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def alg():
    return np.load(os.path.join(G, "algebra.npz"))


@pytest.fixture(scope="module")
def rng():
    return np.load(os.path.join(G, "rng.npz"))


@pytest.mark.parametrize("nc", [3, 2])
@pytest.mark.parametrize("N", [1, 2, 4, 6])
@pytest.mark.parametrize("case", range(5))
def test_series_algebra(alg, nc, N, case):
    k = "nc%d_N%d_c%d" % (nc, N, case)
    a, b = alg[k + "_a"], alg[k + "_b"]
    scale = max(1.0, np.abs(alg[k + "_mul"]).max())
    assert np.abs(O.mul(nc, N, a, b) - alg[k + "_mul"]).max() <= 1e-13 * scale
    for op in ("exp", "log"):
        ref = alg[k + "_" + op]
        assert np.abs(getattr(O, op)(nc, N, a) - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    assert np.abs(O.reh(nc, N, a) - alg[k + "_reh"]).max() <= 1e-15 * np.abs(a).max() * 4
    assert np.array_equal(O.dag(nc, N, a), alg[k + "_dag"])
    assert np.abs(O.tr(nc, N, a) - alg[k + "_tr"]).max() <= 1e-14 * np.abs(a).max() * 4


def test_ranlxd_known_answer():
    kat = json.load(open(os.path.join(G, "ranlxd_kat.json")))
    r = O.Ranlxd(kat["seed"], kat["level"])
    xd = None
    for _ in range(kat["batches"]):
        xd = r.draw(kat["batch_size"])
    got = xd[kat["first_index"]:kat["first_index"] + len(kat["values"])] * 2.0 ** 48
    assert np.array_equal(got, np.array(kat["values"]))


def test_ranlxd_streams(rng):
    assert np.array_equal(O.Ranlxd(12345, 1).draw(3000), rng["ranlxd_l1_seed12345"])
    assert np.array_equal(O.Ranlxd(12345, 2).draw(3000), rng["ranlxd_l2_seed12345"])
    assert np.array_equal(O.Ranlxd(1804289383, 1).draw(3000), rng["ranlxd_l1_seed1804289383"])


def test_su3_noise_ranlxd(rng):
    got = O.Ranlxd(1804289383).noise(3, 200)
    assert np.abs(got - rng["su3rand_ranlxd_seed1804289383"]).max() == 0.0


def test_rand55(rng):
    assert np.array_equal(O.Rand55(1).ranf(3000), rng["rand55_ranf_seed1"])
    assert np.array_equal(O.Rand55(2682328493).gauss(3001), rng["rand55_gauss_seed2682328493"])
    assert np.array_equal(O.Rand55(1).noise(3, 200), rng["su3rand_rand55_seed1"])
    assert np.array_equal(O.Rand55(1).noise(2, 200), rng["su2rand_rand55_seed1"])


def test_glibc_seed_of_ranlxd():
    vals = json.load(open(os.path.join(G, "glibc_rand.json")))["values"]
    for seed, v in vals.items():
        assert O.glibc_first_rand(int(seed)) == v


def test_noise_variance():
    """Box-Muller with r = sqrt(-log(1-u)): each real gaussian has variance
    1/2 (MyMath.cc:305-306), i.e. <|P_12|^2> = 1 for the off-diagonal."""
    eta = O.Ranlxd(4242).noise(3, 40000)
    assert abs(np.mean(np.abs(eta[:, 1]) ** 2) - 1.0) < 0.03
    # anti-Hermitian, traceless
    m = eta.reshape(-1, 3, 3)
    assert np.abs(m + np.conj(np.swapaxes(m, 1, 2))).max() < 1e-15
    assert np.abs(np.trace(m, axis1=1, axis2=2)).max() < 1e-14


def test_philox_known_answer():
    """Philox4x32-10 reference vectors (Salmon et al., Random123 kat_vectors):
    counter 0, key 0 -> 6627e8d5 e169c58d bc57ac4c 9b00dbd8.  Checked through
    the oracle's uniform mapping (u + 0.5) 2^-32 of the first word."""
    import ctypes
    words = [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    u = [(w + 0.5) / 2 ** 32 for w in words]
    # reproduce the su(2) mapping of the oracle from these uniforms
    twopi = 8 * np.arctan(1.0)
    g = list(u)
    for i in (0, 2):
        t = twopi * g[i]
        r = np.sqrt(-np.log(1 - g[i + 1]))
        g[i], g[i + 1] = r * np.cos(t), r * np.sin(t)
    expect = np.array([1j * g[2], g[1] + 1j * g[0], -g[1] + 1j * g[0], -1j * g[2]])
    got = O.noise_philox(2, 0, 0, 1)[0]
    assert np.abs(got - expect).max() < 1e-15
