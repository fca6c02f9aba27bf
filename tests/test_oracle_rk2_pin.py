"""Pin of the RK2 corrector coefficient, which no reference test covers
(the reference has no RK2; the scheme is Batrouni, Katz, Kronfeld, Lepage,
Svetitsky, Wilson, PRD 32 (1985) 2736).

The corrector's (1 + nc eps/6 x^2) factor makes the scheme weak order 2, so
the stationary order-4 plaquette has no term linear in eps.  Long oracle
runs (tests/golden/make_long_runs.py, fixture long_runs_rk2.npz) at five eps
for SU(3) and SU(2) fix it: with the coefficient as restated the linear term
of a quadratic fit is zero within 3 sigma; with the coefficient planted at 0
or at twice its value (oracle.planted) it is not, so a wrong oracle could not
pass.  Order 2 of the plaquette does not see the coefficient (it shifts
drift orders >= 3 only) and must reach the equipartition value at eps -> 0.

Order 4 is estimated with order 2 as a control variate: P4 - kappa (P2 -
P2(0)), kappa the regression slope of the bin means, P2(0) the exact
equipartition value.  The two fluctuate together (correlation ~0.8), which
cuts the error of order 4 by ~40%; and since the order-2 dynamics is that of
a free field, on which RK2 is weak order 2 whatever the coefficient,
P2(eps) - P2(0) = O(eps^2) adds no linear term.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _load_rk2():
    return np.load(os.path.join(GOLD, "long_runs_rk2.npz"))


def run_means(bins):
    """Per-run means and their errors from the bin means (bins: runs x nbin x
    orders): the run mean, and the standard error of the bin means."""
    m = bins.mean(axis=1)
    e = bins.std(axis=1, ddof=1) / np.sqrt(bins.shape[1])
    return m, e


def combine(m, e):
    """Weighted mean of independent runs."""
    w = 1.0 / e ** 2
    return (m * w).sum(axis=0) / w.sum(axis=0), 1.0 / np.sqrt(w.sum(axis=0))


def eps_fit(eps, y, s):
    """Weighted least squares y = a + b eps + c eps^2: (coefficients, sigmas, chi2)."""
    X = np.vstack([np.ones_like(eps), eps, eps ** 2]).T
    W = np.diag(1.0 / s ** 2)
    C = np.linalg.inv(X.T @ W @ X)
    p = C @ X.T @ W @ y
    chi2 = float(((y - X @ p) / s) @ ((y - X @ p) / s))
    return p, np.sqrt(np.diag(C)), chi2


def equipartition(nc, V):
    """Order-2 plaquette at eps -> 0 with the zero modes removed."""
    return -(nc * nc - 1) * (1 - 1.0 / V) / 4


def controlled_order4(bins, p2_exact):
    """Mean and error of order 4 with order 2 as control variate (bins:
    runs x nbin x orders)."""
    p2, p4 = bins[:, :, 2].ravel(), bins[:, :, 4].ravel()
    kappa = np.cov(p2, p4)[0, 1] / np.var(p2, ddof=1)
    cv = bins[:, :, 4] - kappa * (bins[:, :, 2] - p2_exact)
    m, e = combine(*run_means(cv[:, :, None]))
    return m[0], e[0]


def rk2_order4_fit(d, inc, ifac, order=4):
    """Quadratic fit in eps of the plaquette of `order` (order 4: with the
    control variate)."""
    eps = d["eps"]
    nc, V = int(d["nc"][inc]), int(np.prod(d["L"]))
    y, s = [], []
    for ie in range(len(eps)):
        b = d["bins"][inc, ifac, ie]
        if order == 4:
            m, e = controlled_order4(b, equipartition(nc, V))
        else:
            m, e = combine(*run_means(b[:, :, order:order + 1]))
            m, e = m[0], e[0]
        y.append(m)
        s.append(e)
    return eps_fit(eps, np.array(y), np.array(s))


def test_rk2_fixture_reproduces_with_this_oracle():
    """The first steps of the fixture's runs, recomputed now: the long runs
    are those of this oracle."""
    d = _load_rk2()
    L, N = tuple(int(x) for x in d["L"]), int(d["N"])
    for inc, nc in enumerate(d["nc"]):
        U = np.zeros((4 * int(np.prod(L)), 1 + nc * nc * N), np.complex128)
        U[:, 0] = 1.0
        seed = int(d["seeds"][inc, 0, 0])
        st = d["stamps"][inc]
        got = O.run_stats(L, int(nc), N, U, float(d["eps"][0]), float(d["alpha"]), 1, 1, 1, seed, len(st), len(st))
        assert np.abs(got - st).max() <= 1e-12


@pytest.mark.parametrize("inc", [0, 1])
def test_rk2_coefficient_is_weak_order_two(inc):
    """Linear term of the order-4 plaquette in eps: zero within 3 sigma with
    the corrector's coefficient as restated, nonzero beyond 3 sigma with it
    planted at 0 or at twice its value."""
    d = _load_rk2()
    facs = list(d["fac"])
    p, s, chi2 = rk2_order4_fit(d, inc, facs.index(1.0))
    print("nc", d["nc"][inc], "fac 1: a, b, c", p, "+-", s, "chi2", chi2)
    assert abs(p[1]) < 3 * s[1], (p, s)
    for fac in (0.0, 2.0):
        q, t, _ = rk2_order4_fit(d, inc, facs.index(fac))
        print("fac", fac, "b = %.3f +- %.3f" % (q[1], t[1]))
        assert abs(q[1]) > 3 * t[1], (fac, q, t)


@pytest.mark.parametrize("inc", [0, 1])
def test_rk2_order2_equipartition(inc):
    """Order 2 at eps -> 0 (the quadratic fit's intercept): the equipartition
    value -(nc^2 - 1)(1 - 1/V)/4 with the zero modes removed."""
    d = _load_rk2()
    nc = int(d["nc"][inc])
    V = int(np.prod(d["L"]))
    p, s, _ = rk2_order4_fit(d, inc, list(d["fac"]).index(1.0), order=2)
    exact = equipartition(nc, V)
    print("order 2 at eps -> 0", p[0], "+-", s[0], "exact", exact)
    assert abs(p[0] - exact) < 3 * s[0]
