"""Long runs agree statistically with what the theory fixes.

With zero modes removed and eps -> 0, equipartition over the 3(V-1)(NC^2-1)
transverse order-1 modes gives the order-2 plaquette coefficient
<P_2> = -(NC^2-1)(1-1/V)/4 (SU(3), 4^4: -1.99219); odd orders average to
zero.  RK2's eps^2 bias at eps = 0.02 is far below the statistical error
allowed here.  The error is estimated by blocking (blocks longer than the
autocorrelation time).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def blocked(x, nblocks=40):
    b = np.array_split(x, nblocks)
    m = np.array([v.mean() for v in b])
    return x.mean(), m.std(ddof=1) / np.sqrt(nblocks)


@pytest.mark.parametrize("nc", [3, 2])
def test_order2_plaquette_equipartition(nc):
    from prlgt2020_amd import nspt
    L, N, eps = (4, 4, 4, 4), 4, 0.02
    V = 256
    want = -(nc * nc - 1) * (1 - 1 / V) / 4
    print("equipartition run: SU(%d) 4^4 N=4 RK2 eps=%g, device noise" % (nc, eps), flush=True)
    ctx = nspt.Context(L, nc, N)
    p = nspt.Context.params(eps, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_DEVICE, seed=11, step=0)
    ctx.evolve(p, 1000, want_obs=False)
    obs = ctx.evolve(p, 16000)
    plaq = obs[:, :2 * (N + 1)].view(np.complex128)
    m2, e2 = blocked(plaq[:, 2].real)
    m1, e1 = blocked(plaq[:, 1].real)
    m3, e3 = blocked(plaq[:, 3].real)
    print("order 2: %.5f +- %.5f (expected %.5f); order 1: %.2e +- %.1e; order 3: %.2e +- %.1e"
          % (m2, e2, want, m1, e1, m3, e3))
    assert abs(m2 - want) < max(5 * e2, 0.005 * abs(want))
    assert abs(m1) < 5 * e1 + 1e-12
    assert abs(m3) < 5 * e3 + 1e-12
    assert np.allclose(plaq[:, 0], 1.0)


def test_checkerboard_gauge_fixing_equipartition():
    """RK2 with the checkerboard sequential gauge fixing (alpha = -0.05, the
    reference's shipped value): the gauge-invariant order-2 plaquette takes
    the equipartition value as with FA."""
    from prlgt2020_amd import nspt
    L, N, nc, eps = (4, 4, 4, 4), 4, 3, 0.02
    want = -(nc * nc - 1) * (1 - 1 / 256) / 4
    ctx = nspt.Context(L, nc, N)
    p = nspt.Context.params(eps, -0.05, nspt.RK2, nspt.GF_CHECKERBOARD, 1, nspt.NOISE_DEVICE, seed=11, step=0)
    ctx.evolve(p, 1000, want_obs=False)
    plaq = ctx.evolve(p, 16000)[:, :2 * (N + 1)].view(np.complex128)
    m2, e2 = blocked(plaq[:, 2].real)
    m3, e3 = blocked(plaq[:, 3].real)
    print("checkerboard gauge fixing: order 2 %.5f +- %.5f (expected %.5f)" % (m2, e2, want))
    assert abs(m2 - want) < max(5 * e2, 0.005 * abs(want))
    assert abs(m3) < 5 * e3 + 1e-12
    ctx.close()


def test_checkerboard_sweep_converges_to_equipartition():
    """The on-the-fly Euler sweep in checkerboard order carries the Euler
    O(eps) bias; extrapolated linearly from eps = 0.01 and 0.005 to eps = 0 the
    order-2 plaquette is the equipartition value (within 3 sigma)."""
    from prlgt2020_amd import nspt
    L, N, nc = (4, 4, 4, 4), 2, 3
    want = -(nc * nc - 1) * (1 - 1 / 256) / 4
    res = []
    for eps, n in ((0.01, 32000), (0.005, 64000)):
        ctx = nspt.Context(L, nc, N)
        p = nspt.Context.params(eps, -0.05, nspt.EULER_CHECKERBOARD, nspt.GF_CHECKERBOARD, 1, nspt.NOISE_DEVICE,
                                seed=11, step=0)
        ctx.evolve(p, n // 16, want_obs=False)
        plaq = ctx.evolve(p, n)[:, :2 * (N + 1)].view(np.complex128)
        res.append(blocked(plaq[:, 2].real))
        ctx.close()
    (a, ea), (b, eb) = res
    p0 = 2 * b - a  # linear in eps through eps = 0.01 and 0.005
    e0 = np.hypot(2 * eb, ea)
    print("checkerboard sweep: order 2 at eps 0.01: %.5f, 0.005: %.5f, extrapolated %.5f +- %.5f (expected %.5f)"
          % (a, b, p0, e0, want))
    assert abs(p0 - want) < 3 * e0
    assert a < b < 0  # the bias shrinks with eps


def test_twisted_lattice_has_no_zero_modes():
    """Twisted boundary conditions in y and z with the reference's SU(3)
    twist matrices (an irreducible pair: W_y W_z = w W_z W_y): no colour
    component is constant on the lattice, so all 3 V (NC^2-1) transverse
    modes are nonzero and equipartition gives <P_2> = -(NC^2-1)/4 = -2
    exactly, without any zero-mode subtraction -- against -(NC^2-1)(1-1/V)/4
    = -1.99219 on the periodic lattice with its zero modes removed.  RK2 with
    the checkerboard sequential gauge fixing (FA needs periodic conditions)."""
    from prlgt2020_amd import nspt
    L, N, nc, eps = (4, 4, 4, 4), 4, 3, 0.02
    want, periodic = -(nc * nc - 1) / 4, -(nc * nc - 1) * (1 - 1 / 256) / 4
    print("twisted equipartition run: SU(3) 4^4 N=4 RK2 eps=%g, y and z twisted" % eps, flush=True)
    ctx = nspt.Context(L, nc, N)
    ctx.set_twist(0b0110)
    p = nspt.Context.params(eps, -0.05, nspt.RK2, nspt.GF_CHECKERBOARD, 0, nspt.NOISE_DEVICE, seed=23, step=0)
    ctx.evolve(p, 1000, want_obs=False)
    plaq = ctx.evolve(p, 48000)[:, :2 * (N + 1)].view(np.complex128)
    m2, e2 = blocked(plaq[:, 2].real)
    m3, e3 = blocked(plaq[:, 3].real)
    print("twisted: order 2 %.5f +- %.5f (expected %.5f; periodic with zero modes removed %.5f)"
          % (m2, e2, want, periodic))
    assert abs(m2 - want) < 4 * e2
    assert abs(m3) < 5 * e3 + 1e-12
