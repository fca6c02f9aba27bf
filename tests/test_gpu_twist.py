"""Twisted boundary conditions (nspt_set_twist; TWISTED_BC of nspt.cc:97-256,
lattice.h:156-1053, twistmatrices.h) on the GPU against the oracle with the
same twist, through the C ABI: drift, steps (Jacobi Euler, RK2, improved
action, checkerboard sweep with checkerboard gauge fixing), plaquette and
Wilson loops.  Bar: <= 1e-12 relative per order, as the periodic parity
tests.  The oracle's twisted lattice is itself pinned against the periodic
oracle on the covering lattice (test_oracle_twist.py)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _nspt():
    from prlgt2020_amd import nspt
    return nspt


def test_reference_twist_tables_agree():
    """The product's and the oracle's restatements of twistmatrices.h."""
    nspt = _nspt()
    for nc in (3, 2):
        assert np.allclose(nspt.reference_twist(nc), O.reference_twist(nc), atol=1e-15)


# (nc, lattice, N, mask): N = 32 runs the drift's staging path without
# register prefetch; extents 2 make every neighbour cross a boundary
DRIFT_CASES = [(3, (2, 2, 2, 4), 6, 0b0110), (3, (2, 4, 2, 2), 20, 0b1111), (3, (2, 2, 2, 2), 32, 0b1100),
               (2, (2, 2, 4, 2), 6, 0b1110), (2, (4, 2, 2, 2), 24, 0b0110)]


@pytest.mark.parametrize("nc,L,N,mask", DRIFT_CASES)
def test_twisted_drift(nc, L, N, mask, relerr):
    nspt = _nspt()
    print("twisted drift SU(%d) %s N=%d mask=%d" % (nc, "x".join(map(str, L)), N, mask), flush=True)
    U = O.random_field(L, nc, N, seed=11, scale=0.4)
    ctx = nspt.Context(L, nc, N)
    ctx.set_twist(mask)
    ctx.upload(U)
    got = ctx.drift(0.05)
    with O.twisted(nc, mask):
        want = O.drift(L, nc, N, U, 0.05)
        plaq = O.plaquette_measure(L, nc, N, U)
    e = relerr(got, want, nc, N)
    print("per-order rel err", e)
    assert e.max() <= TOL
    assert np.abs(ctx.plaquette() - plaq).max() <= TOL * np.abs(plaq).max()
    ctx.close()


@pytest.mark.parametrize("nc,mask", [(3, 0b0110), (2, 0b1110)])
@pytest.mark.parametrize("integrator,gauge_fix,c1", [(0, 0, 0.0), (1, 0, 0.0), (1, 0, -1.0 / 12), (2, 2, 0.0)])
def test_twisted_steps(nc, mask, integrator, gauge_fix, c1, relerr):
    """Two steps with zero modes (the reference subtracts them on twisted
    lattices as well, nspt.cc:1001-1003): integrator 2 is the checkerboard
    sweep (oracle update 2), gauge_fix 2 the checkerboard sequential gauge
    fixing (oracle gauge_fix 3)."""
    nspt = _nspt()
    L, N = (2, 2, 4, 4), 6
    V = 64
    print("twisted steps SU(%d) mask=%d integrator=%d gf=%d c1=%g" % (nc, mask, integrator, gauge_fix, c1),
          flush=True)
    U = O.random_field(L, nc, N, seed=3, scale=0.4)
    ctx = nspt.Context(L, nc, N)
    ctx.set_twist(mask)
    if c1:
        ctx.set_action(c1)
    ctx.upload(U)
    rng = nspt.Ranlxd(1234)
    Uo = U.copy()
    upd, oint = (2, 0) if integrator == 2 else (0, integrator)
    for s in range(2):
        eta = rng.noise(nc, 4 * V)
        ctx.set_noise(eta)
        obs_g = ctx.step(nspt.Context.params(0.02, -0.05, integrator, gauge_fix, 1, nspt.NOISE_HOST))
        with O.twisted(nc, mask):
            obs_o = O.step(L, nc, N, Uo, 0.02, -0.05, integrator=oint, update=upd, gauge_fix=3 if gauge_fix else 0,
                           zero_modes=1, eta=eta, c1=c1)
        e = relerr(ctx.download(), Uo, nc, N)
        print("step", s, "per-order rel err", e)
        assert e.max() <= TOL, e
        for a, b in zip(nspt.split_obs(obs_g, N), O.split_obs(obs_o, N)):
            assert np.abs(a - b).max() <= TOL * max(1.0, np.abs(b).max())
    ctx.close()


@pytest.mark.parametrize("nc,L,N,mask", [(3, (2, 2, 2, 4), 6, 0b0110), (3, (2, 3, 2, 2), 12, 0b1111),
                                         (2, (2, 2, 4, 2), 6, 0b1110)])
def test_twisted_loops(nc, L, N, mask):
    nspt = _nspt()
    rng = np.random.default_rng(41)
    S = 1 + nc * nc * N
    nl = 4 * int(np.prod(L))
    U = 0.3 * (rng.normal(size=(nl, S)) + 1j * rng.normal(size=(nl, S)))
    U[:, 0] = 1.0 + 0.1 * (rng.normal(size=nl) + 1j * rng.normal(size=nl))
    ctx = nspt.Context(L, nc, N)
    ctx.set_twist(mask)
    ctx.upload(U)
    got = ctx.loops()
    with O.twisted(nc, mask):
        want = O.compute_loops(L, nc, N, U)
    scale = np.maximum(np.abs(want).max(axis=0), 1e-300)
    err = np.abs(got - want).max(axis=0) / scale
    print("per-order rel err", err)
    assert err.max() <= TOL
    ctx.close()


def test_identity_twist_is_periodic():
    """With every direction twisted by the identity the kernels take the
    twisted load path and give the periodic drift bit for bit."""
    nspt = _nspt()
    L, nc, N = (2, 2, 2, 4), 3, 6
    U = O.random_field(L, nc, N, seed=13, scale=0.4)
    ctx = nspt.Context(L, nc, N)
    ctx.upload(U)
    d0 = ctx.drift(0.05)
    ctx.set_twist(0b1111, np.array([np.eye(nc)] * 4))
    assert np.array_equal(ctx.drift(0.05), d0)
    ctx.set_twist(0)
    assert np.array_equal(ctx.drift(0.05), d0)
    ctx.close()


def test_twist_errors():
    """Non-monomial twist matrices, a bad mask and the FA gauge fixing on a
    twisted lattice are refused (no kernel runs)."""
    nspt = _nspt()
    ctx = nspt.Context((2, 2, 2, 2), 3, 4)
    bad = nspt.reference_twist(3)
    bad[1] = np.full((3, 3), 1 / np.sqrt(3))
    with pytest.raises(nspt.NsptError, match="modulus 1"):
        ctx.set_twist(0b0010, bad)
    half = nspt.reference_twist(3)
    half[2] = 0.5 * half[2]
    with pytest.raises(nspt.NsptError, match="modulus 1"):
        ctx.set_twist(0b0100, half)
    with pytest.raises(nspt.NsptError, match="mask"):
        ctx.set_twist(16)
    ctx.set_twist(0b0110)
    with pytest.raises(nspt.NsptError, match="periodic"):
        ctx.step(nspt.Context.params(0.02, 0.05, nspt.EULER, nspt.GF_FA, 1, nspt.NOISE_ZERO))
    with pytest.raises(nspt.NsptError, match="periodic"):
        ctx.fa_gauge_fixing(0.05)
    ctx.close()
