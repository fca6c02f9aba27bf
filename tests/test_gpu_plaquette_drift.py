"""The plaquette form of the Jacobi drift (k_plaq_field + k_drift_p): the
drift W = U Sigma as the sum of the plaquettes through each link, rotated to
start there.  The product path uses it above 16 orders, where a step is
FP64-ill-conditioned anyway; here NSPT_DRIFT_FORM=plaquettes forces it at
orders where the 1e-12 bar holds, against the oracle (which forms the
staple sum as nspt.cc:383-417 does): drift, steps with observables, the
improved action and a twisted lattice."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture
def nspt(monkeypatch):
    monkeypatch.setenv("NSPT_DRIFT_FORM", "plaquettes")
    from prlgt2020_amd import nspt as mod
    return mod


@pytest.mark.parametrize("nc,L,N", [(3, (2, 2, 2, 4), 6), (3, (4, 2, 2, 2), 12), (2, (2, 4, 2, 2), 8),
                                    (3, (2, 2, 2, 2), 16)])
def test_plaquette_drift(nspt, nc, L, N, relerr):
    print("plaquette drift SU(%d) %s N=%d" % (nc, "x".join(map(str, L)), N), flush=True)
    U = O.random_field(L, nc, N, seed=17, scale=0.4)
    ctx = nspt.Context(L, nc, N)
    ctx.upload(U)
    e = relerr(ctx.drift(0.05), O.drift(L, nc, N, U, 0.05), nc, N)
    print("per-order rel err", e)
    assert e.max() <= TOL
    ctx.close()


@pytest.mark.parametrize("nc,integrator,c1,mask", [(3, 1, 0.0, 0), (2, 0, 0.0, 0), (3, 1, -1.0 / 12, 0),
                                                   (3, 1, 0.0, 0b0110), (2, 1, -0.331, 0b1110)])
def test_plaquette_drift_steps(nspt, nc, integrator, c1, mask, relerr):
    """Two steps with zero modes, FA (periodic) or the checkerboard gauge
    fixing (twisted), fields and observables."""
    print("plaquette steps SU(%d) int=%d c1=%g mask=%d" % (nc, integrator, c1, mask), flush=True)
    L, N = (2, 2, 4, 4), 6
    V = 64
    U = O.random_field(L, nc, N, seed=3, scale=0.4)
    ctx = nspt.Context(L, nc, N)
    if mask:
        ctx.set_twist(mask)
    if c1:
        ctx.set_action(c1)
    ctx.upload(U)
    rng = nspt.Ranlxd(1234)
    Uo = U.copy()
    gf, ogf, alpha = (nspt.GF_CHECKERBOARD, 3, -0.05) if mask else (nspt.GF_FA, 1, 0.05)
    for s in range(2):
        eta = rng.noise(nc, 4 * V)
        ctx.set_noise(eta)
        obs_g = ctx.step(nspt.Context.params(0.02, alpha, integrator, gf, 1, nspt.NOISE_HOST))
        with O.twisted(nc, mask):
            obs_o = O.step(L, nc, N, Uo, 0.02, alpha, integrator=integrator, update=0, gauge_fix=ogf, zero_modes=1,
                           eta=eta, c1=c1)
        e = relerr(ctx.download(), Uo, nc, N)
        print("step", s, "per-order rel err", e)
        assert e.max() <= TOL, e
        for a, b in zip(nspt.split_obs(obs_g, N), O.split_obs(obs_o, N)):
            assert np.abs(a - b).max() <= TOL * max(1.0, np.abs(b).max())
    ctx.close()


@pytest.mark.parametrize("nc,N", [(3, 5), (3, 18), (2, 22)])
def test_plaquette_drift_runtime_order_below_instance(nspt, nc, N, relerr):
    """A runtime order below the compiled one (5 on the N = 6 instance, 18
    on 20, 22 on SU(2)'s 24): the products truncate at the runtime order.
    Drift of a random field, against the oracle at 1e-12 per order."""
    L = (2, 2, 2, 4)
    U = O.random_field(L, nc, N, seed=19, scale=0.3)
    ctx = nspt.Context(L, nc, N)
    ctx.upload(U)
    e = relerr(ctx.drift(0.05), O.drift(L, nc, N, U, 0.05), nc, N)
    print("per-order rel err", e)
    assert e.max() <= TOL
    ctx.close()
