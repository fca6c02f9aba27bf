"""Fields larger than 4 GiB (config 5 sizes): SU(3) 32^4 at N = 32 and
SU(2) 32^4 at N = 24, whose order planes lie 0.6 GB and 0.27 GB apart, so
the highest ones start beyond a 32-bit byte offset of the field base.

Size-independent check: a field that is the same on every site (a random
series per direction) has the normalised plaquette, per order, of a 2^4
lattice holding the same links -- computed by the oracle -- and stays
translation invariant under a step whose noise is also the same on every
site.  Every order is populated, so a load from a wrong order plane shows
at O(1).  One RK2 step without gauge fixing or zero modes (on a constant
field both act trivially or remove everything); a second test runs them:
the log of a constant field is all zero mode, so subtracting it must leave
the identity at every order, and FA gauge fixing must keep it there."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nc,N", [(3, 32), (2, 24)])
def test_translation_invariant_field_at_32x4(nc, N):
    from prlgt2020_amd import nspt
    L, small = (32, 32, 32, 32), (2, 2, 2, 2)
    V, Vs = 32 ** 4, 16
    print("large-field run: SU(%d) 32^4 N=%d" % (nc, N), flush=True)
    U4 = O.random_field(small, nc, N, seed=29, scale=0.3)[:4]  # the four links of one site
    Us = np.tile(U4, (Vs, 1))
    ctx = nspt.Context(L, nc, N)
    big = np.tile(U4, (V, 1))
    ctx.upload(big)
    del big
    got0 = ctx.plaquette()
    want0 = O.plaquette_measure(small, nc, N, Us)
    eta4 = nspt.Ranlxd(777).noise(nc, 4)
    eta = np.tile(eta4, (V, 1))
    ctx.set_noise(eta)
    del eta
    ctx.step(nspt.Context.params(0.02, 0.0, nspt.RK2, nspt.GF_NONE, 0, nspt.NOISE_HOST), want_obs=False)
    O.step(small, nc, N, Us, 0.02, 0.0, integrator=1, update=0, gauge_fix=0, zero_modes=0,
           eta=np.tile(eta4, (Vs, 1)))
    got1 = ctx.plaquette()
    want1 = O.plaquette_measure(small, nc, N, Us)
    ctx.close()
    e0 = np.abs(got0 - want0).max() / np.abs(want0).max()
    e1 = np.abs(got1 - want1).max() / np.abs(want1).max()
    print("plaquette rel err %.2e, after a step %.2e" % (e0, e1))
    assert np.abs(want0[N]) > 1e-3 * np.abs(want0).max()  # the top order is populated
    assert e0 <= 1e-12
    # a step at 20+ orders is ill-conditioned in FP64 (see the parity tests)
    assert e1 <= 1e-10


@pytest.mark.parametrize("nc,N", [(3, 32), (2, 24)])
def test_zero_modes_and_fa_at_32x4(nc, N):
    from prlgt2020_amd import nspt
    L, V = (32, 32, 32, 32), 32 ** 4
    print("large-field zero modes + FA: SU(%d) 32^4 N=%d" % (nc, N), flush=True)
    U4 = O.random_field((2, 2, 2, 2), nc, N, seed=31, scale=0.3)[:4]
    ctx = nspt.Context(L, nc, N)
    big = np.tile(U4, (V, 1))
    ctx.upload(big)
    del big
    assert np.abs(ctx.plaquette()[1:]).max() > 1e-2
    eta4 = nspt.Ranlxd(778).noise(nc, 4)
    eta = np.tile(eta4, (V, 1))
    ctx.set_noise(eta)
    del eta
    ctx.step(nspt.Context.params(0.02, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_HOST), want_obs=False)
    got = ctx.plaquette()
    U = ctx.download()
    ctx.close()
    # a series is (flag, orders 1..N): the identity is flag 1, every order 0
    dev = max(np.abs(U[:, 1:]).max(), np.abs(U[:, 0] - 1).max())
    print("max |plaquette order >= 1| %.2e, max |U - 1| %.2e" % (np.abs(got[1:]).max(), dev))
    assert dev <= 1e-12
    assert np.abs(got[1:]).max() <= 1e-12
