"""Momentum norms of the __K_MOM_ANALYSIS__ build (nspt.cc:919-944) through
nspt_momentum_norm: against the oracle (itself pinned against numpy's FFT,
tests/test_oracle_knorm.py) on small lattices, including extents without an
unrolled FFT and a runtime order below the compiled one; the field is left
as it was; and at the headline workload's size (SU(3) 16^4, 20 orders)
Parseval's identity per order and the zero momentum (the squared norm of the
sum of the field over the lattice)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _nspt():
    from prlgt2020_amd import nspt
    return nspt


@pytest.mark.parametrize("nc,L,N", [(3, (3, 5, 2, 7), 4), (2, (2, 6, 3, 2), 5), (3, (4, 16, 2, 2), 6),
                                    (3, (2, 2, 2, 4), 20), (2, (6, 4, 2, 10), 3)])
def test_knorm_matches_oracle(nc, L, N):
    nspt = _nspt()
    U = O.random_field(L, nc, N, seed=3, scale=0.4)
    ctx = nspt.Context(L, nc, N)
    ctx.upload(U)
    kn = ctx.momentum_norm()
    after = ctx.download()
    ctx.close()
    ref = O.knorm(L, nc, N, U)
    err = np.abs(kn - ref).max() / np.abs(ref).max()
    print("L", L, "nc", nc, "N", N, "max rel err %.2e" % err)
    assert err <= 1e-13
    assert np.array_equal(after, U)


def test_knorm_headline_size_parseval_and_zero_momentum():
    nspt = _nspt()
    L, nc, N = (16, 16, 16, 16), 3, 20
    V = 16 ** 4
    ctx = nspt.Context(L, nc, N)
    ctx.evolve(nspt.Context.params(0.02, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_DEVICE, seed=9), 3,
               want_obs=False)
    kn = ctx.momentum_norm()
    U = ctx.download()
    ctx.close()
    ne = nc * nc
    for k in range(N):
        u = U[:, 1 + k * ne:1 + (k + 1) * ne]
        e = (np.abs(u) ** 2).sum()
        assert abs(kn[k].sum() - V * e) <= 1e-12 * V * e, k
        # p = 0: sum over mu and elements of |sum_x U_mu,k(x)|^2
        z = (np.abs(u.reshape(V, 4, ne).sum(axis=0)) ** 2).sum()
        assert abs(kn[k][0] - z) <= 1e-10 * max(z, kn[k].max()), k
