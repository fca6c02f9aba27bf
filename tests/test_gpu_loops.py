"""Wilson loops (nspt_loops, ComputeLoops of nspt.cc:1156-1312) against the
oracle on the same field: per order, |gpu - oracle| <= 1e-12 x the largest
oracle coefficient of that order.  The loops are polynomials in the link
coefficients, so generic complex series (flag near 1) are used: on special
unitary links the order-1 traces vanish and would leave only rounding to
compare.  Lattices with extents 2..4 make the 4-link sides wrap around the
torus more than once."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nc,L,N", [(3, (2, 2, 2, 4), 6), (3, (4, 3, 2, 2), 12), (2, (2, 4, 2, 3), 6),
                                    (3, (2, 2, 2, 2), 32)])
def test_loops_match_oracle(nc, L, N):
    from prlgt2020_amd import nspt
    print("loops SU(%d) %s N=%d" % (nc, "x".join(map(str, L)), N), flush=True)
    rng = np.random.default_rng(41)
    S = 1 + nc * nc * N
    nl = 4 * int(np.prod(L))
    U = 0.3 * (rng.normal(size=(nl, S)) + 1j * rng.normal(size=(nl, S)))
    U[:, 0] = 1.0 + 0.1 * (rng.normal(size=nl) + 1j * rng.normal(size=nl))
    ctx = nspt.Context(L, nc, N)
    ctx.upload(U)
    got = ctx.loops()
    want = O.compute_loops(L, nc, N, U)
    scale = np.maximum(np.abs(want).max(axis=0), 1e-300)
    err = (np.abs(got - want).max(axis=0) / scale)
    print("per-order rel err", err)
    assert err.max() <= 1e-12
    # a second call reuses the buffers and gives the same numbers
    assert np.array_equal(ctx.loops(), got)
    ctx.close()


def test_loops_cold_start():
    from prlgt2020_amd import nspt
    L, nc, N = (4, 4, 2, 2), 3, 4
    V = 64
    ctx = nspt.Context(L, nc, N)
    ctx.upload(O.cold_field(L, nc, N))
    got = ctx.loops()
    assert np.allclose(got[:, 0], nc * np.array([2, 1, 2, 1]) * 6 * V)
    assert np.abs(got[:, 1:]).max() == 0.0
    ctx.close()


@pytest.mark.parametrize("nc", [3, 2])
def test_unitarize_matches_oracle(nc):
    """nspt_unitarize on a field of 3 orders padded to 6 (what loadlowerorder
    loads): exp(reH(log U)) per link; orders <= 3 keep their values and the
    result is unitary through order 6."""
    from prlgt2020_amd import nspt
    L, N, low = (2, 2, 2, 4), 6, 3
    ne = nc * nc
    U3 = O.random_field(L, nc, low, seed=5, scale=0.4)
    U = np.zeros((U3.shape[0], 1 + ne * N), np.complex128)
    U[:, :U3.shape[1]] = U3
    ctx = nspt.Context(L, nc, N)
    ctx.upload(U)
    ctx.unitarize()
    got = ctx.download()
    want = np.array([O.exp(nc, N, O.reh(nc, N, O.log(nc, N, u))) for u in U])
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    assert np.abs(got[:, :U3.shape[1]] - U3).max() < 1e-13
    one = np.zeros(1 + ne * N, np.complex128)
    one[0] = 1.0
    for u in got[:8]:
        assert np.abs(O.mul(nc, N, u, O.dag(nc, N, u)) - one).max() < 1e-13
    ctx.close()
