"""The oracle's ComputeLoops (nspt.cc:1156-1312) pinned independently of its
neighbour chain: every sum is rebuilt from Wilson loops walked link by link
along the rectangle's boundary with plain numpy series products, and the
loops are checked for gauge invariance and their cold-start values.  CPU only.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_invariants import (_gauge_transform, _link, _random_gauge, _unitary_field, pdag, pmul,
                                          shift, to_mats)


def wilson_loop(L, nc, N, U, s, mu, nu, R, T):
    """R x T loop at s in plane (mu, nu): R links along mu, T along nu, back."""
    P = to_mats(np.concatenate([[1.0], np.zeros(nc * nc * N)]), nc, N)
    x = s
    for _ in range(R):
        P = pmul(P, _link(U, x, mu, nc, N))
        x = shift(L, x, mu, 1)
    for _ in range(T):
        P = pmul(P, _link(U, x, nu, nc, N))
        x = shift(L, x, nu, 1)
    for _ in range(R):
        x = shift(L, x, mu, -1)
        P = pmul(P, pdag(_link(U, x, mu, nc, N)))
    for _ in range(T):
        x = shift(L, x, nu, -1)
        P = pmul(P, pdag(_link(U, x, nu, nc, N)))
    return np.array([np.trace(m) for m in P])


def loops_by_walking(L, nc, N, U):
    V = int(np.prod(L))
    out = np.zeros((4, N + 1), complex)
    shapes = [[(2, 1), (1, 2)], [(2, 2)], [(4, 2), (2, 4)], [(4, 4)]]
    for mu in range(4):
        for nu in range(mu + 1, 4):
            for s in range(V):
                for w, rects in enumerate(shapes):
                    for R, T in rects:
                        out[w] += wilson_loop(L, nc, N, U, s, mu, nu, R, T)
    return out


@pytest.mark.parametrize("nc,L", [(3, (2, 2, 2, 3)), (2, (3, 2, 2, 2))])
def test_loops_are_wilson_loops(nc, L):
    N = 3
    U = _unitary_field(L, nc, N, seed=31)
    got = O.compute_loops(L, nc, N, U)
    want = loops_by_walking(L, nc, N, U)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    # a field with orders: the loops are not their cold values
    assert np.abs(want[:, 1:]).max() > 1e-3


@pytest.mark.parametrize("nc", [3, 2])
def test_loops_gauge_invariant(nc):
    L, N = (2, 2, 2, 3), 4
    U = _unitary_field(L, nc, N, seed=32)
    l0 = O.compute_loops(L, nc, N, U)
    lg = O.compute_loops(L, nc, N, _gauge_transform(L, nc, N, U, _random_gauge(L, nc, N)))
    assert np.abs(lg - l0).max() < 1e-12 * np.abs(l0).max()


@pytest.mark.parametrize("nc", [3, 2])
def test_loops_cold_start(nc):
    """Every loop of the cold field is the identity: Tr = nc at order 0."""
    L, N = (2, 2, 2, 2), 4
    V = 16
    loops = O.compute_loops(L, nc, N, O.cold_field(L, nc, N))
    count = np.array([2, 1, 2, 1]) * 6 * V  # loops per sum: rectangles x planes x sites
    assert np.allclose(loops[:, 0], nc * count)
    assert np.abs(loops[:, 1:]).max() == 0.0
