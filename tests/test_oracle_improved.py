"""The improved-action drift of the oracle (staple2x1, QCDenvNODEpt_test.h:579-611;
nspt.cc:385-412) pinned without retyping it: the rectangle staples close the
2x1 and 1x2 rectangles that the path-walked Wilson loops of test_oracle_loops
pin (each rectangle once per link, in both orientations), and the improved
drift is gauge covariant.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_invariants import (_gauge_transform, _random_gauge, _unitary_field, pdag, pmul, to_mats,
                                          from_mats)


@pytest.mark.parametrize("nc,L", [(3, (2, 2, 2, 3)), (2, (3, 2, 2, 2))])
def test_rectangle_staples_close_the_rectangles(nc, L):
    """sum_{x, mu, nu != mu} Tr(U_mu(x) staple2x1(x, mu, nu)) = 3 (l21 + conj(l21))."""
    N = 3
    V = int(np.prod(L))
    U = _unitary_field(L, nc, N, seed=51)
    tot = np.zeros(N + 1, complex)
    for s in range(V):
        for mu in range(4):
            R = sum(O.staple2x1(L, nc, N, U, s, mu, nu) for nu in range(4) if nu != mu)
            tot += O.tr(nc, N, O.mul(nc, N, U[4 * s + mu], R))
    l21 = O.compute_loops(L, nc, N, U)[0]
    assert np.abs(tot - 3 * (l21 + np.conj(l21))).max() < 1e-12 * np.abs(tot).max()


@pytest.mark.parametrize("nc", [3, 2])
def test_improved_drift_gauge_covariant(nc):
    """D[U^G](x, mu) = G(x) D[U](x, mu) G(x)^dag for the DBW2 coefficient."""
    L, N, c1 = (2, 2, 2, 3), 4, -1.4088
    V = int(np.prod(L))
    U = _unitary_field(L, nc, N, seed=52)
    G = _random_gauge(L, nc, N)
    D = O.drift(L, nc, N, U, 0.05, c1=c1)
    DG = O.drift(L, nc, N, _gauge_transform(L, nc, N, U, G), 0.05, c1=c1)
    assert np.abs(D - O.drift(L, nc, N, U, 0.05)).max() > 1e-3  # the rectangles do contribute
    for s in range(V):
        Gs = to_mats(G[s], nc, N)
        for mu in range(4):
            want = from_mats(pmul(pmul(Gs, to_mats(D[4 * s + mu], nc, N)), pdag(Gs)), nc)
            assert np.abs(DG[4 * s + mu] - want).max() < 1e-12
