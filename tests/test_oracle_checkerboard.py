"""The checkerboard forms of the reference's default algorithms in the oracle
(ordered on-the-fly update in parity x direction order, nspt.cc:365-441; the
sequential gauge fixing of nspt.cc:660-862 in checkerboard site order),
pinned independently: the update against a Python sequencing of the
oracle's pinned drift, exp and product; the gauge fixing as a gauge
transformation that shrinks the divergence of log U for the shipped
alpha = -0.05.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_invariants import _unitary_field, coords


def parity(L, s):
    return sum(coords(L, s)) & 1


@pytest.mark.parametrize("nc", [3, 2])
def test_checkerboard_update_is_an_ordered_sweep(nc):
    L, N, eps = (2, 2, 2, 4), 4, 0.05
    V = int(np.prod(L))
    U = _unitary_field(L, nc, N, seed=61)
    rng = np.random.default_rng(62)
    eta = 0.3 * (rng.normal(size=(4 * V, nc * nc)) + 1j * rng.normal(size=(4 * V, nc * nc)))
    got = U.copy()
    O.step(L, nc, N, got, eps, 0.0, integrator=0, update=2, gauge_fix=0, zero_modes=1, eta=eta)
    # the same sweep sequenced here: for parity, for mu, every link of the subset
    # from the drift of the current field
    want = U.copy()
    ne = nc * nc
    for par in range(2):
        for mu in range(4):
            D = O.drift(L, nc, N, want, eps)
            for s in range(V):
                if parity(L, s) != par:
                    continue
                W = D[4 * s + mu].copy()
                W[1:1 + ne] -= np.sqrt(eps) * eta[4 * s + mu]
                want[4 * s + mu] = O.mul(nc, N, O.exp(nc, N, W), want[4 * s + mu])
    MM = np.array([sum(O.log(nc, N, want[4 * s + mu]) for s in range(V)) for mu in range(4)])
    O.zero_modes(L, nc, N, want, MM)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    # not the Jacobi update
    jac = U.copy()
    O.step(L, nc, N, jac, eps, 0.0, integrator=0, update=0, gauge_fix=0, zero_modes=1, eta=eta)
    assert np.abs(jac - got).max() > 1e-6


def _div_norm(L, nc, N, U):
    """sum_x |sum_mu A_mu(x) - A_mu(x-mu)|^2 of the order-1 coefficient of log U."""
    V = int(np.prod(L))
    ne = nc * nc
    A = np.array([O.log(nc, N, u)[1:1 + ne] for u in U]).reshape(V, 4, ne)
    tot = 0.0
    for s in range(V):
        x = coords(L, s)
        d = np.zeros(ne, complex)
        for mu in range(4):
            y = list(x)
            y[mu] = (y[mu] - 1) % L[mu]
            b = ((y[0] * L[1] + y[1]) * L[2] + y[2]) * L[3] + y[3]
            d += A[s, mu] - A[b, mu]
        tot += np.sum(np.abs(d) ** 2)
    return tot


@pytest.mark.parametrize("nc", [3, 2])
def test_checkerboard_gauge_fixing(nc):
    L, N = (2, 2, 4, 4), 4
    U = _unitary_field(L, nc, N, seed=63)
    p0, l0 = O.plaquette_measure(L, nc, N, U), O.compute_loops(L, nc, N, U)
    G = U.copy()
    O.sgf_checkerboard(L, nc, N, G, -0.05)
    assert np.abs(G - U).max() > 1e-4
    assert np.abs(O.plaquette_measure(L, nc, N, G) - p0).max() < 1e-12
    assert np.abs(O.compute_loops(L, nc, N, G) - l0).max() < 1e-12 * np.abs(l0).max()
    assert _div_norm(L, nc, N, G) < 0.8 * _div_norm(L, nc, N, U)
