"""Twisted boundary conditions in the oracle (TWISTED_BC: nspt.cc:97-256,
lattice.h:156-1053, twistmatrices.h) pinned against the periodic oracle: a
twisted field on L is the periodic field on the covering lattice whose
extent along each twisted direction d is p_d L_d (W_d^p_d is a centre
element, so conjugation by it is the identity), unwrapped as
U(x + n L_d e_d) = W_d^n U(x) W_d^-n -- the reference's halo copies
(forward W A W^dag, backward W^dag A W, nspt.cc:163-219).  Drift, step
(Euler, RK2, checkerboard sweep, checkerboard gauge fixing, improved
action), plaquette and loops of the twisted oracle must equal the periodic
oracle's on the unwrapped field.  CPU only."""
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def period(nc):
    return 3 if nc == 3 else 2


def coords(L, s):
    x = [0] * 4
    for d in (3, 2, 1, 0):
        x[d] = s % L[d]
        s //= L[d]
    return x


def index(L, x):
    return ((x[0] * L[1] + x[1]) * L[2] + x[2]) * L[3] + x[3]


def conj_series(u, M, nc):
    out = u.copy()
    N = (len(u) - 1) // (nc * nc)
    m = u[1:].reshape(N, nc, nc)
    out[1:] = (M @ m @ M.conj().T).reshape(-1)
    return out


def unwrap(U, L, nc, mask, om):
    """Field (4V, S) or noise (4V, nc*nc) on the twisted L -> the periodic
    field on the covering lattice."""
    p = [period(nc) if (mask >> d) & 1 else 1 for d in range(4)]
    Lc = [L[d] * p[d] for d in range(4)]
    Vc = int(np.prod(Lc))
    out = np.zeros((4 * Vc, U.shape[1]), np.complex128)
    for sc in range(Vc):
        xc = coords(Lc, sc)
        x = [xc[d] % L[d] for d in range(4)]
        M = np.eye(nc, dtype=np.complex128)
        for d in range(4):
            M = M @ np.linalg.matrix_power(om[d], xc[d] // L[d])
        s = index(L, x)
        for mu in range(4):
            u = U[4 * s + mu]
            if U.shape[1] == nc * nc:  # a noise matrix
                out[4 * sc + mu] = (M @ u.reshape(nc, nc) @ M.conj().T).reshape(-1)
            else:
                out[4 * sc + mu] = conj_series(u, M, nc)
    return out, Lc


CASES = [(3, (2, 2, 2, 2), 0b0110), (3, (2, 2, 2, 2), 0b1100), (2, (2, 4, 2, 2), 0b1110),
         (3, (2, 2, 4, 2), 0b1111)]


def _field(L, nc, N, seed):
    return O.random_field(L, nc, N, seed=seed, scale=0.4)


@pytest.mark.parametrize("nc,L,mask", CASES)
def test_twist_matrices_are_centre_periodic(nc, L, mask):
    om = O.reference_twist(nc)
    for d in range(4):
        P = np.linalg.matrix_power(om[d], period(nc))
        assert np.allclose(P, P[0, 0] * np.eye(nc))
        assert np.allclose(om[d] @ om[d].conj().T, np.eye(nc))
    # the twist eaters of two twisted directions commute up to a centre phase
    for a, b in itertools.combinations(range(4), 2):
        c = om[a] @ om[b] @ np.linalg.inv(om[b] @ om[a])
        assert np.allclose(c, c[0, 0] * np.eye(nc))


@pytest.mark.parametrize("nc,L,mask", CASES)
def test_twisted_drift_is_the_covering_drift(nc, L, mask):
    N = 4
    om = O.reference_twist(nc)
    U = _field(L, nc, N, 3)
    with O.twisted(nc, mask):
        D = O.drift(L, nc, N, U, 0.05)
        D2 = O.drift(L, nc, N, U, 0.05, c1=-1.0 / 12)
        P = O.plaquette_measure(L, nc, N, U)
        W = O.compute_loops(L, nc, N, U)
    Uc, Lc = unwrap(U, L, nc, mask, om)
    ratio = int(np.prod(Lc)) // int(np.prod(L))
    assert np.abs(O.drift(Lc, nc, N, Uc, 0.05) - unwrap(D, L, nc, mask, om)[0]).max() < 1e-13
    assert np.abs(O.drift(Lc, nc, N, Uc, 0.05, c1=-1.0 / 12) - unwrap(D2, L, nc, mask, om)[0]).max() < 1e-13
    assert np.abs(O.plaquette_measure(Lc, nc, N, Uc) - P).max() < 1e-13
    Wc = O.compute_loops(Lc, nc, N, Uc)
    assert np.abs(Wc - ratio * W).max() < 1e-11 * np.abs(Wc).max()
    # and the twist matters: the periodic drift of the same field differs
    assert np.abs(O.drift(L, nc, N, U, 0.05) - D).max() > 1e-3


@pytest.mark.parametrize("nc,L,mask", CASES)
@pytest.mark.parametrize("integrator,update,gauge_fix,c1",
                         [(0, 0, 0, 0.0), (1, 0, 0, 0.0), (0, 2, 3, 0.0), (1, 0, 3, -0.331), (0, 0, 2, 0.0)])
def test_twisted_step_is_the_covering_step(nc, L, mask, integrator, update, gauge_fix, c1):
    """Step without zero modes (the mean of log U over the covering lattice
    is not that of the twisted one) and without FA; gauge_fix 2 is the
    lexicographic sequential sweep, which on the covering lattice visits the
    images in another order, so only its gauge-invariant output (the
    plaquette of the gauge-fixed field) is compared."""
    N = 4
    om = O.reference_twist(nc)
    U = _field(L, nc, N, 5)
    rng = np.random.default_rng(9)
    nl = 4 * int(np.prod(L))
    eta = 0.3 * (rng.normal(size=(nl, nc * nc)) + 1j * rng.normal(size=(nl, nc * nc)))
    kw = dict(alpha=-0.05, integrator=integrator, update=update, gauge_fix=gauge_fix, zero_modes=0, c1=c1)
    Ut = U.copy()
    with O.twisted(nc, mask):
        obs = O.step(L, nc, N, Ut, 0.05, eta=eta, **kw)
    Uc, Lc = unwrap(U, L, nc, mask, om)
    etac, _ = unwrap(eta, L, nc, mask, om)
    obsc = O.step(Lc, nc, N, Uc, 0.05, eta=etac, **kw)
    assert np.abs(obsc - obs).max() < 1e-12
    if gauge_fix == 2:
        with O.twisted(nc, mask):
            P = O.plaquette_measure(L, nc, N, Ut)
        assert np.abs(O.plaquette_measure(Lc, nc, N, Uc) - P).max() < 1e-12
    else:
        assert np.abs(Uc - unwrap(Ut, L, nc, mask, om)[0]).max() < 1e-12


def test_identity_twist_is_periodic():
    L, nc, N = (2, 2, 2, 4), 3, 4
    U = _field(L, nc, N, 7)
    eye = np.array([np.eye(nc)] * 4, np.complex128)
    with O.twisted(nc, 0b1111, eye):
        D = O.drift(L, nc, N, U, 0.05)
        W = O.compute_loops(L, nc, N, U)
    assert np.array_equal(D, O.drift(L, nc, N, U, 0.05))
    assert np.array_equal(W, O.compute_loops(L, nc, N, U))


def test_cold_start_is_a_twisted_field():
    """U = 1 satisfies every twisted boundary condition: zero drift, unit
    plaquette."""
    L, nc, N = (2, 2, 2, 2), 3, 4
    U = O.cold_field(L, nc, N)
    with O.twisted(nc, 0b1110):
        assert np.abs(O.drift(L, nc, N, U, 0.05)).max() == 0.0
        P = O.plaquette_measure(L, nc, N, U)
    assert P[0] == 1.0 and np.abs(P[1:]).max() == 0.0
