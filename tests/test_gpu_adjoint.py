"""The adjoint drift kernel (nspt_adjoint_drift): reverse-mode gradient of the
order-truncated Wilson action of a generalised field (Haar-random SU(N)
order 0, Gaussian su(N) orders 1..N of scale 0.1, seed 7), checked against
the oracle's reverse pass, central finite differences (h = 1e-6) of the
action, and the staple drift tau reH(U Sigma)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def reh(m):
    a = 0.5 * (m - np.conj(np.swapaxes(m, -1, -2)))
    nc = m.shape[-1]
    return a - (np.trace(a, axis1=-2, axis2=-1) / nc)[..., None, None] * np.eye(nc)


@pytest.mark.parametrize("nc,L,N", [(3, (8, 8, 8, 8), 12), (3, (4, 4, 2, 4), 6), (2, (4, 2, 4, 4), 6),
                                    (3, (2, 3, 2, 2), 4), (3, (2, 4, 2, 3), 8), (2, (2, 2, 3, 2), 16),
                                    (3, (2, 2, 2, 3), 16), (3, (2, 2, 2, 2), 20), (2, (3, 2, 2, 2), 20),
                                    (3, (2, 2, 2, 4), 20), (2, (2, 2, 2, 4), 20),
                                    # orders between the compiled instances (run on the next one up)
                                    (3, (2, 2, 3, 2), 10), (2, (3, 2, 2, 2), 3), (3, (2, 2, 2, 2), 17),
                                    (3, (2, 2, 2, 3), 1)])
def test_adjoint_drift(nc, L, N):
    from oracle import oracle as O
    from prlgt2020_amd import nspt
    eps = 0.02
    V = int(np.prod(L))
    ne = nc * nc
    print("adjoint drift SU(%d) %s N=%d" % (nc, "x".join(map(str, L)), N), flush=True)
    U = O.random_gfield(L, nc, N, seed=7, scale=0.1)
    G, D = nspt.adjoint_drift(L, nc, N, U, eps)
    # the oracle's reverse pass over the whole lattice
    Go = O.gaction_grad(L, nc, N, U, N)
    assert np.abs(G - Go).max() < 1e-12 * np.abs(Go).max()
    # finite differences of the action (the six plaquettes of the link)
    rng = np.random.default_rng(5)
    h = 1e-6
    for t in range(12):
        site, mu = int(rng.integers(V)), int(rng.integers(4))
        col = int(rng.integers((N + 1) * ne))
        for unit, part in ((1.0, "re"), (1j, "im")):
            Up, Um = U.copy(), U.copy()
            Up[4 * site + mu, col] += h * unit
            Um[4 * site + mu, col] -= h * unit
            fd = (O.gaction(L, nc, N, Up, N, site, mu) - O.gaction(L, nc, N, Um, N, site, mu)) / (2 * h)
            g = G[4 * site + mu, col]
            want = g.real if part == "re" else g.imag
            assert abs(fd - want) < 1e-6 * max(1.0, abs(want)), (t, part, fd, want)
    # the staple drift: D = -eps/(2 nc) reH(U Sigma) per order, on a sample of links
    tau = -eps / (2 * nc)
    for l in rng.choice(4 * V, size=min(4 * V, 48), replace=False):
        site, mu = divmod(int(l), 4)
        S = O.gstaple_sum(L, nc, N, U, site, mu)
        W = O.gmul(nc, N, U[l], S).reshape(N + 1, nc, nc)
        want = tau * reh(W)
        assert np.abs(D[l].reshape(N + 1, nc, nc) - want).max() < 1e-12 * max(1e-3, np.abs(want).max())


@pytest.mark.parametrize("nc,L,N", [(3, (4, 4, 2, 4), 12), (3, (2, 2, 2, 4), 20), (2, (2, 4, 2, 2), 12)])
def test_adjoint_drift_equals_the_production_drift_on_an_nspt_field(nc, L, N):
    """On a field of the engine's own form (order 0 the flag times 1; here
    U = exp(A), A antihermitian at orders 1..N) the adjoint drift equals the
    production drift of nspt_drift (the staple sum of k_drift up to 16
    orders, the plaquette-field form above) at every order 1..N, and its
    order 0 (reH of the order-0 staple sum, a multiple of 1) is zero."""
    from oracle import oracle as O
    from prlgt2020_amd import nspt
    eps = 0.02
    ne = nc * nc
    U = O.random_field(L, nc, N, seed=2, scale=0.3)
    Ug = np.zeros((U.shape[0], (N + 1) * ne), np.complex128)
    Ug[:, :ne] = U[:, :1] * np.eye(nc).ravel()
    Ug[:, ne:] = U[:, 1:]
    _, Da = nspt.adjoint_drift(L, nc, N, Ug, eps)
    ctx = nspt.Context(L, nc, N)
    ctx.upload(U)
    Dp = ctx.drift(eps)
    ctx.close()
    scale = np.abs(Dp[:, 1:]).max()
    err = np.abs(Da[:, ne:] - Dp[:, 1:]).max() / scale
    print("SU(%d) %s N=%d: adjoint vs production drift %.2e" % (nc, "x".join(map(str, L)), N, err))
    assert err <= 1e-12
    assert np.abs(Da[:, :ne]).max() <= 1e-14 * scale
