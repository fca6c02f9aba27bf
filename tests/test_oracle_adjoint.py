"""The oracle's reverse-mode gradient of the truncated Wilson action
(generalised series with a Haar-random SU(N) order 0): against central finite
differences (h = 1e-6) and against the staple sum it must reproduce (the
Langevin drift is tau_g reH(U Sigma)).  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O


def to_mats(s, nc, N):
    return [np.array(s[k * nc * nc:(k + 1) * nc * nc]).reshape(nc, nc) for k in range(N + 1)]


@pytest.mark.parametrize("nc", [3, 2])
def test_generalised_product_is_polynomial_product(nc):
    rng = np.random.default_rng(3)
    N = 5
    a = rng.normal(size=O.gser_len(nc, N)) + 1j * rng.normal(size=O.gser_len(nc, N))
    b = rng.normal(size=O.gser_len(nc, N)) + 1j * rng.normal(size=O.gser_len(nc, N))
    A, B = to_mats(a, nc, N), to_mats(b, nc, N)
    C = to_mats(O.gmul(nc, N, a, b), nc, N)
    for k in range(N + 1):
        assert np.allclose(C[k], sum(A[i] @ B[k - i] for i in range(k + 1)), atol=1e-13)


def test_random_gfield_order0_is_haar_su_n():
    U = O.random_gfield((2, 2, 2, 2), 3, 2, seed=7)
    m = U[:, :9].reshape(-1, 3, 3)
    assert np.allclose(m @ np.conj(np.swapaxes(m, 1, 2)), np.eye(3), atol=1e-13)
    assert np.allclose(np.linalg.det(m), 1.0, atol=1e-13)


@pytest.mark.parametrize("nc,L,N,K", [(3, (2, 2, 2, 3), 4, 4), (2, (2, 3, 2, 2), 5, 3), (3, (8, 8, 8, 8), 12, 12)])
def test_gradient_matches_central_differences(nc, L, N, K):
    V = int(np.prod(L))
    U = O.random_gfield(L, nc, N, seed=7, scale=0.1)
    G = O.gaction_grad(L, nc, N, U, K) if V <= 64 else None
    rng = np.random.default_rng(11)
    h = 1e-6
    ne = nc * nc
    for t in range(24):
        site, mu = int(rng.integers(V)), int(rng.integers(4))
        j, e = int(rng.integers(K + 1)), int(rng.integers(ne))
        col = j * ne + e
        if G is None:  # large lattice: gradient of the link's six plaquettes only
            g = _local_grad(L, nc, N, U, K, site, mu)
            gv = g[col]
        else:
            gv = G[4 * site + mu, col]
        for part, unit in (("re", 1.0), ("im", 1j)):
            Up, Um = U.copy(), U.copy()
            Up[4 * site + mu, col] += h * unit
            Um[4 * site + mu, col] -= h * unit
            fd = (O.gaction(L, nc, N, Up, K, site, mu) - O.gaction(L, nc, N, Um, K, site, mu)) / (2 * h)
            want = gv.real if part == "re" else gv.imag
            assert abs(fd - want) < 1e-6 * max(1.0, abs(want)), (t, part, fd, want)


def _local_grad(L, nc, N, U, K, site, mu):
    """Gradient w.r.t. one link from a small lattice holding its neighbourhood
    is not available; use the staple identity G_j = Sigma_{K-j}^dag instead."""
    S = to_mats(O.gstaple_sum(L, nc, N, U, site, mu), nc, N)
    return np.concatenate([S[K - j].conj().T.reshape(-1) if j <= K else np.zeros(nc * nc)
                           for j in range(N + 1)])


@pytest.mark.parametrize("nc", [3, 2])
def test_gradient_is_the_daggered_staple_sum(nc):
    """G_{l, j} = (Sigma_l)^dag_{K-j} for j <= K and 0 above: one reverse pass
    seeded at order K yields every order of the staple sum."""
    L, N, K = (2, 2, 3, 2), 5, 5
    V = int(np.prod(L))
    U = O.random_gfield(L, nc, N, seed=9, scale=0.2)
    G = O.gaction_grad(L, nc, N, U, K)
    for site in range(V):
        for mu in range(4):
            S = to_mats(O.gstaple_sum(L, nc, N, U, site, mu), nc, N)
            g = to_mats(G[4 * site + mu], nc, N)
            for j in range(N + 1):
                want = S[K - j].conj().T if j <= K else np.zeros((nc, nc))
                assert np.allclose(g[j], want, atol=1e-13), (site, mu, j)
