"""The oracle pinned by mathematics: closed forms, invariants and brute-force
restatements written independently of oracle/nspt_oracle.cpp (plain numpy
polynomial arithmetic over matrix coefficients, numpy's FFT).  Each test is
chosen so that a plausible slip in the oracle (a dropped flag term, a wrong
sign or index, a transposed factor) fails it.  CPU only.
"""
import numpy as np
import pytest

from oracle import oracle as O

rng = np.random.default_rng(1234)


# ---------------------------------------------------------------- numpy model
def to_mats(s, nc, N):
    """Series -> list of N+1 matrices, order 0 = flag * identity."""
    ne = nc * nc
    out = [s[0] * np.eye(nc, dtype=complex)]
    for k in range(N):
        out.append(np.array(s[1 + k * ne:1 + (k + 1) * ne]).reshape(nc, nc))
    return out


def from_mats(M, nc):
    assert np.allclose(M[0], M[0][0, 0] * np.eye(nc), atol=1e-13)
    return np.concatenate([[M[0][0, 0]]] + [m.reshape(-1) for m in M[1:]])


def pmul(A, B):
    """Truncated polynomial product of matrix-coefficient series."""
    N = len(A) - 1
    return [sum(A[i] @ B[k - i] for i in range(k + 1)) for k in range(N + 1)]


def pdag(A):
    return [m.conj().T for m in A]


def rand_series(nc, N, flag=None, scale=1.0):
    s = scale * (rng.normal(size=1 + nc * nc * N) + 1j * rng.normal(size=1 + nc * nc * N))
    s[0] = rng.normal() + 1j * rng.normal() if flag is None else flag
    return s


def site_index(L, x):
    return ((x[0] * L[1] + x[1]) * L[2] + x[2]) * L[3] + x[3]


def coords(L, s):
    x3 = s % L[3]
    x2 = (s // L[3]) % L[2]
    x1 = (s // (L[3] * L[2])) % L[1]
    x0 = s // (L[3] * L[2] * L[1])
    return [x0, x1, x2, x3]


def shift(L, s, d, delta):
    x = coords(L, s)
    x[d] = (x[d] + delta) % L[d]
    return site_index(L, x)


# ---------------------------------------------------------------- product
@pytest.mark.parametrize("nc", [3, 2])
@pytest.mark.parametrize("N", [1, 3, 7])
@pytest.mark.parametrize("flags", [(0, 0), (1, 0), (0, 1), (0.3 - 1.2j, 2.0 + 0.5j)])
def test_product_is_polynomial_product(nc, N, flags):
    a, b = rand_series(nc, N, flags[0]), rand_series(nc, N, flags[1])
    got = O.mul(nc, N, a, b)
    want = from_mats(pmul(to_mats(a, nc, N), to_mats(b, nc, N)), nc)
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_product_associative():
    nc, N = 3, 6
    a, b, c = (rand_series(nc, N) for _ in range(3))
    l = O.mul(nc, N, O.mul(nc, N, a, b), c)
    r = O.mul(nc, N, a, O.mul(nc, N, b, c))
    assert np.abs(l - r).max() < 1e-11 * np.abs(l).max()


def test_scalar_series_reduce_to_polynomials():
    nc, N = 3, 8
    pa, pb = rng.normal(size=N + 1), rng.normal(size=N + 1)
    a = from_mats([x * np.eye(nc) for x in pa], nc)
    b = from_mats([x * np.eye(nc) for x in pb], nc)
    got = to_mats(O.mul(nc, N, a, b), nc, N)
    want = np.polynomial.polynomial.polymul(pa, pb)[:N + 1]
    for k in range(N + 1):
        assert np.allclose(got[k], want[k] * np.eye(nc), atol=1e-12)


# ---------------------------------------------------------------- exp / log
def algebra_series(nc, N, scale):
    return O.random_algebra(rng, nc, N, scale, 1)[0]


@pytest.mark.parametrize("nc", [3, 2])
def test_log_exp_inverse(nc):
    N = 10
    X = algebra_series(nc, N, 0.7)
    assert np.abs(O.log(nc, N, O.exp(nc, N, X)) - X).max() < 1e-12
    U = O.exp(nc, N, algebra_series(nc, N, 0.5))
    assert np.abs(O.exp(nc, N, O.log(nc, N, U)) - U).max() < 1e-12


def test_exp_commuting_generator_closed_form():
    """X = x(g) H with one fixed matrix H: exp(X)_k = sum_n [g^k] x(g)^n / n! H^n."""
    nc, N = 3, 9
    H = rng.normal(size=(nc, nc)) + 1j * rng.normal(size=(nc, nc))
    x = np.concatenate([[0.0], rng.normal(size=N)])  # x_0 = 0
    X = from_mats([x[k] * H for k in range(N + 1)], nc)
    got = to_mats(O.exp(nc, N, X), nc, N)
    P = np.polynomial.polynomial
    want = [np.eye(nc, dtype=complex)] + [np.zeros((nc, nc), complex) for _ in range(N)]
    xn = np.array([1.0])
    fact = 1.0
    for n in range(1, N + 1):
        xn = P.polymul(xn, x)[:N + 1]
        fact *= n
        Hn = np.linalg.matrix_power(H, n)
        for k in range(len(xn)):
            want[k] = want[k] + xn[k] / fact * Hn
    for k in range(N + 1):
        assert np.allclose(got[k], want[k], atol=1e-11 * max(1, np.abs(want[k]).max()))


def test_log_is_mercator_series():
    nc, N = 3, 7
    Xs = algebra_series(nc, N, 0.8)
    Xs[0] = 0
    X = to_mats(Xs, nc, N)
    U = from_mats([np.eye(nc)] + X[1:], nc)  # U = 1 + X
    want = [np.zeros((nc, nc), complex) for _ in range(N + 1)]
    Xp = [np.eye(nc, dtype=complex)] + [np.zeros((nc, nc), complex)] * N
    for k in range(1, N + 1):
        Xp = pmul(Xp, X)
        for o in range(N + 1):
            want[o] = want[o] + (-1) ** (k + 1) / k * Xp[o]
    got = to_mats(O.log(nc, N, U), nc, N)
    for o in range(N + 1):
        assert np.allclose(got[o], want[o], atol=1e-12)


@pytest.mark.parametrize("nc", [3, 2])
def test_exp_of_algebra_is_unitary_order_by_order(nc):
    N = 12
    U = to_mats(O.exp(nc, N, algebra_series(nc, N, 0.9)), nc, N)
    P = pmul(U, pdag(U))
    assert np.allclose(P[0], np.eye(nc), atol=1e-13)
    for k in range(1, N + 1):
        assert np.abs(P[k]).max() < 1e-11 * max(1.0, np.abs(U[k]).max() ** 2)


# ---------------------------------------------------------------- reH, Tr, dag
@pytest.mark.parametrize("nc", [3, 2])
def test_reh_antihermitian_traceless_idempotent(nc):
    N = 5
    s = rand_series(nc, N)
    r = O.reh(nc, N, s)
    assert r[0] == 0
    M = to_mats(r, nc, N)
    for k in range(1, N + 1):
        assert np.allclose(M[k], -M[k].conj().T, atol=1e-14)
        assert abs(np.trace(M[k])) < 1e-14
    assert np.allclose(O.reh(nc, N, r), r, atol=1e-15)
    # on an anti-Hermitian traceless input reH is the identity
    A = algebra_series(nc, N, 1.0)
    assert np.allclose(O.reh(nc, N, A), A, atol=1e-15)


@pytest.mark.parametrize("nc", [3, 2])
def test_trace_and_dagger(nc):
    N = 4
    s = rand_series(nc, N)
    M = to_mats(s, nc, N)
    t = O.tr(nc, N, s)
    assert np.isclose(t[0], nc * s[0])
    for k in range(1, N + 1):
        assert np.isclose(t[k], np.trace(M[k]))
    d = to_mats(O.dag(nc, N, s), nc, N)
    for k in range(N + 1):
        assert np.allclose(d[k], M[k].conj().T)


# ---------------------------------------------------------------- lattice stages
def _unitary_field(L, nc, N, seed, scale=0.4):
    return O.random_field(L, nc, N, seed=seed, scale=scale)


def _link(U, s, mu, nc, N):
    return to_mats(U[4 * s + mu], nc, N)


def _plaq_brute(L, nc, N, U):
    """sum_{x, mu<nu} Re Tr P_{mu nu}(x) / (6 nc V), the textbook plaquette."""
    V = int(np.prod(L))
    acc = np.zeros(N + 1, complex)
    for s in range(V):
        for mu in range(4):
            for nu in range(mu + 1, 4):
                P = pmul(pmul(_link(U, s, mu, nc, N), _link(U, shift(L, s, mu, 1), nu, nc, N)),
                         pmul(pdag(_link(U, shift(L, s, nu, 1), mu, nc, N)), pdag(_link(U, s, nu, nc, N))))
                acc += np.array([np.trace(m) for m in P])
    return acc.real / (6 * nc * V)


@pytest.mark.parametrize("nc", [3, 2])
def test_plaquette_is_textbook_plaquette(nc):
    L, N = (2, 2, 2, 3), 3
    U = _unitary_field(L, nc, N, seed=5)
    got = O.plaquette_measure(L, nc, N, U)
    want = _plaq_brute(L, nc, N, U)
    assert np.allclose(got.real, want, atol=1e-13)
    assert np.abs(got.imag).max() < 1e-13
    cold = O.plaquette_measure(L, nc, N, O.cold_field(L, nc, N))
    assert np.allclose(cold, np.eye(1, N + 1)[0], atol=1e-15)


def _gauge_transform(L, nc, N, U, G):
    """U_mu(x) -> G(x) U_mu(x) G(x+mu)^dag with numpy products."""
    V = int(np.prod(L))
    out = np.empty_like(U)
    for s in range(V):
        Gs = to_mats(G[s], nc, N)
        for mu in range(4):
            Gp = to_mats(G[shift(L, s, mu, 1)], nc, N)
            out[4 * s + mu] = from_mats(pmul(pmul(Gs, _link(U, s, mu, nc, N)), pdag(Gp)), nc)
    return out


def _random_gauge(L, nc, N, scale=0.5):
    V = int(np.prod(L))
    A = O.random_algebra(rng, nc, N, scale, V)
    return np.array([O.exp(nc, N, a) for a in A])


@pytest.mark.parametrize("nc", [3, 2])
def test_drift_gauge_covariant(nc):
    """D[U^G](x, mu) = G(x) D[U](x, mu) G(x)^dag."""
    L, N = (2, 2, 2, 3), 4
    V = int(np.prod(L))
    U = _unitary_field(L, nc, N, seed=11)
    G = _random_gauge(L, nc, N)
    D = O.drift(L, nc, N, U, 0.05)
    DG = O.drift(L, nc, N, _gauge_transform(L, nc, N, U, G), 0.05)
    for s in range(V):
        Gs = to_mats(G[s], nc, N)
        for mu in range(4):
            want = from_mats(pmul(pmul(Gs, to_mats(D[4 * s + mu], nc, N)), pdag(Gs)), nc)
            assert np.abs(DG[4 * s + mu] - want).max() < 1e-12


@pytest.mark.parametrize("nc", [3, 2])
def test_plaquette_gauge_invariant_and_fa_is_a_gauge_transform(nc):
    L, N = (2, 2, 2, 3), 4
    U = _unitary_field(L, nc, N, seed=12)
    p0 = O.plaquette_measure(L, nc, N, U)
    UG = _gauge_transform(L, nc, N, U, _random_gauge(L, nc, N))
    assert np.abs(O.plaquette_measure(L, nc, N, UG) - p0).max() < 1e-12
    UF = U.copy()
    O.fa_gauge_fixing(L, nc, N, UF, 0.05)
    assert np.abs(UF - U).max() > 1e-6  # it does change the links
    assert np.abs(O.plaquette_measure(L, nc, N, UF) - p0).max() < 1e-12
    US = U.copy()
    O.sgf_default(L, nc, N, US, 0.05)
    assert np.abs(O.plaquette_measure(L, nc, N, US) - p0).max() < 1e-12


@pytest.mark.parametrize("nc", [3, 2])
def test_cold_start_first_jacobi_step_closed_form(nc):
    """Staples of the cold field sum to 6: zero drift, U' = exp(-sqrt(eps) eta)."""
    L, N, eps = (2, 2, 2, 2), 5, 0.05
    V = int(np.prod(L))
    U = O.cold_field(L, nc, N)
    eta = O.Ranlxd(1804289383).noise(nc, 4 * V)
    O.step(L, nc, N, U, eps, integrator=0, update=0, gauge_fix=0, zero_modes=0, eta=eta)
    for l in range(4 * V):
        W = -np.sqrt(eps) * eta[l].reshape(nc, nc)
        want = [np.eye(nc, dtype=complex)] + [np.linalg.matrix_power(W, k) / np.prod(range(1, k + 1))
                                               for k in range(1, N + 1)]
        assert np.abs(U[l] - from_mats(want, nc)).max() < 1e-14


@pytest.mark.parametrize("nc", [3, 2])
def test_zero_mode_subtraction_removes_the_mean_log(nc):
    L, N = (2, 2, 2, 3), 4
    V = int(np.prod(L))
    U = _unitary_field(L, nc, N, seed=13)
    MM = np.zeros((4, O.ser_len(nc, N)), complex)
    for s in range(V):
        for mu in range(4):
            MM[mu] += O.log(nc, N, U[4 * s + mu])
    assert np.abs(MM).max() > 1e-3
    O.zero_modes(L, nc, N, U, MM)
    S = np.zeros_like(MM)
    for s in range(V):
        for mu in range(4):
            S[mu] += O.log(nc, N, U[4 * s + mu])
    assert np.abs(S).max() < 1e-12


@pytest.mark.parametrize("sign", [-1, 1])
def test_dft_matches_numpy_fft(sign):
    L = (2, 3, 4, 5)
    V = int(np.prod(L))
    howmany, stride = 3, 4
    data = rng.normal(size=(V, stride)) + 1j * rng.normal(size=(V, stride))
    got = O.dft_many(L, data.copy(), howmany, stride, sign).reshape(V, stride)
    x = data[:, :howmany].reshape(*L, howmany)
    want = np.fft.fftn(x, axes=(0, 1, 2, 3)) if sign == -1 else np.fft.ifftn(x, axes=(0, 1, 2, 3)) * V
    assert np.allclose(got[:, :howmany], want.reshape(V, howmany), atol=1e-12)
    assert np.array_equal(got[:, howmany:], data[:, howmany:])


def test_rk2_differs_from_euler_at_second_order():
    """Without noise both schemes agree to O(eps): their one-step difference
    scales as eps^2 (ratio 4 when eps halves)."""
    L, nc, N = (2, 2, 2, 2), 3, 4
    V = int(np.prod(L))
    U0 = _unitary_field(L, nc, N, seed=21, scale=0.6)
    zero = np.zeros((4 * V, nc * nc), complex)
    diffs = []
    for eps in (0.02, 0.01, 0.005):
        Ue, Ur = U0.copy(), U0.copy()
        O.step(L, nc, N, Ue, eps, integrator=0, update=0, gauge_fix=0, zero_modes=0, eta=zero)
        O.step(L, nc, N, Ur, eps, integrator=1, update=0, gauge_fix=0, zero_modes=0, eta=zero)
        diffs.append(np.abs(Ur - Ue).max())
    assert diffs[0] > 1e-8
    for a, b in zip(diffs, diffs[1:]):
        assert 3.5 < a / b < 4.5, diffs


def test_canonical_flop_counts_match_enumeration():
    """SP and EX of bench.py are the trip counts of QCDpt.h:95-99, 181-187."""
    import bench
    for N in (2, 6, 12, 20, 32):
        pairs = sum(1 for i in range(1, N + 1) for j in range(1, N + 1) if i + j <= N)
        assert bench.sp_flops(N) == 216 * pairs + 144 * N
        # exp: A^k = A^{k-1} A with A^{k-1} starting at order k-1
        exp_mm = sum(1 for k in range(2, N + 1) for i in range(k - 1, N + 1) for j in range(1, N + 1)
                     if i + j <= N)
        assert exp_mm == (N - 1) * N * (N + 1) // 6
        assert bench.ex_flops(N) == 216 * exp_mm + 36 * N * (N - 1)
