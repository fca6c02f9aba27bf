"""Pin of the FA gauge-fixing normalisation, which no reference test or
golden vector covers (the RK2 coefficient's pin is test_oracle_rk2_pin.py).
The pin is shown to fail on an oracle with a planted mistake
(oracle.planted), so a wrong oracle could not pass it.

FA (nspt.cc:871-916, p^2 of lattice.h:1144-1168).  Take U = exp(A) with A an
antihermitian traceless field at order 1 only.  At order 1 one FA step is
  A'_mu(x) = A_mu(x) + W(x) - W(x+mu),  W = IDFT[-alpha/p^2 DFT[div A]]/V,
so the backward divergence div A(x) = sum_mu A_mu(x) - A_mu(x-mu) becomes
  div A'(k) = div A(k) + p^2(k) W(k) = (1 - alpha) div A(k)
for every k != 0 (and 0 at k = 0), p^2(k) = sum_mu 4 sin^2(k_mu/2) being the
symbol of the lattice Laplacian, whatever the DFT sign convention.  A
flipped alpha gives 1 + alpha, a dropped 1/V gives 1 - alpha V, a wrong p^2
a k-dependent factor.
"""
import numpy as np
import pytest

from oracle import oracle as O

def order1_field(L, nc, N, seed):
    """U = exp(A), A random antihermitian traceless at order 1 only (4V
    links); returns (U, A) with A of shape L + (4, nc, nc)."""
    rng = np.random.default_rng(seed)
    V = int(np.prod(L))
    M = rng.normal(size=(4 * V, nc, nc)) + 1j * rng.normal(size=(4 * V, nc, nc))
    A = 0.5 * (M - np.conj(np.swapaxes(M, 1, 2)))
    A -= np.trace(A, axis1=1, axis2=2)[:, None, None] * np.eye(nc) / nc
    U = np.zeros((4 * V, 1 + nc * nc * N), np.complex128)
    U[:, 0] = 1.0
    P = np.broadcast_to(np.eye(nc), A.shape).astype(np.complex128)
    fact = 1.0
    for k in range(1, N + 1):
        P = P @ A
        fact *= k
        U[:, 1 + (k - 1) * nc * nc:1 + k * nc * nc] = (P / fact).reshape(4 * V, nc * nc)
    return U, A.reshape(tuple(L) + (4, nc, nc))


def divergence_ft(A):
    """DFT over the four lattice axes of sum_mu A_mu(x) - A_mu(x - mu)."""
    div = np.zeros(A.shape[:4] + A.shape[5:], np.complex128)
    for mu in range(4):
        a = A[..., mu, :, :]
        div += a - np.roll(a, 1, axis=mu)
    return np.fft.fftn(div, axes=(0, 1, 2, 3))


def order1(U, L, nc):
    return U[:, 1:1 + nc * nc].reshape(tuple(L) + (4, nc, nc))


def fa_divergence_ratio_error(L, nc, U0, U1, alpha):
    """max over k != 0 of |divA'(k) - (1 - alpha) divA(k)| / max |divA|, and
    max |divA'(0)| / max |divA|."""
    D0 = divergence_ft(order1(U0, L, nc))
    D1 = divergence_ft(order1(U1, L, nc))
    scale = np.abs(D0).max()
    err = np.abs(D1 - (1 - alpha) * D0)
    err[0, 0, 0, 0] = 0.0
    return err.max() / scale, np.abs(D1[0, 0, 0, 0]).max() / scale


@pytest.mark.parametrize("nc,L", [(3, (3, 5, 2, 7)), (2, (4, 3, 3, 5)), (3, (2, 4, 6, 4))])
def test_fa_order1_divergence_closed_form(nc, L):
    N, alpha = 3, 0.05
    U, _ = order1_field(L, nc, N, seed=11)
    U1 = U.copy()
    O.fa_gauge_fixing(L, nc, N, U1, alpha)
    e, e0 = fa_divergence_ratio_error(L, nc, U, U1, alpha)
    assert e <= 1e-13 and e0 <= 1e-13, (e, e0)


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_fa_pin_fails_on_planted_mistakes(variant):
    """alpha sign flipped, 1/V dropped, p^2 = 4 sin^2(p): the closed form
    catches each."""
    nc, L, N, alpha = 3, (3, 5, 2, 7), 3, 0.05
    U, _ = order1_field(L, nc, N, seed=11)
    U1 = U.copy()
    with O.planted(fa=variant):
        O.fa_gauge_fixing(L, nc, N, U1, alpha)
    e, _ = fa_divergence_ratio_error(L, nc, U, U1, alpha)
    assert e > 1e-3, (variant, e)


@pytest.mark.parametrize("nc,integrator", [(3, 1), (2, 0)])
def test_extended_oracle_is_the_same_step(nc, integrator):
    """The long-double build of the oracle (the accuracy reference of FP64
    steps at high orders) computes the same step: at 6 orders, where one
    FP64 step is well conditioned, it agrees with the FP64 oracle to a few
    roundings per order, fields and observables."""
    from tests.conftest import rel_err_per_order
    L, N = (2, 4, 2, 3), 6
    V = 48
    U = O.random_field(L, nc, N, seed=4, scale=0.4)
    eta = O.Ranlxd(3).noise(nc, 4 * V)
    Ud, hi, lo = U.copy(), U.copy(), np.zeros_like(U)
    od = O.step(L, nc, N, Ud, 0.02, 0.05, integrator=integrator, update=0, gauge_fix=1, zero_modes=1, eta=eta)
    oe = O.step_extended(L, nc, N, hi, lo, 0.02, 0.05, integrator=integrator, gauge_fix=1, zero_modes=1, eta=eta)
    assert rel_err_per_order(Ud, hi + lo, nc, N).max() <= 2e-14
    assert np.abs(lo).max() > 0  # the extended result carries digits beyond FP64
    assert np.abs(od - oe).max() <= 1e-14 * np.abs(od).max()


def test_long_run_fixture_of_the_cpu_runnable_workload_reproduces():
    """The first steps of run 0 of tests/golden/long_runs_config1.npz (SU(3)
    4^4, 6 orders, Euler eps 0.05, zero modes + FA), recomputed now: the
    long CPU runs the GPU statistics are compared with are this oracle's."""
    import os
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "long_runs_config1.npz"))
    L, N = tuple(int(x) for x in d["L"]), int(d["N"])
    U = O.cold_field(L, 3, N)
    st = d["stamps"]
    got = O.run_stats(L, 3, N, U, float(d["eps"]), float(d["alpha"]), 0, 1, 1, int(d["seeds"][0]), len(st), len(st))
    assert np.abs(got - st).max() <= 1e-12
