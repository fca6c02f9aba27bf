"""GPU side of the oracle pins and of the accuracy statement at high orders.

- FA normalisation: the exact order-1 identity div A'(k) = (1 - alpha) div A(k)
  (tests/test_oracle_pins.py) checked on the output of the HIP FA gauge
  fixing itself, including the 16^4 lattice whose lines are the headline
  workload's.
- Orders 20, 24, 32: one RK2 + zero modes + FA step in FP64 is ill-conditioned
  there (a rounding of the input moves order 20 by ~1e-11), so "<= 1e-12 of
  the CPU step" is not a property either FP64 step can have.  The accuracy
  statement instead: against the same step computed in x87 long double by
  the oracle (oracle/_build/liboracle_ext.so), the GPU's error per order is
  at most twice the FP64 oracle's own error (or 2e-15 of that order's scale,
  whichever is larger), for both association forms of the drift.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_oracle_pins import fa_divergence_ratio_error, order1_field

pytestmark = pytest.mark.gpu


def _nspt():
    from prlgt2020_amd import nspt
    return nspt


@pytest.mark.parametrize("nc,L", [(3, (3, 5, 2, 7)), (2, (4, 3, 3, 5)), (3, (2, 4, 6, 4)), (3, (16, 16, 16, 16)),
                                  (2, (16, 2, 20, 10))])
def test_gpu_fa_order1_divergence_closed_form(nc, L):
    nspt = _nspt()
    N, alpha = 3, 0.05
    U, _ = order1_field(L, nc, N, seed=11)
    ctx = nspt.Context(L, nc, N)
    ctx.upload(U)
    ctx.fa_gauge_fixing(alpha)
    e, e0 = fa_divergence_ratio_error(L, nc, U, ctx.download(), alpha)
    ctx.close()
    print("L", L, "k != 0: %.2e, k = 0: %.2e" % (e, e0))
    assert e <= 1e-13 and e0 <= 1e-13, (e, e0)


def err_vs_truth(a, hi, lo, nc, N):
    """Per order k = 1..N: max |a - (hi + lo)| / max |hi| over the field."""
    ne = nc * nc
    d = (a - hi) - lo
    out = []
    for k in range(1, N + 1):
        sl = slice(1 + (k - 1) * ne, 1 + k * ne)
        out.append(np.abs(d[:, sl]).max() / np.abs(hi[:, sl]).max())
    return np.array(out)


def gpu_step(L, nc, N, U, eta, form):
    nspt = _nspt()
    old = os.environ.get("NSPT_DRIFT_FORM")
    os.environ["NSPT_DRIFT_FORM"] = form
    try:
        ctx = nspt.Context(L, nc, N)
    finally:
        if old is None:
            del os.environ["NSPT_DRIFT_FORM"]
        else:
            os.environ["NSPT_DRIFT_FORM"] = old
    ctx.upload(U)
    ctx.set_noise(eta)
    ctx.step(nspt.Context.params(0.02, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_HOST), want_obs=False)
    out = ctx.download()
    ctx.close()
    return out


@pytest.mark.parametrize("nc,N", [(3, 20), (3, 24), (3, 32), (2, 20), (2, 24)])
def test_high_orders_accuracy_against_extended_precision(nc, N):
    from tests.test_gpu_parity import evolved_field
    nspt = _nspt()
    L = (2, 2, 2, 4)
    V = 32
    U = evolved_field(L, nc, N)
    eta = nspt.Ranlxd(4321).noise(nc, 4 * V)
    hi, lo = U.copy(), np.zeros_like(U)
    O.step_extended(L, nc, N, hi, lo, 0.02, 0.05, integrator=1, gauge_fix=1, zero_modes=1, eta=eta)
    Uc = U.copy()
    O.step(L, nc, N, Uc, 0.02, 0.05, integrator=1, update=0, gauge_fix=1, zero_modes=1, eta=eta)
    e_cpu = err_vs_truth(Uc, hi, lo, nc, N)
    bar = np.maximum(2 * e_cpu, 2e-15)
    np.set_printoptions(precision=1, linewidth=200)
    print("\nFP64 oracle vs long double:", e_cpu)
    for form in ("plaquettes", "staples"):
        e_gpu = err_vs_truth(gpu_step(L, nc, N, U, eta, form), hi, lo, nc, N)
        print("GPU %-10s vs long double:" % form, e_gpu)
        assert (e_gpu <= bar).all(), (form, e_gpu, bar)


def test_headline_step_16x4_N20():
    """One RK2 + zero modes + FA step of the headline workload's lattice and
    orders (SU(3) 16^4, N = 20, eps 0.02, alpha 0.05) on a field evolved 20
    steps from the cold start, with ranlxd host noise: the GPU step's error
    per order against the long-double oracle step is at most twice the FP64
    oracle's own (or 2e-15 of the order's scale), and the two FP64 steps
    agree to 1e-12 relative at every order where the FP64 oracle itself is
    that accurate.  The oracle runs on the host's cores (OpenMP)."""
    import time
    nspt = _nspt()
    L, nc, N = (16, 16, 16, 16), 3, 20
    V = 16 ** 4
    ctx = nspt.Context(L, nc, N)
    ctx.evolve(nspt.Context.params(0.02, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_DEVICE, seed=3), 20,
               want_obs=False)
    U = ctx.download()
    eta = nspt.Ranlxd(1804289383).noise(nc, 4 * V)
    ctx.set_noise(eta)
    obs_g = ctx.step(nspt.Context.params(0.02, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_HOST))
    Ug = ctx.download()
    ctx.close()
    t = time.time()
    Uc = U.copy()
    obs_c = O.step(L, nc, N, Uc, 0.02, 0.05, integrator=1, update=0, gauge_fix=1, zero_modes=1, eta=eta)
    print("\nFP64 oracle step: %.1f s" % (time.time() - t), flush=True)
    t = time.time()
    hi, lo = U.copy(), np.zeros_like(U)
    O.step_extended(L, nc, N, hi, lo, 0.02, 0.05, integrator=1, gauge_fix=1, zero_modes=1, eta=eta)
    print("long-double oracle step: %.1f s" % (time.time() - t), flush=True)
    e_cpu = err_vs_truth(Uc, hi, lo, nc, N)
    e_gpu = err_vs_truth(Ug, hi, lo, nc, N)
    from tests.conftest import rel_err_per_order
    e_gc = rel_err_per_order(Ug, Uc, nc, N)[1:]
    np.set_printoptions(precision=1, linewidth=200)
    print("FP64 oracle vs long double:", e_cpu)
    print("GPU vs long double:        ", e_gpu)
    print("GPU vs FP64 oracle:        ", e_gc)
    assert (e_gpu <= np.maximum(2 * e_cpu, 2e-15)).all()
    assert (e_gc[e_cpu < 2.5e-13] <= 1e-12).all()
    assert np.abs(obs_g - obs_c).max() <= 1e-12 * np.abs(obs_c).max()
