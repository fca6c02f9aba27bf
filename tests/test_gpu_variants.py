"""Kernel builds other than an order's default: tile edge 3 and the
occupancy variants (4 links per wave at one and two waves per SIMD), chosen
at context creation through NSPT_TILE_EDGE / NSPT_LINKS_PER_WAVE /
NSPT_WAVES_PER_SIMD, as the batch autotuner chooses them.

Each build is one RK2 + zero modes + FA step checked as the default build
is: 1e-12 per order against the oracle on the evolved 16-order field of the
default build's test, and on larger evolved fields at 16 and 20 orders the
long-double accuracy bar; and
against the default build: the results differ (the build's own code ran --
builds of one kernel template once shared a single launch handle) by
rounding only.  A batch of replicas on a variant equals each
replica alone on that variant bit for bit, and the autotuner's candidate
list covers the occupancy variants of the edge it keeps.
"""
import os
from contextlib import contextmanager

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_parity import evolved_field
from tests.test_gpu_pins import err_vs_truth

pytestmark = pytest.mark.gpu

ENV = ("NSPT_TILE_EDGE", "NSPT_LINKS_PER_WAVE", "NSPT_WAVES_PER_SIMD")


def _nspt():
    from prlgt2020_amd import nspt
    return nspt


@contextmanager
def build_env(qt, lpw, wpe):
    old = {k: os.environ.get(k) for k in ENV}
    os.environ.update(dict(zip(ENV, (str(qt), str(lpw), str(wpe)))))
    try:
        yield
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def step_on(shape, L, nc, N, U, eta):
    nspt = _nspt()
    if shape is None:
        ctx = nspt.Context(L, nc, N)
    else:
        with build_env(*shape):
            ctx = nspt.Context(L, nc, N)
    got = ctx.kernel_shape()
    ctx.upload(U)
    ctx.set_noise(eta)
    obs = ctx.step(nspt.Context.params(0.02, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_HOST))
    out = ctx.download()
    ctx.close()
    return out, obs, got


# (tile edge, links per wave asked, waves per SIMD) -> links per wave run
VARIANTS = [((3, 0, 1), 7), ((4, 4, 1), 4), ((4, 4, 2), 4)]


@pytest.mark.parametrize("shape,lpb", VARIANTS)
def test_variant_step_matches_oracle_16_orders(shape, lpb):
    """The default build's 16-order case (test_orders_evolved_field: an
    evolved 2x2x2x4 field, one RK2 + zero modes + FA step with ranlxd noise,
    <= 1e-12 per order and in the observables) run on the variant."""
    from tests.conftest import rel_err_per_order
    from tests.test_gpu_parity import _run_pair
    nspt = _nspt()
    L, nc, N = (2, 2, 2, 4), 3, 16  # 32 sites: a ragged last block at 7 links per wave, 8 blocks at 4
    V = 32
    U = evolved_field(L, nc, N)
    with build_env(*shape):
        _run_pair(L, nc, N, 1, 1, 1, eps=0.02, relerr=rel_err_per_order, U=U)
    # the variant's own code ran: its bits are not the default build's
    eta = nspt.Ranlxd(97).noise(nc, 4 * V)
    Uv, _, got = step_on(shape, L, nc, N, U, eta)
    assert got == (shape[0], lpb, shape[2]), got
    Ud, _, got_d = step_on(None, L, nc, N, U, eta)
    assert got_d == (4, 7, 1), got_d
    assert not np.array_equal(Uv, Ud)


@pytest.mark.parametrize("shape,lpb,N", [((3, 0, 1), 7, 16), ((4, 4, 1), 4, 16), ((4, 4, 2), 4, 16),
                                         ((3, 0, 1), 7, 20), ((4, 4, 2), 4, 20)])
def test_variant_step_accuracy_on_evolved_fields(shape, lpb, N):
    """On an evolved field a high-order FP64 step is ill-conditioned (its
    orders above ~12 move by more than 1e-12 under a rounding of the input),
    so the bar is the accuracy against the long-double oracle step: at most
    twice the FP64 oracle's own error, as for the default build."""
    nspt = _nspt()
    L, nc = (2, 2, 4, 6), 3
    V = 96
    U = evolved_field(L, nc, N)
    eta = nspt.Ranlxd(4321).noise(nc, 4 * V)
    hi, lo = U.copy(), np.zeros_like(U)
    O.step_extended(L, nc, N, hi, lo, 0.02, 0.05, integrator=1, gauge_fix=1, zero_modes=1, eta=eta)
    Uc = U.copy()
    O.step(L, nc, N, Uc, 0.02, 0.05, integrator=1, update=0, gauge_fix=1, zero_modes=1, eta=eta)
    e_cpu = err_vs_truth(Uc, hi, lo, nc, N)
    bar = np.maximum(2 * e_cpu, 2e-15)
    Uv, _, got = step_on(shape, L, nc, N, U, eta)
    assert got == (shape[0], lpb, shape[2]), got
    Ud, _, _ = step_on(None, L, nc, N, U, eta)
    e, e_d = err_vs_truth(Uv, hi, lo, nc, N), err_vs_truth(Ud, hi, lo, nc, N)
    np.set_printoptions(precision=1, linewidth=200)
    print("\nFP64 oracle vs long double:", e_cpu)
    print("default build vs long double:", e_d)
    print("build", shape, "vs long double:", e)
    assert (e <= bar).all(), (e, bar)
    assert not np.array_equal(Uv, Ud)


def test_unknown_shape_runs_the_default_build():
    nspt = _nspt()
    with build_env(5, 3, 4):  # not compiled for 16 orders
        ctx = nspt.Context((2, 2, 2, 4), 3, 16)
    assert ctx.kernel_shape() == (4, 7, 1)
    ctx.close()


def test_batch_on_a_variant_equals_its_replicas_alone():
    nspt = _nspt()
    from tests.test_gpu_batch import _sweep
    N = 16
    Ls = [(4, 4, 4, 4), (6, 4, 4, 4), (4, 6, 4, 4), (4, 4, 6, 4)]
    ps = _sweep(nspt, nspt.RK2, nspt.GF_FA, 1, len(Ls))
    with build_env(4, 4, 2):
        b = nspt.Batch(Ls, 3, N, ps)
        singles = [nspt.Context(L, 3, N) for L in Ls]
    assert b.kernel_shape() == (4, 4, 2)
    b.set_form(2)
    obs = b.evolve(3)
    for r, (L, p, ctx) in enumerate(zip(Ls, ps, singles)):
        q = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
        o = ctx.evolve(q, 3)
        assert np.array_equal(obs[r], o), r
        assert np.array_equal(b.replicas[r].download(), ctx.download()), r
        ctx.close()
    b.close()


def test_tuner_times_the_occupancy_variants():
    nspt = _nspt()
    from tests.test_gpu_batch import _sweep
    N = 16
    Ls = [(4, 4, 4, 4), (6, 4, 4, 4), (4, 6, 4, 4), (4, 4, 6, 4)]
    ps = _sweep(nspt, nspt.RK2, nspt.GF_FA, 1, len(Ls))
    b = nspt.Batch(Ls, 3, N, ps)
    b.evolve(1, want_obs=False)
    before = [r.download() for r in b.replicas]
    q, ms = b.tune(steps=2)
    cands = b.tune_candidates()
    print("\ncandidates:", cands, "\nms:", ms[:len(cands)])
    assert len(cands) <= 16 and (ms[:len(cands)] > 0).all() and (ms[len(cands):] == 0).all()
    stage1 = [c for c in cands if c[2] == 0 and c[3] == 1]
    stage2 = cands[len(stage1):]
    assert {c[1] for c in stage1} == {3, 4}
    best1 = stage1[int(np.argmin(ms[:len(stage1)]))]
    if best1[1] == 4:
        assert sorted(c[2:] for c in stage2) == [(4, 1), (4, 2)]
        assert all(c[:2] == best1[:2] for c in stage2)
    best = cands[int(np.argmin(ms[:len(cands)]))]
    shape = b.kernel_shape()
    assert shape[0] == best[1] and shape[2] == best[3] and (best[2] == 0 or shape[1] == best[2])
    for x, r in zip(before, b.replicas):
        assert np.array_equal(x, r.download())
    b.evolve(1)
    b.close()
