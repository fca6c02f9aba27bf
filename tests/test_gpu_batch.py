"""Batched replicas (nspt_batch_*): a graph-replayed step of several
replicas of different volume, time step, integrator and seed gives the same
fields and observables as each replica evolved alone with nspt_evolve."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _replicas(nspt):
    Ls = [(4, 4, 4, 4), (4, 4, 6, 4), (6, 4, 4, 4)]
    ps = [nspt.Context.params(0.02, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_DEVICE, seed=11, step=0),
          nspt.Context.params(0.01, 0.05, nspt.EULER, nspt.GF_FA, 1, nspt.NOISE_DEVICE, seed=12, step=0),
          nspt.Context.params(0.03, 0.05, nspt.RK2, nspt.GF_NONE, 0, nspt.NOISE_DEVICE, seed=13, step=5)]
    return Ls, ps


@pytest.mark.parametrize("nc", [3, 2])
def test_batch_equals_single_replicas(nc):
    from prlgt2020_amd import nspt
    N = 6
    Ls, ps = _replicas(nspt)
    print("batch of %d SU(%d) replicas, N=%d" % (len(Ls), nc, N), flush=True)
    b = nspt.Batch(Ls, nc, N, ps)
    obs_b = b.evolve(4)
    obs_b2 = b.evolve(3)
    for r, (L, p) in enumerate(zip(Ls, ps)):
        c = nspt.Context(L, nc, N)
        q = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
        o1 = c.evolve(q, 4)
        o2 = c.evolve(q, 3)
        assert np.array_equal(b.replicas[r].download(), c.download()), r
        assert np.array_equal(obs_b[r], o1) and np.array_equal(obs_b2[r], o2), r
        c.close()
    b.close()


def test_batch_without_observables_and_long_run():
    from prlgt2020_amd import nspt
    N = 4
    Ls, ps = _replicas(nspt)
    b = nspt.Batch(Ls, 3, N, ps)
    b.evolve(300, want_obs=False)
    obs = b.evolve(300)  # more steps than one observable chunk (256)
    for r, (L, p) in enumerate(zip(Ls, ps)):
        c = nspt.Context(L, 3, N)
        q = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
        c.evolve(q, 300, want_obs=False)
        o = c.evolve(q, 300)
        assert np.array_equal(b.replicas[r].download(), c.download()), r
        assert np.array_equal(obs[r], o), r
        c.close()
    b.close()


def test_tuner_restores_the_run():
    from prlgt2020_amd import nspt
    N = 16  # compiled with tile edges 4 and 3
    Ls, ps = _replicas(nspt)
    b = nspt.Batch(Ls, 3, N, ps)
    b.evolve(2, want_obs=False)
    before = [r.download() for r in b.replicas]
    q, ms = b.tune(steps=2)
    assert q in (3, 4) and ms[0] > 0 and ms[1] > 0
    after = [r.download() for r in b.replicas]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    # the run continues with the chosen tiles from the restored state and counters
    obs = b.evolve(2)
    ref = nspt.Context(Ls[0], 3, N)
    p = ps[0]
    q0 = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
    ref.evolve(q0, 2, want_obs=False)
    o = ref.evolve(q0, 2)
    # a different tile edge only changes rounding
    assert np.abs(obs[0] - o).max() < 1e-12
    b.close()
    ref.close()


def test_batch_replicas_refuse_action_and_twist():
    """A replica's step is replayed from graphs captured at its first
    evolve: the action and the boundary conditions cannot change under it."""
    from prlgt2020_amd import nspt
    Ls, ps = _replicas(nspt)
    b = nspt.Batch(Ls[:1], 3, 4, ps[:1])
    rep = b.replicas[0]
    with pytest.raises(nspt.NsptError, match="Wilson action"):
        rep.set_action(-1.0 / 12)
    with pytest.raises(nspt.NsptError, match="periodic"):
        rep.set_twist(0b0110)
    b.close()


def test_batch_replicas_refuse_their_own_steps():
    """A replica steps only through the batch: its captured graphs track its
    U/F buffer roles and observable rows.  The run goes on unharmed."""
    from prlgt2020_amd import nspt
    Ls, ps = _replicas(nspt)
    b = nspt.Batch(Ls[:1], 3, 4, ps[:1])
    b.evolve(2)
    rep = b.replicas[0]
    p = ps[0]
    q = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, 100)
    for call in (lambda: rep.step(q), lambda: rep.evolve(q, 1), lambda: rep.gauge_wilson(q),
                 lambda: rep.set_noise(np.zeros((4 * rep.V, 9), np.complex128))):
        with pytest.raises(nspt.NsptError, match="replica of a batch"):
            call()
    obs = b.evolve(1)
    ref = nspt.Context(Ls[0], 3, 4)
    q0 = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
    ref.evolve(q0, 2, want_obs=False)
    o = ref.evolve(q0, 1)
    assert np.array_equal(b.replicas[0].download(), ref.download())
    assert np.array_equal(obs[0], o)
    ref.close()
    b.close()


def test_eps_sweep_batch_full_size():
    """The batched epsilon sweep at its real size: SU(3) 4 x 12^4 + 4 x 16^4,
    16 orders, RK2 + zero modes + FA, eps {0.01, 0.015, 0.02, 0.03} per
    volume, distinct seeds.  Three steps of the batch equal, bit for bit,
    three steps of each replica alone (fields and observables)."""
    from prlgt2020_amd import nspt
    N = 16
    Ls = [(12, 12, 12, 12)] * 4 + [(16, 16, 16, 16)] * 4
    eps = [0.01, 0.015, 0.02, 0.03] * 2
    ps = [nspt.Context.params(e, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_DEVICE, seed=100 + r, step=0)
          for r, e in enumerate(eps)]
    print("batch of 8 SU(3) replicas 4 x 12^4 + 4 x 16^4, N=16", flush=True)
    b = nspt.Batch(Ls, 3, N, ps)
    obs_b = b.evolve(3)
    for r, (L, p) in enumerate(zip(Ls, ps)):
        c = nspt.Context(L, 3, N)
        q = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
        o = c.evolve(q, 3)
        assert np.array_equal(b.replicas[r].download(), c.download()), r
        assert np.array_equal(obs_b[r], o), r
        c.close()
    b.close()


def _sweep(nspt, integrator, gauge_fix, zero_modes, n):
    eps = [0.01, 0.02, 0.015, 0.03, 0.025][:n]
    return [nspt.Context.params(e, 0.05, integrator, gauge_fix, zero_modes, nspt.NOISE_DEVICE, seed=40 + r, step=r)
            for r, e in enumerate(eps)]


@pytest.mark.parametrize("nc,N,integrator,gf,zm", [(3, 6, 1, 1, 1), (3, 20, 1, 1, 1), (2, 6, 0, 0, 1), (3, 4, 0, 0, 0),
                                                   (3, 16, 0, 1, 0), (2, 20, 1, 1, 1)])
def test_table_form_equals_single_replicas(nc, N, integrator, gf, zm):
    """One launch per stage over the replica table (uneven volumes, odd
    extents, eps and seeds per replica): fields and observables equal, bit
    for bit, each replica evolved alone; the staple drift (N <= 16) and the
    plaquette drift (N = 20), RK2 and Euler, FA or no gauge fixing, with and
    without zero modes."""
    from prlgt2020_amd import nspt
    Ls = [(4, 4, 4, 4), (4, 4, 6, 4), (6, 4, 4, 4), (2, 2, 2, 4), (3, 4, 5, 2)]
    ps = _sweep(nspt, integrator, gf, zm, len(Ls))
    print("table batch of %d SU(%d) replicas, N=%d" % (len(Ls), nc, N), flush=True)
    b = nspt.Batch(Ls, nc, N, ps)
    b.set_form(True)
    assert b.table
    obs_b = b.evolve(4)
    obs_b2 = b.evolve(3)
    for r, (L, p) in enumerate(zip(Ls, ps)):
        c = nspt.Context(L, nc, N)
        q = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
        o1 = c.evolve(q, 4)
        o2 = c.evolve(q, 3)
        assert np.array_equal(b.replicas[r].download(), c.download()), r
        assert np.array_equal(obs_b[r], o1) and np.array_equal(obs_b2[r], o2), r
        c.close()
    b.close()


@pytest.mark.parametrize("nc,N,groups", [(3, 6, 2), (3, 6, 5), (2, 20, 3), (3, 16, 4)])
def test_grouped_table_form_equals_single_replicas(nc, N, groups):
    """The table form split into concurrent groups of replicas (each group's
    launches a graph branch of its own): fields and observables equal, bit
    for bit, each replica evolved alone."""
    from prlgt2020_amd import nspt
    Ls = [(4, 4, 4, 4), (4, 4, 6, 4), (6, 4, 4, 4), (2, 2, 2, 4), (3, 4, 5, 2)]
    ps = _sweep(nspt, nspt.RK2, nspt.GF_FA, 1, len(Ls))
    print("table batch of %d SU(%d) replicas in %d groups, N=%d" % (len(Ls), nc, groups, N), flush=True)
    b = nspt.Batch(Ls, nc, N, ps)
    b.set_form(groups)
    assert b.table and b.form == groups
    obs_b = b.evolve(4)
    obs_b2 = b.evolve(3)
    for r, (L, p) in enumerate(zip(Ls, ps)):
        c = nspt.Context(L, nc, N)
        q = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
        o1 = c.evolve(q, 4)
        o2 = c.evolve(q, 3)
        assert np.array_equal(b.replicas[r].download(), c.download()), r
        assert np.array_equal(obs_b[r], o1) and np.array_equal(obs_b2[r], o2), r
        c.close()
    with pytest.raises(nspt.NsptError, match="form"):
        b.set_form(9)
    b.close()


def test_table_form_refused_for_mixed_replicas_and_forms_switch():
    """Replicas with different integrators or gauge fixings cannot share one
    launch per stage; switching the form of a run keeps it exact."""
    from prlgt2020_amd import nspt
    Ls, ps = _replicas(nspt)
    b = nspt.Batch(Ls, 3, 4, ps)
    with pytest.raises(nspt.NsptError, match="differ"):
        b.set_form(True)
    b.close()
    Ls = [(4, 4, 4, 4), (2, 4, 2, 4)]
    ps = _sweep(nspt, nspt.RK2, nspt.GF_FA, 1, 2)
    b = nspt.Batch(Ls, 3, 6, ps)
    b.evolve(2)
    b.set_form(True)
    b.evolve(3)
    b.set_form(False)
    obs = b.evolve(2)
    for r, (L, p) in enumerate(zip(Ls, ps)):
        c = nspt.Context(L, 3, 6)
        q = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
        c.evolve(q, 5, want_obs=False)
        o = c.evolve(q, 2)
        assert np.array_equal(b.replicas[r].download(), c.download()), r
        assert np.array_equal(obs[r], o), r
        c.close()
    b.close()


def test_tuner_over_forms_restores_the_run():
    """The tuner times both forms and the compiled tile edges, then the run
    goes on from the restored fields and counters with the chosen ones."""
    from prlgt2020_amd import nspt
    N = 16
    Ls = [(4, 4, 4, 4), (6, 4, 4, 4), (4, 6, 4, 4)]
    ps = _sweep(nspt, nspt.RK2, nspt.GF_FA, 1, 3)
    b = nspt.Batch(Ls, 3, N, ps)
    b.evolve(2, want_obs=False)
    before = [r.download() for r in b.replicas]
    q, ms = b.tune(steps=2)
    assert (ms[:4] > 0).all()  # branches and table, tile edges 4 and 3
    after = [r.download() for r in b.replicas]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    obs = b.evolve(2)
    ref = nspt.Context(Ls[1], 3, N)
    p = ps[1]
    q1 = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
    ref.evolve(q1, 2, want_obs=False)
    o = ref.evolve(q1, 2)
    assert np.abs(obs[1] - o).max() < 1e-12  # a different tile edge only changes rounding
    b.close()
    ref.close()


def test_eps_sweep_batch_full_size_table_form():
    """The real-size epsilon sweep (4 x 12^4 + 4 x 16^4, N=16, RK2 + zero
    modes + FA) in the table form equals its replicas alone, bit for bit."""
    from prlgt2020_amd import nspt
    N = 16
    Ls = [(12, 12, 12, 12)] * 4 + [(16, 16, 16, 16)] * 4
    eps = [0.01, 0.015, 0.02, 0.03] * 2
    ps = [nspt.Context.params(e, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_DEVICE, seed=100 + r, step=0)
          for r, e in enumerate(eps)]
    print("table batch of 8 SU(3) replicas 4 x 12^4 + 4 x 16^4, N=16", flush=True)
    b = nspt.Batch(Ls, 3, N, ps)
    b.set_form(True)
    obs_b = b.evolve(3)
    for r, (L, p) in enumerate(zip(Ls, ps)):
        c = nspt.Context(L, 3, N)
        q = nspt.Context.params(p.eps, p.alpha, p.integrator, p.gauge_fix, p.zero_modes, p.noise, p.seed, p.step)
        o = c.evolve(q, 3)
        assert np.array_equal(b.replicas[r].download(), c.download()), r
        assert np.array_equal(obs_b[r], o), r
        c.close()
    b.close()
