"""Throughput of the NSPT Langevin step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2|3|...]

Workload (default): SU(3) Wilson NSPT on a 16^4 periodic lattice, 20 orders
in beta^-1/2, FP64; RK2 integrator eps = 0.02, zero modes removed and
Fourier-accelerated stochastic gauge fixing (alpha = 0.05) every step; noise
generated on the device (Philox4x32-10) inside the timed step.  The field
starts cold and stays resident in HBM.

Metric: link-order updates/s = 4 V N steps / time.  One JSON line is
printed with a roofline object for the dominant kernel (FP64 flops per
launch / its average HIP-event duration over the timed steps) and the CPU
baseline (the oracle port timed on a bounded sample on this host).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X published FP64 (vector and matrix) peak
PEAK_HBM_GBS = 8000.0


def sp_flops(N, nc=3):
    """Canonical flops of one truncated series product: 216 per 3x3 complex
    matrix product of two orders (64 for 2x2), plus the flag terms."""
    if nc == 3:
        return 216 * N * (N - 1) // 2 + 144 * N
    return 64 * N * (N - 1) // 2 + 64 * N


def ex_flops(N, nc=3):
    """Canonical flops of one series exp or log."""
    if nc == 3:
        return 216 * (N - 1) * N * (N + 1) // 6 + 36 * N * (N - 1)
    return 64 * (N - 1) * N * (N + 1) // 6 + 16 * N * (N - 1)


ACTIONS = {"wil": 0.0, "tls": -1.0 / 12.0, "iwa": -0.331, "dbw": -1.4088}  # c1, Quenched.cc:313-322

# instances of at least this many orders evaluate exp and log in the
# Paterson-Stockmeyer form (NSPT_PS_MIN_N of the library's build)
PS_MIN_N = 12


def pairs(T, p, q):
    """Pairs of orders (i >= p, j >= q, i + j <= T) of a truncated product."""
    m = T - p - q + 1
    return m * (m + 1) // 2 if m > 0 else 0


def ps_flops(N, nc=3):
    """Executed flops of one exp or log in the Paterson-Stockmeyer form: X^2,
    Y = X^3 and the N/3 steps in Y, each product over the pairs of its
    operands' nonzero orders (a matrix product of two orders is 216 flops,
    64 for SU(2)), plus each step's lane-local addend (12 flops per element
    and order)."""
    J = N // 3
    pr = pairs(N, 1, 1) + pairs(N, 1, 2) + sum(pairs(N - 3 * j, 3, 1) for j in range(J))
    return (216 if nc == 3 else 64) * pr + 12 * nc * nc * N * J


def kernel_flops(name, N, V, nc=3, improved=False, cb_update=False, cb_gauge=False, fused=True, executed=False,
                 plaq_form=True):
    """FP64 flops of one launch of the kernel behind a profile name.

    Canonical (executed=False): the reference's algorithm, a product SP, an
    exp or log EX (Horner), exp(X) B = EX + SP.  Executed: what this
    build's kernel forms, where exp(X) B by Horner from the right costs EX
    alone and from 12 orders on exp and log are ps_flops.
    The Jacobi drift is the plaquette field (3 products per plane and site,
    4.5 per link) plus 6 products per link (plaq_form), or the staple sum, 13
    products per link, as up to 16 orders and in the checkerboard sweep
    (cb_update: one launch per parity and direction, V/2 links); the checkerboard gauge fixing
    (cb_gauge) is one launch per parity (an exp and 8 products per site).
    With zero modes folded into the FA transform (fused), that kernel also
    forms the zero-mode exponential.  An improved
    action adds U R to the drift; the rectangle kernels count 72 products
    per link as the reference forms them (six 5-link staples in each of three
    planes, QCDenvNODEpt_test.h:591-607) and execute 38."""
    SP, EX = sp_flops(N, nc), ex_flops(N, nc)
    if executed:
        ps = N >= PS_MIN_N
        E = ps_flops(N, nc) if ps else EX
        ET = E + SP if ps else EX
        rect = 38 * SP
    else:
        E, ET, rect = EX, EX + SP, 72 * SP
    drift = (13 if cb_update or not plaq_form else 6) + (1 if improved else 0)
    if cb_update and name in ("drift", "update", "update_log"):
        return {"drift": drift * SP, "update": ET, "update_log": ET + E}[name] * V / 2
    if cb_gauge and name == "transform":
        return (E + 8 * SP) * V / 2
    per_link = {"drift": drift * SP, "plaq_field": 4.5 * SP, "update": ET, "update_log": ET + E,
                "zero_mode": E, "transform": (ET + SP) if fused else 2 * SP, "rect": rect}
    if name in per_link:
        return per_link[name] * 4 * V
    if name == "fa_gexp":
        return E * V
    return 0


# one drift evaluation of the plaquette form: the plaquette field, then the
# drift from it (one launch of each)
DRIFT_PAIR = "k_plaq_field+k_drift_p"


def kernel_function(pname, plaquette_drift, checkerboard_gauge=False):
    """The kernel function(s) behind one of the library's profile names; the
    two kernels of a plaquette-form drift evaluation count as one, whose
    algorithmic flops are the drift's 13 products per link."""
    if pname in ("drift", "plaq_field") and plaquette_drift:
        return DRIFT_PAIR
    if pname == "drift":
        return "k_drift"
    if pname == "transform":
        return "k_sgf" if checkerboard_gauge else "k_gauge_transform"
    return {"update_log": "k_update", "fa_div": "k_fa_fft", "dft": "k_fa_fft", "rect": "k_link_pair+k_rect",
            "noise": "k_noise_philox", "reduce": "k_reduce_rows"}.get(pname, "k_" + pname)


def kernel_groups(prof, N, V, nc=3, improved=False, **how):
    """Sum the library's profile (name -> (ms, launches)) per kernel function:
    {function: [ms, launches, canonical flops, profile names, executed
    flops]}, and the executed flops of all launches.  A launch of the
    plaquette-form pair (k_plaq_field+k_drift_p) is one drift evaluation,
    counted on the drift kernel and worth the drift's algorithmic 13 products
    per link (14 with an improved action)."""
    plaq_form = how.get("plaq_form", True)
    groups = {}
    step_exec = 0.0
    for pname, (ms, c) in prof.items():
        if not c:
            continue
        fname = kernel_function(pname, plaq_form, how.get("cb_gauge", False))
        fl1 = kernel_flops(pname, N, V, nc, improved, **how)
        fx1 = kernel_flops(pname, N, V, nc, improved, executed=True, **how)
        step_exec += fx1 * c
        c_fn = c
        if fname == DRIFT_PAIR:
            c_fn = c if pname == "drift" else 0
            fl1 = kernel_flops("drift", N, V, nc, improved, plaq_form=False) if pname == "drift" else 0.0
        g = groups.setdefault(fname, [0.0, 0, 0.0, [], 0.0])
        g[0] += ms
        g[1] += c_fn
        g[2] += fl1 * c
        g[3].append(pname)
        g[4] += fx1 * c
    return groups, step_exec


def step_flops(N, V, integrator, nc=3, improved=False):
    SP, EX = sp_flops(N, nc), ex_flops(N, nc)
    per_link = (30 * SP + 4.25 * EX) if integrator == 1 else (16 * SP + 3.25 * EX)
    if improved:
        per_link += (2 if integrator == 1 else 1) * 73 * SP
    return per_link * 4 * V


def host_cpu():
    """CPU model and the cores this process may use."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"model": model, "nproc": avail, "cpu_count": os.cpu_count()}


def calibration():
    """Oracle / reference throughput ratio measured where the reference's
    sources are (oracle/calibrate.py), if recorded."""
    path = os.path.join(ROOT, "oracle", "calibration.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        c = json.load(f)
    return {"ratio_oracle_over_reference": c["ratio"], "measured_on": c["cpu"],
            "how": "the SU(3) series algebra of one RK2 + zero-mode + FA link-step (30 products, 3 exp, 1 log) "
                   "at %d orders, timed on one core for the oracle and for the reference's own QCDpt.cc / "
                   "MyMath.cc compiled from its sources (oracle/_ref); the reference's step itself needs FFTW3, "
                   "absent from the image; rows in oracle/calibration.json" % c.get("orders", 6)}


def cpu_baseline(N, integrator, budget_s=12.0):
    """Oracle port, one core, 4^4 lattice with the workload's orders and step."""
    from oracle import oracle as O
    L, nc = (4, 4, 4, 4), 3
    V = 256
    U = O.cold_field(L, nc, N)
    done, t0 = 0, time.time()
    while True:
        eta = O.noise_philox(nc, 1, done, 4 * V)
        O.step(L, nc, N, U, 0.02, 0.05, integrator=integrator, update=0, gauge_fix=1, zero_modes=1, eta=eta)
        done += 1
        if time.time() - t0 > budget_s or done >= 50:
            break
    dt = time.time() - t0
    return {"value": 4 * V * N * done / dt, "unit": "link-order updates/s", "cores": 1, "kind": "port",
            "sample": "oracle/nspt_oracle.cpp, SU(3) 4^4, N=%d, %s + zero modes + FA, %d steps in %.1f s, 1 core"
                      % (N, "RK2" if integrator == 1 else "Euler", done, dt)}


_FARM_CHILD = """
import sys, time
sys.path.insert(0, %r)
from oracle import oracle as O
N, integrator, budget, seed = %d, %d, %f, int(sys.argv[1])
L, nc, V = (4, 4, 4, 4), 3, 256
U = O.cold_field(L, nc, N)
done, t0 = 0, time.time()
while time.time() - t0 < budget:
    eta = O.noise_philox(nc, seed, done, 4 * V)
    O.step(L, nc, N, U, 0.02, 0.05, integrator=integrator, update=0, gauge_fix=1, zero_modes=1, eta=eta)
    done += 1
print(done, time.time() - t0)
"""


def cpu_farm(N, integrator, budget_s=8.0, procs=None):
    """The reference's farm model (one independent replica per core,
    Quenched.cc:262): the oracle step of cpu_baseline in `procs` child
    processes at once for about budget_s seconds; throughput summed.  Run
    before the GPU is touched (the children are separate programs)."""
    import subprocess
    procs = procs or min(16, os.cpu_count() or 1)
    code = _FARM_CHILD % (ROOT, N, integrator, budget_s)
    ps = [subprocess.Popen([sys.executable, "-c", code, str(r + 1)], stdout=subprocess.PIPE, text=True)
          for r in range(procs)]
    V, total, steps = 256, 0.0, 0
    for p in ps:
        out, _ = p.communicate()
        if p.returncode != 0:
            return None
        n, dt = out.split()
        total += 4 * V * N * int(n) / float(dt)
        steps += int(n)
    return {"value": total, "unit": "link-order updates/s", "cores": procs, "kind": "port",
            "sample": "%d processes x oracle SU(3) 4^4, N=%d, %s + zero modes + FA, %.0f s each, %d steps in all"
                      % (procs, N, "RK2" if integrator == 1 else "Euler", budget_s, steps)}


def host_noise_path(nspt, L, nc, N, eps, p_dev, steps=3):
    """The reference's noise path beside `value`: the su(N) noise of every
    link drawn on the host by ranlxd (the reference's stream, nspt_rng_noise)
    and uploaded from pageable memory before each step, then the same step.
    Wall time per phase; the upload completes inside the step call, which
    ends with a stream synchronisation."""
    V = int(np.prod(L))
    ctx = nspt.Context(L, nc, N)
    rng = nspt.Ranlxd(1)
    p = nspt.Context.params(eps, p_dev.alpha, p_dev.integrator, p_dev.gauge_fix, p_dev.zero_modes,
                            nspt.NOISE_HOST)
    t_gen = t_set = t_step = 0.0
    for s in range(steps + 1):  # the first step is a warmup
        t0 = time.perf_counter()
        eta = rng.noise(nc, 4 * V)
        t1 = time.perf_counter()
        ctx.set_noise(eta)
        t2 = time.perf_counter()
        ctx.step(p, want_obs=False)
        t3 = time.perf_counter()
        if s:
            t_gen, t_set, t_step = t_gen + t1 - t0, t_set + t2 - t1, t_step + t3 - t2
    ctx.close()
    tot = t_gen + t_set + t_step
    return {"value": 4 * V * N * steps / tot, "unit": "link-order updates/s", "ms_per_step": tot / steps * 1e3,
            "ms_ranlxd_host": t_gen / steps * 1e3, "ms_set_noise": t_set / steps * 1e3,
            "ms_step_with_upload": t_step / steps * 1e3, "noise_bytes_per_step": 4 * V * nc * nc * 16,
            "steps": steps, "host_threads": 1}


def adjoint_bench(L=(8, 8, 8, 8), nc=3, N=12, calls=20):
    """The adjoint drift at its configuration (SU(3) 8^4, 12 orders): the
    entry point on device buffers (nspt_adjoint_drift_dev, which returns when
    its kernels are done), timed over `calls` calls after two untimed ones.
    It forms 13 full products of N+1 orders per link (8 per plaquette, 6
    plaquettes per site; one per link for the drift), (N+1)(N+2)/2 pairs of
    orders each."""
    import ctypes
    import torch
    from prlgt2020_amd import nspt
    V = int(np.prod(L))
    ng, ne = N + 1, nc * nc
    n = 4 * V * ng * ne  # complex
    gen = torch.Generator(device="cuda").manual_seed(7)
    U = 0.1 * torch.randn(2 * n, dtype=torch.float64, device="cuda", generator=gen)
    G = torch.empty_like(U)
    D = torch.empty_like(U)
    geom = nspt._Geom((ctypes.c_int * 4)(*L), nc, N, N)
    f = nspt.lib().nspt_adjoint_drift_dev
    args = (ctypes.byref(geom), ctypes.c_void_p(U.data_ptr()), 0.02, ctypes.c_void_p(G.data_ptr()),
            ctypes.c_void_p(D.data_ptr()))
    torch.cuda.synchronize()
    for _ in range(2):
        if f(*args) != 0:
            raise RuntimeError("nspt_adjoint_drift_dev failed")
    t0 = time.perf_counter()
    for _ in range(calls):
        f(*args)
    dt = (time.perf_counter() - t0) / calls
    flops = 13 * 4 * V * (216 if nc == 3 else 64) * ng * (ng + 1) // 2
    return {"workload": "adjoint drift, SU(%d) %s, %d orders (order 0 a full matrix)"
                        % (nc, "x".join(map(str, L)), N),
            "ms_per_call": dt * 1e3, "link_orders_per_s": 4 * V * ng / dt, "flops_per_call": flops,
            "tflops": flops / dt / 1e12, "frac": flops / dt / 1e12 / PEAK_FP64_TFLOPS, "calls": calls,
            "note": "a call is two kernels (the plaquette adjoints, then the per-link sum and drift) on the null "
                    "stream plus the host synchronisation; at 8^4 they are 3 516 and 2 344 one-wave workgroups, "
                    "3.4 and 2.3 rounds of the chip's 1 024 wave slots"}


def batched_sweep(nsteps, warmup, sizes=(12, 16), per_volume=4):
    """Epsilon-replica sweep: SU(3) per_volume x a^4 + per_volume x b^4 for
    sizes (a, b) (the workload is 4 x 12^4 + 4 x 16^4), N = 16, RK2 + zero
    modes + FA, eps in {0.01, 0.015, 0.02, 0.03} per volume (repeated with
    further seeds when per_volume > 4), distinct seeds.  Plain: each replica
    a context of its own, stepped alone (nspt_evolve, its own launches).
    Batched: one graph-replayed step of all replicas, the step form
    (branches, or the replica table in 1, 2 or 4 concurrent groups), tile
    edge and occupancy (links per wave, waves per SIMD) autotuned; the other
    forms are timed too."""
    from prlgt2020_amd import nspt
    N = 16
    Ls = [(sizes[0],) * 4] * per_volume + [(sizes[1],) * 4] * per_volume
    eps = [0.01, 0.015, 0.02, 0.03] * (per_volume // 2)
    ps = [nspt.Context.params(e, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_DEVICE, seed=100 + r, step=0)
          for r, e in enumerate(eps)]
    work = sum(4 * L[0] ** 4 * N for L in Ls)  # link-orders per step of all replicas
    b = nspt.Batch(Ls, 3, N, ps)
    b.evolve(warmup, want_obs=False)
    # stage 1: step form (branches / table in 1, 2, 4 groups) x tile edge;
    # stage 2: the occupancy variants of the chosen edge in the chosen form
    qt, tune_ms = b.tune(steps=3)
    cands = b.tune_candidates()
    shape = b.kernel_shape()
    chosen = b.form
    # plain: the same replicas as contexts of their own, one after the other,
    # each with its own launches
    plain = [nspt.Context(L, 3, N) for L in Ls]
    pp = [nspt.Context.params(e, 0.05, nspt.RK2, nspt.GF_FA, 1, nspt.NOISE_DEVICE, seed=100 + r, step=0)
          for r, e in enumerate(eps)]
    for c, p in zip(plain, pp):
        c.evolve(p, warmup, want_obs=False)
    t_rep = []
    for c, p in zip(plain, pp):
        t0 = time.perf_counter()
        c.evolve(p, nsteps, want_obs=True)
        t_rep.append(time.perf_counter() - t0)
    t_plain = sum(t_rep)
    for c in plain:
        c.close()
    # each volume's plain rate: where every replica fills the GPU by itself,
    # a schedule of them can only remove launch gaps and tail rounds, so the
    # batch stays near the faster volume's own rate
    per_volume_value = {}
    for size in sorted(set(sizes)):
        rs = [r for r, L in enumerate(Ls) if L[0] == size]
        per_volume_value[str(size)] = sum(4 * size ** 4 * N for _ in rs) * nsteps / sum(t_rep[r] for r in rs)
    t0 = time.perf_counter()
    b.evolve(nsteps, want_obs=True)
    t_batch = time.perf_counter() - t0
    # every other step form, timed the same way (the kernel build stays the tuned one)
    forms = [0, 1] + [g for g in (2, 4) if g <= len(Ls) // 2]
    name = lambda f: "branches" if f == 0 else ("table" if f == 1 else "table in %d groups" % f)
    form_values = {name(chosen): work * nsteps / t_batch}
    for f in forms:
        if f == chosen:
            continue
        b.set_form(f)
        b.evolve(1, want_obs=True)
        t0 = time.perf_counter()
        b.evolve(nsteps, want_obs=True)
        form_values[name(f)] = work * nsteps / (time.perf_counter() - t0)
    b.close()
    # the tuner's candidates: form x tile edge, then the occupancy variants
    label = lambda c: "%s, edge %d%s" % (name(c[0]), c[1], "" if c[2] == 0 and c[3] == 1 else
                                        ", %d links/wave, %d waves/SIMD" % (c[2], c[3]))
    tune = {label(c): round(float(tune_ms[i]), 3) for i, c in enumerate(cands)}
    return {"workload": "SU(3) %dx%d^4 + %dx%d^4, 16 orders, RK2 + zero modes + FA, eps {0.01,0.015,0.02,0.03} "
                        "per volume%s, distinct seeds, device noise"
                        % (per_volume, sizes[0], per_volume, sizes[1],
                           " (x%d seeds)" % (per_volume // 4) if per_volume > 4 else ""),
            "steps": nsteps, "plain_value": work * nsteps / t_plain, "batched_value": work * nsteps / t_batch,
            "plain_value_per_volume": per_volume_value, "best_volume_plain_value": max(per_volume_value.values()),
            "unit": "link-order updates/s", "speedup": t_plain / t_batch,
            "form": name(chosen), "tile_edge": qt, "links_per_wave": shape[1], "waves_per_simd": shape[2],
            "form_values": form_values, "tune_ms_per_step": tune}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--L", type=int, default=16)
    ap.add_argument("--N", type=int, default=20)
    ap.add_argument("--nc", type=int, choices=[3, 2], default=3)
    ap.add_argument("--eps", type=float, default=0.02)
    ap.add_argument("--integrator", choices=["rk2", "euler", "checkerboard"], default="rk2",
                    help="checkerboard: the on-the-fly Euler sweep in checkerboard order")
    ap.add_argument("--gauge", choices=["fa", "checkerboard"], default="fa",
                    help="gauge fixing: Fourier accelerated, or the sequential one in checkerboard order")
    ap.add_argument("--action", choices=list(ACTIONS), default="wil",
                    help="gauge action: Wilson (the default workload) or an improved one")
    ap.add_argument("--twist", type=int, default=0,
                    help="twisted boundary conditions: bit d twists direction d with the reference's twist "
                         "matrices (6: y and z, its default); needs --gauge checkerboard")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched epsilon-replica sweep")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-drift line")
    ap.add_argument("--no-host-noise", action="store_true",
                    help="skip the host-noise path (ranlxd on the host, upload before each step)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="before the step, copy a buffer of this many MiB device to device (a plain copy kernel "
                         "of known size that scales the FETCH_SIZE / WRITE_SIZE counters of a --pmc run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: measures one GPU; --gpus %d is not supported (independent replicas would run one "
              "or more per GPU; one lattice would need halo exchanges not implemented here)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    integ_cpu = {"rk2": 1, "euler": 0, "checkerboard": 0}[args.integrator]
    farm = None if args.no_cpu else cpu_farm(args.N, integ_cpu)

    if args.copy_calib_mib or not args.no_adjoint:
        # torch's device memory (the adjoint line's buffers, the calibration
        # copy) needs torch to initialise the GPU before the library does
        import torch
        torch.empty(1, device="cuda")
    from prlgt2020_amd import nspt
    if args.copy_calib_mib:
        import torch
        n = args.copy_calib_mib << 20
        src = torch.ones(n // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    # the roofline's peaks are AMD's published figures; what this GPU reaches
    # (an FP64 FMA-chain kernel, a 1 GiB device copy) is reported beside them
    meas_tflops, meas_gbs = nspt.measure_peaks()
    peak_fp64, peak_hbm = PEAK_FP64_TFLOPS, PEAK_HBM_GBS
    L = (args.L,) * 4
    N, nc = args.N, args.nc
    V = args.L ** 4
    integ = {"rk2": nspt.RK2, "euler": nspt.EULER, "checkerboard": nspt.EULER_CHECKERBOARD}[args.integrator]
    gfix = nspt.GF_FA if args.gauge == "fa" else nspt.GF_CHECKERBOARD
    alpha = 0.05 if args.gauge == "fa" else -0.05  # FA needs alpha > 0; -0.05 is the shipped value
    ctx = nspt.Context(L, nc, N)
    if args.twist:
        ctx.set_twist(args.twist)
    improved = args.action != "wil"
    if improved:
        ctx.set_action(ACTIONS[args.action])
    p = nspt.Context.params(args.eps, alpha, integ, gfix, 1, nspt.NOISE_DEVICE, seed=1, step=0)
    ctx.evolve(p, args.warmup, want_obs=False)
    import ctypes
    lib = nspt.lib()
    lib.nspt_profile(ctx.h, 1)
    t0 = time.perf_counter()
    # the per-step observables (plaquette, norm, Tr U per order) are formed
    # and copied back, as the reference writes them every step
    obs = ctx.evolve(p, args.steps, want_obs=True)  # returns after the stream is synchronised
    dt = time.perf_counter() - t0
    prof = ctx.profile_read()
    lib.nspt_profile(ctx.h, 0)
    plaq = ctx.plaquette()
    value = 4 * V * N * args.steps / dt

    # dominant kernel function by total time: the library's profile names
    # split some functions by what a launch does (k_update with and without
    # the log), so launches, time and flops are summed per function first
    uses_pf = bool(prof.get("plaq_field", (0, 0))[1])
    how = dict(cb_update=integ == nspt.EULER_CHECKERBOARD, cb_gauge=args.gauge != "fa", fused=args.gauge == "fa",
               plaq_form=uses_pf)
    groups, step_exec = kernel_groups(prof, N, V, nc, improved, **how)
    name = max(groups, key=lambda k: groups[k][0])
    tot_ms, cnt, fl_tot, pnames, fx_tot = groups[name]
    avg_s = tot_ms / cnt / 1e3
    fl = fl_tot / cnt
    achieved = fl / avg_s / 1e12 if fl else 0.0
    fx = fx_tot / cnt
    executed = {"flops_per_launch": fx, "achieved": fx / avg_s / 1e12, "frac": fx / avg_s / 1e12 / peak_fp64}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_su3_16x4_N20_rk2.md")
    if os.path.exists(tpath) and (args.L, N, args.integrator, args.action, args.twist) == (16, 20, "rk2", "wil", 0):
        # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE counter runs
        # recorded there (calibrated against a copy of known size)
        for line in open(tpath):
            f = line.split()
            if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == name:
                traffic = float(f[2])
    step_tf = step_flops(N, V, integ, nc, improved) * args.steps / dt / 1e12
    step_xtf = step_exec / dt / 1e12

    out = {
        "metric": "link-order updates/s per Langevin step (FP64), plain single replica",
        "value": value,
        "unit": "link-order updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: cold start, device Philox noise",
        "config": {"workload": "SU(%d) %s NSPT %d^4, %d orders, %s eps=%g, zero modes + FA gauge fixing "
                               "(alpha=0.05) every step"
                               % (nc, {"wil": "Wilson", "tls": "tree-level Symanzik", "iwa": "Iwasaki",
                                       "dbw": "DBW2"}[args.action], args.L, N, args.integrator.upper(), args.eps)
                               if args.gauge == "fa" else
                               "SU(%d) Wilson NSPT %d^4, %d orders, %s eps=%g, zero modes + sequential gauge fixing "
                               "in checkerboard order (alpha=-0.05) every step%s"
                               % (nc, args.L, N, args.integrator.upper(), args.eps,
                                  ", twisted boundary conditions in %s" % "".join(
                                      "xyzt"[d] for d in range(4) if (args.twist >> d) & 1) if args.twist else ""),
                   "lattice": list(L), "orders": N, "integrator": args.integrator, "replicas": 1},
        "roofline": {"bound": "fp64-valu", "kernel": name, "profile_names": pnames,
                     "achieved": achieved, "peak": peak_fp64, "unit": "TFLOP/s",
                     "frac": achieved / peak_fp64, "traffic": traffic,
                     "avg_launch_ms": avg_s * 1e3, "flops_per_launch": fl,
                     "executed": executed,
                     "traffic_gbs": traffic / avg_s / 1e9 if traffic else None,
                     "traffic_frac_of_hbm_peak": traffic / avg_s / 1e9 / peak_hbm if traffic else None,
                     "step_fp64_tflops": step_tf, "step_frac": step_tf / peak_fp64,
                     "step_executed_tflops": step_xtf, "step_executed_frac": step_xtf / peak_fp64,
                     "flop_model": "achieved, frac, step_fp64_tflops and step_frac price the algorithmic flops of "
                                   "the reference's algorithm (a product SP, an exp or log EX by Horner, 13 SP per "
                                   "drift evaluation, exp(X)B = EX + SP: the canonical count per link-step); "
                                   "executed and step_executed_* price the flops these kernels execute "
                                   "(the plaquette-field drift, exp(X)B by Horner from the right, from 12 orders on "
                                   "exp and log by Paterson-Stockmeyer), fewer for the same result; the plaquette "
                                   "field and the drift from it (k_plaq_field+k_drift_p) are one drift evaluation, "
                                   "their launch time summed"},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in prof.items() if v[1]},
        "plaquette_order2": float(plaq[2].real),
        "plaquette_order2_last_step": float(obs[-1][4]),
        "peaks": {"fp64_tflops_spec": PEAK_FP64_TFLOPS, "fp64_tflops_measured": meas_tflops,
                  "hbm_gbs_spec": PEAK_HBM_GBS, "hbm_gbs_measured": meas_gbs,
                  "note": "the roofline's peak is AMD's published MI355X figure (FP64 vector and matrix alike; the "
                          "path runs on the FP64 VALU, no MFMA); measured: an FP64 FMA-chain kernel at full "
                          "occupancy and a 1 GiB device-to-device copy (bytes read + written), best of three"},
    }
    ctx.close()
    if not args.no_host_noise and not args.twist and not improved:
        out["host_noise"] = host_noise_path(nspt, L, nc, N, args.eps, p)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_bench()
    if not args.no_batched:
        out["batched"] = batched_sweep(max(4, args.steps), 2)
        # replicas too small to fill the GPU alone, where the graph and the
        # concurrent branches pay
        out["batched_small"] = batched_sweep(max(20, 4 * args.steps), 4, sizes=(4, 6))
        # the same sweep with four seeds per (volume, eps): 32 replicas
        out["batched_many"] = batched_sweep(max(20, 4 * args.steps), 4, sizes=(4, 6), per_volume=16)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(N, 1 if integ == nspt.RK2 else 0)
        out["cpu_baseline"]["farm"] = farm
        out["cpu_baseline"]["host"] = host_cpu()
        cal = calibration()
        if cal:
            cal["reference_equivalent_value"] = out["cpu_baseline"]["value"] / cal["ratio_oracle_over_reference"]
            out["cpu_baseline"]["calibration"] = cal
    print(json.dumps(out))


if __name__ == "__main__":
    main()
