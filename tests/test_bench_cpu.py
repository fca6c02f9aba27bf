"""Host logic of bench.py that needs no GPU: the kernel function behind each
profile name of the library (the roofline kernel is chosen per function),
the flop model, and the counter traffic that the roofline line reads from
the committed profile."""
import os

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_kernel_function_groups_profile_names():
    kf = bench.kernel_function
    # the two launches of the RK2 update are one kernel function
    assert kf("update", True) == kf("update_log", True) == "k_update"
    # the plaquette field and the drift from it are one drift evaluation
    assert kf("drift", True) == kf("plaq_field", True) == "k_plaq_field+k_drift_p"
    assert kf("drift", False) == "k_drift"
    assert kf("fa_div", True) == kf("dft", True) == "k_fa_fft"
    assert kf("transform", True) == "k_gauge_transform"
    assert kf("transform", False, checkerboard_gauge=True) == "k_sgf"
    assert kf("fa_gexp", True) == "k_fa_gexp"


def test_flop_model():
    # SU(3), N = 20: 43 920 flops per series product, 300 960 per exp or log
    assert bench.sp_flops(20) == 43920
    assert bench.ex_flops(20) == 300960
    V = 16 ** 4
    # an RK2 + zero modes + FA link step is 30 products and 4.25 exp/log
    assert bench.step_flops(20, V, 1) == (30 * 43920 + 4.25 * 300960) * 4 * V
    # the update with log adds one log to the update without
    assert bench.kernel_flops("update_log", 20, V) - bench.kernel_flops("update", 20, V) == 300960 * 4 * V


def test_traffic_lines_for_the_roofline_kernels():
    path = os.path.join(ROOT, "profiles", "traffic_su3_16x4_N20_rk2.md")
    found = {}
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch":
            found[f[1]] = float(f[2])
    assert "k_update" in found and "k_drift_p" in found and "k_plaq_field+k_drift_p" in found
    # HBM bytes per launch within a factor of three of the algorithmic bytes
    # (2.69 GB for k_update, 2.65-3.41 GB for k_drift_p, 4.55-5.31 GB for a
    # drift evaluation)
    assert 2.69e9 / 3 < found["k_update"] < 3 * 2.69e9
    assert 2.65e9 / 3 < found["k_drift_p"] < 3 * 3.41e9
    # the evaluation's line is the sum of its two kernels' lines
    pair = found["k_plaq_field"] + found["k_drift_p"]
    assert abs(found["k_plaq_field+k_drift_p"] - pair) <= 1e-3 * pair
    assert 4.55e9 / 3 < pair < 3 * 5.31e9


def test_executed_flop_model():
    """The executed count of a Paterson-Stockmeyer exp: X^2 (orders >= 1 x
    >= 1), Y = X^2 X (>= 2 x >= 1) and the steps R Y (>= 1 x >= 3, truncated
    at N - 3j), the pairs enumerated here one by one."""
    def brute(T, p, q):
        return sum(1 for i in range(p, T + 1) for j in range(q, T + 1) if i + j <= T)
    for N in (12, 16, 20, 24, 32):
        J = N // 3
        pr = brute(N, 1, 1) + brute(N, 2, 1) + sum(brute(N - 3 * j, 1, 3) for j in range(J))
        assert bench.ps_flops(N) == 216 * pr + 12 * 9 * N * J
    assert bench.ps_flops(20) == 216 * 739 + 12 * 9 * 20 * 6
    V = 16 ** 4
    SP, EX = bench.sp_flops(20), bench.ex_flops(20)
    # canonical: exp(X) U = EX + SP; the FA transform with the zero-mode exp folded in: EX + 2 SP
    assert bench.kernel_flops("update", 20, V) == (EX + SP) * 4 * V
    assert bench.kernel_flops("transform", 20, V) == (EX + 2 * SP) * 4 * V
    # executed: Horner from the right costs the exp alone below 12 orders
    assert bench.kernel_flops("update", 8, V, executed=True) == bench.ex_flops(8) * 4 * V
    assert bench.kernel_flops("update", 16, V, executed=True) == (bench.ps_flops(16) + bench.sp_flops(16)) * 4 * V
    assert bench.kernel_flops("update", 20, V, executed=True) == (bench.ps_flops(20) + SP) * 4 * V
    # the rectangle kernels execute 38 of the 72 products the reference forms
    assert bench.kernel_flops("rect", 20, V, executed=True) * 72 == bench.kernel_flops("rect", 20, V) * 38


def test_kernel_groups_count_a_drift_evaluation_once():
    """An RK2 step's profile: two plaquette fields and two drifts are two
    drift evaluations at 13 products per link, their time summed; the two
    updates are one function of two launches."""
    N, V = 20, 16 ** 4
    SP = bench.sp_flops(N)
    prof = {"plaq_field": (3.0, 2), "drift": (5.0, 2), "update": (2.0, 1), "update_log": (3.0, 1),
            "noise": (0.1, 1), "reduce": (0.0, 0)}
    groups, step_exec = bench.kernel_groups(prof, N, V, plaq_form=True)
    ms, cnt, fl, names, fx = groups["k_plaq_field+k_drift_p"]
    assert (ms, cnt, sorted(names)) == (8.0, 2, ["drift", "plaq_field"])
    assert fl == 2 * 13 * SP * 4 * V
    assert fx == 2 * (4.5 + 6) * SP * 4 * V
    assert groups["k_update"][:2] == [5.0, 2]
    assert "k_reduce_rows" not in groups
    assert step_exec == sum(g[4] for g in groups.values())
    # the staple form is its own kernel, at the same 13 products per link
    groups, _ = bench.kernel_groups({"drift": (4.0, 2)}, N, V, plaq_form=False)
    assert groups["k_drift"][1] == 2 and groups["k_drift"][2] == 2 * 13 * SP * 4 * V
