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
    assert kf("drift", True) == "k_drift_p"
    assert kf("drift", False) == "k_drift"
    assert kf("fa_div", True) == kf("dft", True) == "k_fa_fft"
    assert kf("transform", True) == "k_gauge_transform"
    assert kf("transform", False, checkerboard_gauge=True) == "k_sgf"
    assert kf("plaq_field", True) == "k_plaq_field"
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
    assert "k_update" in found and "k_drift_p" in found
    # HBM bytes per launch within a factor of three of the algorithmic bytes
    # (2.69 GB for k_update, 2.65-3.41 GB for k_drift_p)
    assert 2.69e9 / 3 < found["k_update"] < 3 * 2.69e9
    assert 2.65e9 / 3 < found["k_drift_p"] < 3 * 3.41e9
