"""CPU checks of the C-ABI library: it loads, exports every function that
include/nspt.h declares, its host noise source (ranlxd + su(N) mapping)
matches the oracle, and the host package fails loudly without the library.
No GPU call is made.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nspt.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nspt_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_entry_points():
    names = declared_functions()
    for must in ("nspt_create", "nspt_upload", "nspt_download", "nspt_set_noise", "nspt_step", "nspt_plaquette",
                 "nspt_destroy", "nspt_last_error"):
        assert must in names


def test_library_exports_every_declared_function():
    from prlgt2020_amd import nspt
    lib = nspt.lib()
    missing = [f for f in declared_functions() if not hasattr(lib, f)]
    assert not missing, missing
    assert sorted(nspt.EXPORTS) == sorted(declared_functions())


def test_library_exports_are_c_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "prlgt2020_amd", "libnspt.so")],
                         capture_output=True, text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    for f in declared_functions():
        assert f in syms, f  # unmangled: extern "C"


@pytest.mark.parametrize("level", [1, 2])
def test_host_ranlxd_matches_oracle(level):
    from oracle import oracle as O
    from prlgt2020_amd import nspt
    a = nspt.Ranlxd(1804289383, level).uniform(5000)
    b = O.Ranlxd(1804289383, level).draw(5000)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("nc", [3, 2])
def test_host_noise_matches_oracle(nc):
    from oracle import oracle as O
    from prlgt2020_amd import nspt
    a = nspt.Ranlxd(12345).noise(nc, 700)
    b = O.Ranlxd(12345).noise(nc, 700)
    assert np.array_equal(a, b)


def test_host_noise_matches_reference_golden():
    """The su(3) mapping of MyMath.cc:295-322 on the reference's own stream."""
    from prlgt2020_amd import nspt
    g = np.load(os.path.join(ROOT, "tests", "golden", "rng.npz"))
    want = g["su3rand_ranlxd_seed1804289383"]
    got = nspt.Ranlxd(1804289383).noise(3, want.shape[0])
    assert np.allclose(got, want.reshape(got.shape), rtol=0, atol=0)


def test_host_rand55_matches_reference_golden():
    """RAND55 of the reference's shipped build (MyRand.cc:30-130), seeded as
    Rand[0].init(1): uniforms, and the su(3) / su(2) noise of SU3rand /
    SU2rand (MyMath.cc:256-292, 326-346), bit for bit against the vectors the
    reference's own sources wrote."""
    from prlgt2020_amd import nspt
    g = np.load(os.path.join(ROOT, "tests", "golden", "rng.npz"))
    assert np.array_equal(nspt.Rand55(1).uniform(3000), g["rand55_ranf_seed1"])
    for nc, key in ((3, "su3rand_rand55_seed1"), (2, "su2rand_rand55_seed1")):
        want = g[key]
        got = nspt.Rand55(1).noise(nc, want.shape[0])
        assert np.array_equal(got, want.reshape(got.shape)), nc


def test_host_rand55_matches_oracle():
    """Longer streams and a large seed (the reference's signed seeding sum
    wraps) against the oracle's restatement."""
    from oracle import oracle as O
    from prlgt2020_amd import nspt
    for seed in (7, 2682328493):
        assert np.array_equal(nspt.Rand55(seed).noise(3, 900), O.Rand55(seed).noise(3, 900))
        assert np.array_equal(nspt.Rand55(seed).noise(2, 900), O.Rand55(seed).noise(2, 900))


def test_package_fails_loudly_without_library(tmp_path):
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from prlgt2020_amd import nspt\n"
            "try:\n    nspt.Context((2,2,2,2), 3, 2)\nexcept nspt.NsptError as e:\n    print('raised', e)\n"
            % ROOT)
    env = dict(os.environ, NSPT_LIB=str(tmp_path / "missing.so"))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert "raised" in out.stdout and "not built" in out.stdout, out.stdout + out.stderr


def test_every_kernel_build_has_its_own_launch_handles():
    """Builds of one kernel template (tile edge, occupancy, field size) live
    in namespaces of their own: each has its own host launch handle, so the
    runtime launches each build's own code (one shared handle would make every
    build run the code registered first)."""
    from prlgt2020_amd import build
    out = subprocess.run(["nm", "-C", "--defined-only", os.path.join(ROOT, "prlgt2020_amd", "libnspt.so")],
                         capture_output=True, text=True, check=True).stdout
    handles = {}
    for line in out.splitlines():
        m = re.match(r"([0-9a-f]+) V void nspt::(\w+)::k_update<(\d), (\d+)>\(", line)
        if m:
            handles[(m.group(2), int(m.group(3)), int(m.group(4)))] = m.group(1)
    want = []
    for nc, n, q, lpw, wpe in build.instances():
        for f in ("std_fields", "big_fields"):
            want.append(("%s_q%d_l%d_w%d" % (f, q, lpw, wpe), nc, n))
    missing = [w for w in want if w not in handles]
    assert not missing, missing
    addrs = [handles[w] for w in want]
    assert len(set(addrs)) == len(addrs)
