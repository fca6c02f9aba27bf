"""nspt_adjoint_drift_dev, the entry point that allocates nothing (its
scratch is a static device array), against nspt_adjoint_drift, which
allocates its own, on the same fields: two calls in a row, the second on a
smaller input, each equal bit for bit (same kernels, same order of sums);
and the size limit refused before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def to_dev_layout(A, L, nc, N):
    """Host image (4V, (N+1) nc^2), link 4*site + mu -> the engine's layout,
    element (k*nc^2 + e) * 4V + mu*V + site."""
    V = int(np.prod(L))
    S = (N + 1) * nc * nc
    return np.ascontiguousarray(A.reshape(V, 4, S).transpose(2, 1, 0)).reshape(-1)


class DeviceBuffer:
    """Device memory from the HIP runtime libnspt.so itself uses (the test is
    the caller, which may allocate; the entry point under test does not)."""
    hip = None

    def __init__(self, host):
        if DeviceBuffer.hip is None:
            DeviceBuffer.hip = ctypes.CDLL("libamdhip64.so.7")
        self.n = host.nbytes
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.n)) == 0
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.n), 1) == 0

    def get(self, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.n), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def test_no_alloc_adjoint_equals_allocating_entry_point():
    from oracle import oracle as O
    from prlgt2020_amd import nspt
    lib = nspt.lib()
    eps = 0.02
    for nc, L, N in [(3, (4, 4, 4, 6), 6), (3, (2, 4, 2, 3), 4), (2, (2, 2, 2, 4), 6)]:
        print("no-alloc adjoint SU(%d) %s N=%d" % (nc, "x".join(map(str, L)), N), flush=True)
        U = O.random_gfield(L, nc, N, seed=7, scale=0.1)
        G, D = nspt.adjoint_drift(L, nc, N, U, eps)
        u = to_dev_layout(U, L, nc, N)
        nan = np.full_like(u, np.nan)
        dU, dG, dD = DeviceBuffer(u), DeviceBuffer(nan), DeviceBuffer(nan)
        g = nspt._Geom((ctypes.c_int * 4)(*L), nc, N, N)
        rc = lib.nspt_adjoint_drift_dev(ctypes.byref(g), dU.ptr, eps, dG.ptr, dD.ptr)
        assert rc == 0
        assert np.array_equal(dG.get(u), to_dev_layout(G, L, nc, N))
        assert np.array_equal(dD.get(u), to_dev_layout(D, L, nc, N))
        # past the scratch array's size (SU(3) 16^4 at 12 orders): refused, nothing runs
        big = nspt._Geom((ctypes.c_int * 4)(16, 16, 16, 16), 3, 12, 12)
        assert lib.nspt_adjoint_drift_dev(ctypes.byref(big), dU.ptr, eps, None, None) == -4
        for b in (dU, dG, dD):
            b.free()
