"""ctypes bindings of the oracle (oracle/nspt_oracle.cpp) and of the
reference-algebra harness (oracle/_ref, built only where the reference
sources exist).

Test infrastructure: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py, never by the product package.

Series are numpy complex128 arrays of length 1 + nc*nc*N (flag first);
fields are complex128 arrays of shape (4*V, 1 + nc*nc*N) in link order
4*site + mu (QCDenvNODEpt_test.h:521).
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
EXT_PATH = os.path.join(HERE, "_build", "liboracle_ext.so")
REF_PATH = os.path.join(HERE, "_ref", "libprlgt_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(quiet=True):
    """Compile the oracle (and oracle/_ref when the reference is present)."""
    out = subprocess.run(["make", "-C", HERE], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _ptr(a):
    return a.ctypes.data_as(_dp)



_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.orc_ranlxd_new.restype = ctypes.c_void_p
        _lib.orc_ranlxd_free.argtypes = [ctypes.c_void_p]
        _lib.orc_ranlxd_draw.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        _lib.orc_noise_ranlxd.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        _lib.orc_rand55_new.restype = ctypes.c_void_p
        _lib.orc_rand55_new.argtypes = [ctypes.c_long]
        _lib.orc_rand55_free.argtypes = [ctypes.c_void_p]
        _lib.orc_rand55_ranf.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        _lib.orc_rand55_gauss.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        _lib.orc_noise_rand55.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        _lib.orc_noise_philox.argtypes = [ctypes.c_int, ctypes.c_ulonglong, ctypes.c_ulonglong,
                                          ctypes.c_int, _dp]
        _lib.orc_glibc_first_rand.argtypes = [ctypes.c_uint]
        _lib.orc_drift.argtypes = [_ip, _dp, ctypes.c_double, _dp]
        _lib.orc_gauge_wilson_jacobi.argtypes = [_ip, _dp, _dp, _dp, ctypes.c_double, _dp, _dp]
        _lib.orc_fa_gauge_fixing.argtypes = [_ip, _dp, ctypes.c_double]
        _lib.orc_sgf_default.argtypes = [_ip, _dp, ctypes.c_double]
        _lib.orc_gaction.restype = ctypes.c_double
        _lib.orc_gaction.argtypes = [_ip, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _lib.orc_gaction_grad.argtypes = [_ip, _dp, ctypes.c_int, _dp]
        _lib.orc_gstaple_sum.argtypes = [_ip, _dp, ctypes.c_int, ctypes.c_int, _dp]
        _lib.orc_set_plant.argtypes = [ctypes.c_int, ctypes.c_double]
        _lib.orc_run_stats.argtypes = [_ip, _dp, _dp, ctypes.c_int, ctypes.c_long, ctypes.c_int, _dp]
    return _lib


_ext = None


def ext_lib():
    """The oracle built in x87 long double (64-bit significand)."""
    global _ext
    if _ext is None:
        if not os.path.exists(EXT_PATH):
            build()
        _ext = ctypes.CDLL(EXT_PATH)
        _ext.orcx_step.argtypes = [_ip, _dp, _dp, _dp, _dp, _dp]
    return _ext


def ser_len(nc, N):
    return 1 + nc * nc * N


def _ser_unary(name, nc, N, a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.zeros(ser_len(nc, N), np.complex128)
    getattr(lib(), name)(nc, N, _ptr(a), _ptr(out))
    return out


def mul(nc, N, a, b):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    out = np.zeros(ser_len(nc, N), np.complex128)
    lib().orc_ser_mul(nc, N, _ptr(a), _ptr(b), _ptr(out))
    return out


def exp(nc, N, a):
    return _ser_unary("orc_ser_exp", nc, N, a)


def log(nc, N, a):
    return _ser_unary("orc_ser_log", nc, N, a)


def reh(nc, N, a):
    return _ser_unary("orc_ser_reh", nc, N, a)


def trless(nc, N, a):
    return _ser_unary("orc_ser_trless", nc, N, a)


def dag(nc, N, a):
    return _ser_unary("orc_ser_dag", nc, N, a)


def tr(nc, N, a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    out = np.zeros(N + 1, np.complex128)
    lib().orc_ser_tr(nc, N, _ptr(a), _ptr(out))
    return out


def _geom(L, nc, N):
    return (ctypes.c_int * 6)(L[0], L[1], L[2], L[3], nc, N)


def obs_len(N, improved=False):
    """plaq (N+1 complex) + norm (N real) + trU (4N complex), as doubles;
    improved actions append the rectangle plaquette w2 (N+1 complex)."""
    return 2 * (N + 1) + N + 8 * N + (2 * (N + 1) if improved else 0)


def split_obs(obs, N):
    """plaq, norm, trU (and w2 for an improved action's vector)."""
    plaq = obs[0:2 * (N + 1)].view(np.complex128)
    norm = obs[2 * (N + 1):2 * (N + 1) + N]
    o = 2 * (N + 1) + N
    tru = obs[o:o + 8 * N].view(np.complex128).reshape(4, N)
    if len(obs) > o + 8 * N:
        return plaq, norm, tru, obs[o + 8 * N:].view(np.complex128)
    return plaq, norm, tru


def step(L, nc, N, U, eps, alpha=0.0, integrator=0, update=0, gauge_fix=0, zero_modes=1, eta=None, c1=0.0):
    """One NsptEvolve step in place on U; returns the observable vector.
    c1 != 0: improved gauge action (c0 = 1 - 8 c1)."""
    V = L[0] * L[1] * L[2] * L[3]
    assert U.shape == (4 * V, ser_len(nc, N)) and U.dtype == np.complex128 and U.flags.c_contiguous
    eta = np.ascontiguousarray(eta, dtype=np.complex128)
    assert eta.shape == (4 * V, nc * nc)
    pp = (ctypes.c_double * 7)(eps, alpha, integrator, update, gauge_fix, zero_modes, c1)
    obs = np.zeros(obs_len(N, c1 != 0.0))
    lib().orc_step(_geom(L, nc, N), _ptr(U), pp, _ptr(eta), _ptr(obs))
    return obs


def step_extended(L, nc, N, Uhi, Ulo, eps, alpha=0.0, integrator=0, gauge_fix=0, zero_modes=1, eta=None):
    """One Jacobi step of the oracle in long double on the field hi + lo (two
    complex128 arrays, updated in place): an accuracy reference for FP64
    steps.  The noise and parameters are the double inputs; returns obs."""
    V = L[0] * L[1] * L[2] * L[3]
    for a in (Uhi, Ulo):
        assert a.shape == (4 * V, ser_len(nc, N)) and a.dtype == np.complex128 and a.flags.c_contiguous
    eta = np.ascontiguousarray(eta, dtype=np.complex128)
    pp = (ctypes.c_double * 7)(eps, alpha, integrator, 0, gauge_fix, zero_modes, 0.0)
    obs = np.zeros(obs_len(N))
    ext_lib().orcx_step(_geom(L, nc, N), _ptr(Uhi), _ptr(Ulo), pp, _ptr(eta), _ptr(obs))
    return obs


class planted:
    """with planted(fa=..., rk2=...): oracle calls carry a planted mistake
    (fa 1: alpha with the wrong sign, 2: the 1/V dropped, 3: p^2 = 4 sin^2(p);
    rk2: factor on the nc*eps/6 coefficient).  For the tests that show a pin
    of the oracle fails on a wrong oracle; restored on exit."""

    def __init__(self, fa=0, rk2=1.0):
        self.fa, self.rk2 = fa, rk2

    def __enter__(self):
        lib().orc_set_plant(self.fa, self.rk2)
        return self

    def __exit__(self, *exc):
        lib().orc_set_plant(0, 1.0)
        return False


def run_stats(L, nc, N, U, eps, alpha, integrator, gauge_fix, zero_modes, seed, nsteps, nbin):
    """nsteps Jacobi steps from U (in place) with ranlxd(seed) noise; returns
    the (nbin, N+1) array of per-bin means of the real plaquette per order."""
    assert nsteps % nbin == 0
    pp = (ctypes.c_double * 7)(eps, alpha, integrator, 0, gauge_fix, zero_modes, 0.0)
    bins = np.zeros((nbin, N + 1))
    lib().orc_run_stats(_geom(L, nc, N), _ptr(U), pp, seed, nsteps, nbin, _ptr(bins))
    return bins


def drift(L, nc, N, U, eps, c1=0.0):
    D = np.zeros_like(U)
    if c1 == 0.0:
        lib().orc_drift(_geom(L, nc, N), _ptr(U), eps, _ptr(D))
    else:
        lib().orc_drift_c1(_geom(L, nc, N), _ptr(U), ctypes.c_double(eps), ctypes.c_double(c1), _ptr(D))
    return D


def staple2x1(L, nc, N, U, site, mu, nu):
    out = np.zeros(ser_len(nc, N), np.complex128)
    lib().orc_staple2x1(_geom(L, nc, N), _ptr(U), site, mu, nu, _ptr(out))
    return out


def gauge_wilson_jacobi(L, nc, N, U, eta, eps):
    F = np.zeros_like(U)
    MM = np.zeros((4, ser_len(nc, N)), np.complex128)
    obs = np.zeros(obs_len(N))
    eta = np.ascontiguousarray(eta, dtype=np.complex128)
    lib().orc_gauge_wilson_jacobi(_geom(L, nc, N), _ptr(U), _ptr(F), _ptr(eta), eps, _ptr(MM), _ptr(obs))
    return F, MM, obs


def zero_modes(L, nc, N, U, MM):
    MM = np.ascontiguousarray(MM, dtype=np.complex128)
    lib().orc_zero_modes(_geom(L, nc, N), _ptr(U), _ptr(MM))


def fa_gauge_fixing(L, nc, N, U, alpha):
    lib().orc_fa_gauge_fixing(_geom(L, nc, N), _ptr(U), alpha)


def sgf_checkerboard(L, nc, N, U, alpha):
    """The sequential gauge fixing with sites in checkerboard order, in place."""
    lib().orc_sgf_checkerboard(_geom(L, nc, N), _ptr(U), ctypes.c_double(alpha))


def sgf_default(L, nc, N, U, alpha):
    lib().orc_sgf_default(_geom(L, nc, N), _ptr(U), alpha)


def reference_twist(nc):
    """The twist matrices of twistmatrices.h (not IDENTITY_TWIST), 4 x nc x nc:
    SU(3) y = diag(e^{-2 pi i/3}, 1, e^{2 pi i/3}), z the cyclic shift, t the
    shift with phases; SU(2) y = diag(-i, i), z = [[0, 1], [-1, 0]],
    t = [[0, -i], [-i, 0]].  x is the identity (the reference's SU(2) x twist
    is the identity and its SU(3) one is left undefined)."""
    w = np.exp(2j * np.pi / 3)
    if nc == 3:
        om = [np.eye(3), np.diag([np.conj(w), 1.0, w]),
              np.array([[0, 1, 0], [0, 0, 1], [1, 0, 0]]),
              np.array([[0, np.conj(w), 0], [0, 0, 1], [w, 0, 0]])]
    else:
        om = [np.eye(2), np.diag([-1j, 1j]), np.array([[0, 1], [-1, 0]]), np.array([[0, -1j], [-1j, 0]])]
    return np.array(om, np.complex128)


def set_twist(nc, mask, omegas=None):
    """Twisted boundary conditions for the oracle calls that follow (mask bit d:
    U(x + L_d e_d) = W_d U(x) W_d^dag); mask 0 restores periodic ones."""
    om = np.ascontiguousarray(omegas if omegas is not None else np.zeros((4, nc, nc)), np.complex128)
    lib().orc_set_twist(nc, mask, _ptr(om))


class twisted:
    """with twisted(nc, mask[, omegas]): ... -- oracle calls with twisted
    boundary conditions, periodic again on exit."""

    def __init__(self, nc, mask, omegas=None):
        self.args = (nc, mask, reference_twist(nc) if omegas is None else omegas)

    def __enter__(self):
        set_twist(*self.args)
        return self

    def __exit__(self, *exc):
        set_twist(self.args[0], 0)
        return False


def plaquette_measure(L, nc, N, U):
    out = np.zeros(N + 1, np.complex128)
    lib().orc_plaquette_measure(_geom(L, nc, N), _ptr(U), _ptr(out))
    return out


def knorm(L, nc, N, U):
    """Momentum norms of the __K_MOM_ANALYSIS__ build (nspt.cc:919-944):
    array (N, V), row k-1 = sum_mu Re Tr(Ut Ut^dag) of order k per momentum."""
    V = int(np.prod(L))
    out = np.zeros((N, V), np.float64)
    lib().orc_knorm(_geom(L, nc, N), _ptr(U), _ptr(out))
    return out


def compute_loops(L, nc, N, U):
    """Tr of the ComputeLoops sums w21, w22, w42, w44 (nspt.cc:1156-1312):
    array (4, N+1) complex, unnormalised."""
    out = np.zeros(4 * (N + 1), np.complex128)
    lib().orc_compute_loops(_geom(L, nc, N), _ptr(U), _ptr(out))
    return out.reshape(4, N + 1)


def staple_sum(L, nc, N, U, site, mu):
    out = np.zeros(ser_len(nc, N), np.complex128)
    lib().orc_staple_sum(_geom(L, nc, N), _ptr(U), site, mu, _ptr(out))
    return out


def dft_many(L, data, howmany, stride, sign):
    data = np.ascontiguousarray(data, dtype=np.complex128)
    Lc = (ctypes.c_int * 4)(*L)
    lib().orc_dft_many(Lc, howmany, stride, _ptr(data), sign)
    return data


class Ranlxd:
    def __init__(self, seed, level=1):
        self.h = lib().orc_ranlxd_new(seed, level)

    def draw(self, n):
        out = np.zeros(n)
        lib().orc_ranlxd_draw(self.h, n, _ptr(out))
        return out

    def noise(self, nc, nlinks):
        out = np.zeros((nlinks, nc * nc), np.complex128)
        lib().orc_noise_ranlxd(self.h, nc, nlinks, _ptr(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_ranlxd_free(self.h)
            self.h = None


class Rand55:
    def __init__(self, seed):
        self.h = lib().orc_rand55_new(seed)

    def ranf(self, n):
        out = np.zeros(n)
        lib().orc_rand55_ranf(self.h, n, _ptr(out))
        return out

    def gauss(self, n):
        out = np.zeros(n)
        lib().orc_rand55_gauss(self.h, n, _ptr(out))
        return out

    def noise(self, nc, nlinks):
        out = np.zeros((nlinks, nc * nc), np.complex128)
        lib().orc_noise_rand55(self.h, nc, nlinks, _ptr(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_rand55_free(self.h)
            self.h = None


def noise_philox(nc, seed, step, nlinks):
    out = np.zeros((nlinks, nc * nc), np.complex128)
    lib().orc_noise_philox(nc, seed, step, nlinks, _ptr(out))
    return out


def glibc_first_rand(seed):
    return lib().orc_glibc_first_rand(seed)


# ---------------------------------------------------------------- fields
def cold_field(L, nc, N):
    V = L[0] * L[1] * L[2] * L[3]
    U = np.zeros((4 * V, ser_len(nc, N)), np.complex128)
    U[:, 0] = 1.0
    return U


def random_algebra(rng, nc, N, scale=1.0, nlinks=1):
    """Random anti-Hermitian traceless series (flag 0), shape (nlinks, S)."""
    S = ser_len(nc, N)
    out = np.zeros((nlinks, S), np.complex128)
    m = rng.normal(size=(nlinks, N, nc, nc)) + 1j * rng.normal(size=(nlinks, N, nc, nc))
    m = 0.5 * (m - np.conj(np.swapaxes(m, -1, -2)))
    tr = np.trace(m, axis1=-2, axis2=-1) / nc
    m -= tr[..., None, None] * np.eye(nc)
    out[:, 1:] = (scale * m).reshape(nlinks, N * nc * nc)
    return out


def random_field(L, nc, N, seed=0, scale=0.3):
    """U = exp(A) with A a random algebra series on every link."""
    V = L[0] * L[1] * L[2] * L[3]
    rng = np.random.default_rng(seed)
    A = random_algebra(rng, nc, N, scale, 4 * V)
    U = np.empty_like(A)
    for i in range(4 * V):
        U[i] = exp(nc, N, A[i])
    return U


# ---------------------------------------------------------------- adjoint
# Generalised series: orders 0..N, all NC x NC matrices, shape ((N+1)*NC^2,);
# fields (4V, (N+1)*NC^2) in link order 4*site + mu.
def gser_len(nc, N):
    return (N + 1) * nc * nc


def gmul(nc, N, a, b):
    a = np.ascontiguousarray(a, np.complex128)
    b = np.ascontiguousarray(b, np.complex128)
    out = np.zeros(gser_len(nc, N), np.complex128)
    lib().orc_gser_mul(nc, N, _ptr(a), _ptr(b), _ptr(out))
    return out


def gaction(L, nc, N, U, K, site=-1, mu=0):
    """sum_{x, mu<nu} Re Tr P_K, or (site >= 0) the six plaquettes of a link."""
    U = np.ascontiguousarray(U, np.complex128)
    return lib().orc_gaction(_geom(L, nc, N), _ptr(U), K, site, mu)


def gaction_grad(L, nc, N, U, K):
    U = np.ascontiguousarray(U, np.complex128)
    G = np.zeros_like(U)
    lib().orc_gaction_grad(_geom(L, nc, N), _ptr(U), K, _ptr(G))
    return G


def gstaple_sum(L, nc, N, U, site, mu):
    U = np.ascontiguousarray(U, np.complex128)
    out = np.zeros(gser_len(nc, N), np.complex128)
    lib().orc_gstaple_sum(_geom(L, nc, N), _ptr(U), site, mu, _ptr(out))
    return out


def haar_sun(rng, nc, n):
    """n Haar-random SU(nc) matrices (QR of complex Gaussians, phases fixed)."""
    z = (rng.normal(size=(n, nc, nc)) + 1j * rng.normal(size=(n, nc, nc))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=1, axis2=2)
    q = q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(q)
    return q / (det ** (1.0 / nc))[:, None, None]


def random_gfield(L, nc, N, seed=7, scale=0.1):
    """Order 0 Haar-random SU(nc), orders 1..N Gaussian su(nc) of the given scale."""
    V = L[0] * L[1] * L[2] * L[3]
    rng = np.random.default_rng(seed)
    U = np.zeros((4 * V, gser_len(nc, N)), np.complex128)
    U[:, :nc * nc] = haar_sun(rng, nc, 4 * V).reshape(4 * V, nc * nc)
    U[:, nc * nc:] = random_algebra(rng, nc, N, scale, 4 * V)[:, 1:]
    return U


# ---------------------------------------------------------------- _ref
class Ref:
    """The reference's own ptSU3/ptSU2 algebra and RNGs (allocORD = 6)."""

    def __init__(self):
        if not os.path.exists(REF_PATH):
            raise FileNotFoundError(REF_PATH)
        self.l = ctypes.CDLL(REF_PATH)
        self.A = self.l.ref_alloc_ord()
        self.l.ref_rand55_ranf.argtypes = [ctypes.c_long, ctypes.c_int, _dp]
        self.l.ref_rand55_gauss.argtypes = [ctypes.c_long, ctypes.c_int, _dp]
        self.l.ref_su3rand_rand55.argtypes = [ctypes.c_long, ctypes.c_int, _dp]
        self.l.ref_su2rand_rand55.argtypes = [ctypes.c_long, ctypes.c_int, _dp]

    def _pad(self, nc, a):
        out = np.zeros(ser_len(nc, self.A), np.complex128)
        out[:len(a)] = a
        return out

    def op(self, name, nc, N, *args):
        self.l.ref_set_ptord(N)
        bufs = [self._pad(nc, a) for a in args]
        out = np.zeros(ser_len(nc, self.A), np.complex128)
        getattr(self.l, "ref_su%d_%s" % (nc, name))(*[_ptr(b) for b in bufs], _ptr(out))
        return out[:ser_len(nc, N)]

    def tr(self, nc, N, a):
        self.l.ref_set_ptord(N)
        out = np.zeros(N + 1, np.complex128)
        getattr(self.l, "ref_su%d_tr" % nc)(_ptr(self._pad(nc, a)), _ptr(out))
        return out

    def ranlxd(self, seed, n, level=1):
        out = np.zeros(n)
        self.l.ref_ranlxd(seed, level, n, _ptr(out))
        return out

    def su3rand_ranlxd(self, seed, n):
        out = np.zeros((n, 9), np.complex128)
        self.l.ref_su3rand_ranlxd(seed, n, _ptr(out))
        return out

    def rand55_ranf(self, seed, n):
        out = np.zeros(n)
        self.l.ref_rand55_ranf(seed, n, _ptr(out))
        return out

    def rand55_gauss(self, seed, n):
        out = np.zeros(n)
        self.l.ref_rand55_gauss(seed, n, _ptr(out))
        return out

    def su3rand_rand55(self, seed, n):
        out = np.zeros((n, 9), np.complex128)
        self.l.ref_su3rand_rand55(seed, n, _ptr(out))
        return out

    def su2rand_rand55(self, seed, n):
        out = np.zeros((n, 4), np.complex128)
        self.l.ref_su2rand_rand55(seed, n, _ptr(out))
        return out
