"""ctypes binding of libnspt.so (include/nspt.h).

The product path: every call goes to the HIP library; there is no CPU
fallback.  Importing this module without the built library raises.

Host field images are numpy complex128 arrays of shape (4*V, 1 + nc*nc*A) in
link order 4*site + mu, A = alloc_ord (the ptGluon[V] layout of
QCDpt.h:47-48 / QCDenvNODEpt_test.h:338).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("NSPT_LIB") or os.path.join(HERE, "libnspt.so")

EULER, RK2, EULER_CHECKERBOARD = 0, 1, 2
GF_NONE, GF_FA, GF_CHECKERBOARD = 0, 1, 2
NOISE_HOST, NOISE_DEVICE, NOISE_ZERO = 0, 1, 2


class NsptError(RuntimeError):
    pass


def reference_twist(nc):
    """Twist matrices W_x, W_y, W_z, W_t of twistmatrices.h (not
    IDENTITY_TWIST), shape (4, nc, nc).  SU(3): W_y = diag(w*, 1, w) with
    w = exp(2 pi i/3), W_z the cyclic shift (rows (0 1 0), (0 0 1), (1 0 0)),
    W_t the shift with phases (rows (0 w* 0), (0 0 1), (w 0 0)); SU(2):
    W_y = diag(-i, i), W_z = ((0, 1), (-1, 0)), W_t = ((0, -i), (-i, 0)).
    W_x is the identity (the reference's SU(2) x twist is the identity, its
    SU(3) one is left undefined); the reference twists y and z by default
    (choices.h:45-50)."""
    w = np.exp(2j * np.pi / 3)
    if nc == 3:
        m = [np.eye(3), np.diag([w.conjugate(), 1.0, w]), [[0, 1, 0], [0, 0, 1], [1, 0, 0]],
             [[0, w.conjugate(), 0], [0, 0, 1], [w, 0, 0]]]
    elif nc == 2:
        m = [np.eye(2), np.diag([-1j, 1j]), [[0, 1], [-1, 0]], [[0, -1j], [-1j, 0]]]
    else:
        raise ValueError("nc must be 2 or 3")
    return np.array(m, np.complex128)


class _Geom(ctypes.Structure):
    _fields_ = [("sz", ctypes.c_int * 4), ("nc", ctypes.c_int), ("ptord", ctypes.c_int),
                ("alloc_ord", ctypes.c_int)]


class StepParams(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_double), ("alpha", ctypes.c_double), ("integrator", ctypes.c_int),
                ("gauge_fix", ctypes.c_int), ("zero_modes", ctypes.c_int), ("noise", ctypes.c_int),
                ("seed", ctypes.c_ulonglong), ("step", ctypes.c_ulonglong)]


_dp = ctypes.POINTER(ctypes.c_double)
_lib = None

EXPORTS = ["nspt_create", "nspt_destroy", "nspt_last_error", "nspt_obs_len", "nspt_upload", "nspt_download",
           "nspt_cold_start", "nspt_set_noise", "nspt_step", "nspt_evolve", "nspt_plaquette",
           "nspt_fa_gauge_fixing", "nspt_momentum_norm", "nspt_drift", "nspt_rng_create_ranlxd", "nspt_rng_destroy",
           "nspt_rng_uniform", "nspt_rng_noise", "nspt_device_noise", "nspt_profile", "nspt_profile_read",
           "nspt_batch_create", "nspt_batch_destroy", "nspt_batch_last_error", "nspt_batch_evolve",
           "nspt_batch_replica", "nspt_batch_tune", "nspt_adjoint_drift", "nspt_loops",
           "nspt_unitarize", "nspt_set_action", "nspt_set_twist", "nspt_gauge_wilson",
           "nspt_zero_modes_subtraction", "nspt_stochastic_gauge_fixing", "nspt_rng_create_rand55",
           "nspt_adjoint_drift_dev", "nspt_batch_set_form", "nspt_batch_form", "nspt_measure_peaks",
           "nspt_kernel_shape", "nspt_batch_tune_candidates", "nspt_batch_kernel_shape"]

PROF_NAMES = ["drift", "update", "update_log", "zero_mode", "fa_div", "dft", "fa_gexp", "transform", "noise",
              "reduce", "rect", "plaq_field"]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NsptError("libnspt.so is not built: run python -m prlgt2020_amd.build")
        l = ctypes.CDLL(LIB_PATH)
        l.nspt_create.argtypes = [ctypes.POINTER(_Geom), ctypes.POINTER(ctypes.c_void_p)]
        l.nspt_destroy.argtypes = [ctypes.c_void_p]
        l.nspt_last_error.argtypes = [ctypes.c_void_p]
        l.nspt_last_error.restype = ctypes.c_char_p
        l.nspt_obs_len.argtypes = [ctypes.c_void_p]
        for f in ("nspt_upload", "nspt_download", "nspt_set_noise", "nspt_plaquette"):
            getattr(l, f).argtypes = [ctypes.c_void_p, _dp]
        l.nspt_cold_start.argtypes = [ctypes.c_void_p]
        l.nspt_step.argtypes = [ctypes.c_void_p, ctypes.POINTER(StepParams), _dp]
        l.nspt_evolve.argtypes = [ctypes.c_void_p, ctypes.POINTER(StepParams), ctypes.c_int, _dp]
        l.nspt_fa_gauge_fixing.argtypes = [ctypes.c_void_p, ctypes.c_double]
        l.nspt_momentum_norm.argtypes = [ctypes.c_void_p, _dp]
        l.nspt_drift.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        l.nspt_rng_create_ranlxd.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        l.nspt_rng_create_rand55.argtypes = [ctypes.c_long, ctypes.POINTER(ctypes.c_void_p)]
        l.nspt_gauge_wilson.argtypes = [ctypes.c_void_p, ctypes.POINTER(StepParams), _dp]
        l.nspt_zero_modes_subtraction.argtypes = [ctypes.c_void_p]
        l.nspt_stochastic_gauge_fixing.argtypes = [ctypes.c_void_p, ctypes.c_double]
        l.nspt_rng_destroy.argtypes = [ctypes.c_void_p]
        l.nspt_rng_uniform.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        l.nspt_rng_noise.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        l.nspt_device_noise.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_ulonglong, _dp]
        l.nspt_profile.argtypes = [ctypes.c_void_p, ctypes.c_int]
        l.nspt_profile_read.argtypes = [ctypes.c_void_p, _dp, ctypes.POINTER(ctypes.c_longlong)]
        l.nspt_batch_create.argtypes = [ctypes.POINTER(_Geom), ctypes.POINTER(StepParams), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]
        l.nspt_batch_destroy.argtypes = [ctypes.c_void_p]
        l.nspt_batch_last_error.argtypes = [ctypes.c_void_p]
        l.nspt_batch_last_error.restype = ctypes.c_char_p
        l.nspt_batch_evolve.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        l.nspt_batch_replica.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        l.nspt_batch_tune.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp]
        l.nspt_batch_set_form.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _ip = ctypes.POINTER(ctypes.c_int)
        l.nspt_kernel_shape.argtypes = [ctypes.c_void_p, _ip, _ip, _ip]
        l.nspt_batch_kernel_shape.argtypes = [ctypes.c_void_p, _ip, _ip, _ip]
        l.nspt_batch_tune_candidates.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int]
        l.nspt_batch_form.argtypes = [ctypes.c_void_p]
        l.nspt_adjoint_drift.argtypes = [ctypes.POINTER(_Geom), _dp, ctypes.c_double, _dp, _dp]
        l.nspt_adjoint_drift_dev.argtypes = [ctypes.POINTER(_Geom), ctypes.c_void_p, ctypes.c_double,
                                             ctypes.c_void_p, ctypes.c_void_p]
        _lib = l
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Ranlxd:
    """Host noise source of the engine (ranlxd, MyMath.cc:295-322 mapping)."""

    def __init__(self, seed, level=1):
        h = ctypes.c_void_p()
        rc = lib().nspt_rng_create_ranlxd(seed, level, ctypes.byref(h))
        if rc != 0:
            raise NsptError("nspt_rng_create_ranlxd failed (%d)" % rc)
        self.h = h

    def uniform(self, n):
        out = np.zeros(n)
        lib().nspt_rng_uniform(self.h, n, _ptr(out))
        return out

    def noise(self, nc, nlinks):
        out = np.zeros((nlinks, nc * nc), np.complex128)
        lib().nspt_rng_noise(self.h, nc, nlinks, _ptr(out))
        return out

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.nspt_rng_destroy(self.h)
            self.h = None


class Rand55(Ranlxd):
    """Host noise source of the reference's shipped build: RAND55 (MyRand,
    MyRand.cc:30-130) seeded as Rand[0].init(seed), su(N) noise of SU3rand /
    SU2rand (MyMath.cc:256-292, 326-346)."""

    def __init__(self, seed):
        h = ctypes.c_void_p()
        rc = lib().nspt_rng_create_rand55(seed, ctypes.byref(h))
        if rc != 0:
            raise NsptError("nspt_rng_create_rand55 failed (%d)" % rc)
        self.h = h


class Context:
    """One lattice field on the GPU (an nspt_ctx)."""

    def __init__(self, L, nc, ptord, alloc_ord=None):
        self.L = tuple(int(x) for x in L)
        self.nc, self.N = nc, ptord
        self.A = ptord if alloc_ord is None else alloc_ord
        self.V = self.L[0] * self.L[1] * self.L[2] * self.L[3]
        g = _Geom((ctypes.c_int * 4)(*self.L), nc, ptord, self.A)
        h = ctypes.c_void_p()
        rc = lib().nspt_create(ctypes.byref(g), ctypes.byref(h))
        self.h = h
        if rc != 0:
            msg = lib().nspt_last_error(h).decode() if h.value else "nspt_create failed"
            self.close()
            raise NsptError("%s (%d)" % (msg, rc))
        self.obs_len = lib().nspt_obs_len(self.h)

    def _check(self, rc, what):
        if rc != 0:
            raise NsptError("%s: %s (%d)" % (what, lib().nspt_last_error(self.h).decode(), rc))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().nspt_destroy(self.h)
        self.h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    @property
    def S(self):
        return 1 + self.nc * self.nc * self.A

    def upload(self, U):
        U = np.ascontiguousarray(U, dtype=np.complex128)
        assert U.shape == (4 * self.V, self.S), U.shape
        self._check(lib().nspt_upload(self.h, _ptr(U)), "upload")

    def kernel_shape(self):
        """(tile edge, links per wave, waves per SIMD) of the kernel build
        this context runs."""
        v = [ctypes.c_int() for _ in range(3)]
        self._check(lib().nspt_kernel_shape(self.h, *[ctypes.byref(x) for x in v]), "kernel_shape")
        return tuple(x.value for x in v)

    def download(self):
        U = np.zeros((4 * self.V, self.S), np.complex128)
        self._check(lib().nspt_download(self.h, _ptr(U)), "download")
        return U

    def cold_start(self):
        self._check(lib().nspt_cold_start(self.h), "cold_start")

    def set_noise(self, eta):
        eta = np.ascontiguousarray(eta, dtype=np.complex128)
        assert eta.shape == (4 * self.V, self.nc * self.nc)
        self._check(lib().nspt_set_noise(self.h, _ptr(eta)), "set_noise")

    @staticmethod
    def params(eps, alpha=0.0, integrator=EULER, gauge_fix=GF_NONE, zero_modes=1, noise=NOISE_HOST,
               seed=0, step=0):
        return StepParams(eps, alpha, integrator, gauge_fix, zero_modes, noise, seed, step)

    def step(self, params, want_obs=True):
        obs = np.zeros(self.obs_len) if want_obs else None
        self._check(lib().nspt_step(self.h, ctypes.byref(params), _ptr(obs) if want_obs else None), "step")
        return obs

    def evolve(self, params, nsteps, want_obs=True):
        obs = np.zeros(nsteps * self.obs_len) if want_obs else None
        self._check(lib().nspt_evolve(self.h, ctypes.byref(params), nsteps, _ptr(obs) if want_obs else None),
                    "evolve")
        return obs.reshape(nsteps, self.obs_len) if want_obs else None

    def plaquette(self):
        out = np.zeros(self.N + 1, np.complex128)
        self._check(lib().nspt_plaquette(self.h, _ptr(out)), "plaquette")
        return out

    def loops(self):
        """Tr of the Wilson-loop sums w21, w22, w42, w44 (nspt.cc:1156-1312):
        array (4, N+1) complex, unnormalised."""
        out = np.zeros(4 * (self.N + 1), np.complex128)
        self._check(lib().nspt_loops(self.h, _ptr(out)), "loops")
        return out.reshape(4, self.N + 1)

    def set_action(self, c1):
        """Gauge action: c1 = 0 Wilson, -1/12 TLSYM, -0.331 IWA, -1.4088 DBW2
        (c0 = 1 - 8 c1); the observable row then ends with the rectangle
        plaquette."""
        self._check(lib().nspt_set_action(self.h, ctypes.c_double(c1)), "set_action")
        self.obs_len = lib().nspt_obs_len(self.h)

    def set_twist(self, mask, omegas=None):
        """Twisted boundary conditions: bit d of mask twists direction d,
        U(x + L_d e_d) = W_d U(x) W_d^dag (TWISTED_BC, nspt.cc:97-256).
        omegas: (4, nc, nc) monomial matrices, default reference_twist(nc);
        mask 0 restores periodic conditions."""
        om = reference_twist(self.nc) if omegas is None else np.asarray(omegas, np.complex128)
        om = np.ascontiguousarray(om.reshape(4, self.nc, self.nc))
        self._check(lib().nspt_set_twist(self.h, int(mask), _ptr(om)), "set_twist")

    def unitarize(self):
        """U <- exp(reH(log U)) for every link (QCDenvNODEpt_test.h:415-418)."""
        self._check(lib().nspt_unitarize(self.h), "unitarize")

    def fa_gauge_fixing(self, alpha):
        self._check(lib().nspt_fa_gauge_fixing(self.h, alpha), "fa_gauge_fixing")

    def momentum_norm(self):
        """knorm of the __K_MOM_ANALYSIS__ build (nspt.cc:919-944): (ptord, V)
        array, row k-1 = sum_mu Re Tr(Ut Ut^dag) of order k per momentum (site
        order), Ut the unnormalised forward 4-D DFT of the field."""
        out = np.zeros((self.N, self.V), np.float64)
        self._check(lib().nspt_momentum_norm(self.h, out.ctypes.data_as(_dp)), "momentum_norm")
        return out

    def gauge_wilson(self, params, want_obs=True):
        """The Langevin update stage alone (no zero modes, no gauge fixing)."""
        obs = np.zeros(self.obs_len) if want_obs else None
        self._check(lib().nspt_gauge_wilson(self.h, ctypes.byref(params), _ptr(obs) if want_obs else None),
                    "gauge_wilson")
        return obs

    def zero_modes_subtraction(self):
        self._check(lib().nspt_zero_modes_subtraction(self.h), "zero_modes_subtraction")

    def stochastic_gauge_fixing(self, alpha):
        """The sequential gauge fixing in checkerboard site order."""
        self._check(lib().nspt_stochastic_gauge_fixing(self.h, alpha), "stochastic_gauge_fixing")

    def drift(self, eps):
        D = np.zeros((4 * self.V, self.S), np.complex128)
        self._check(lib().nspt_drift(self.h, eps, _ptr(D)), "drift")
        return D

    def profile(self, enable=True):
        self._check(lib().nspt_profile(self.h, 1 if enable else 0), "profile")

    def profile_read(self):
        """{kernel: (total_ms, launches)} since the last profile() call."""
        ms = np.zeros(len(PROF_NAMES))
        cnt = (ctypes.c_longlong * len(PROF_NAMES))()
        self._check(lib().nspt_profile_read(self.h, _ptr(ms), cnt), "profile_read")
        return {n: (ms[i], int(cnt[i])) for i, n in enumerate(PROF_NAMES)}

    def device_noise(self, seed, step):
        eta = np.zeros((4 * self.V, self.nc * self.nc), np.complex128)
        self._check(lib().nspt_device_noise(self.h, seed, step, _ptr(eta)), "device_noise")
        return eta


class _ReplicaView(Context):
    """A replica of a Batch: a Context that the batch owns."""

    def __init__(self, h, L, nc, ptord):
        self.L = tuple(int(x) for x in L)
        self.nc, self.N, self.A = nc, ptord, ptord
        self.V = self.L[0] * self.L[1] * self.L[2] * self.L[3]
        self.h = h
        self.obs_len = lib().nspt_obs_len(self.h)

    def close(self):
        self.h = None


class Batch:
    """Independent replicas stepped together (nspt_batch_*): lattices L[r],
    step parameters params[r] (device noise, own seeds), common nc, ptord."""

    def __init__(self, Ls, nc, ptord, params):
        n = len(Ls)
        geoms = (_Geom * n)(*[_Geom((ctypes.c_int * 4)(*L), nc, ptord, ptord) for L in Ls])
        ps = (StepParams * n)(*params)
        h = ctypes.c_void_p()
        rc = lib().nspt_batch_create(geoms, ps, n, ctypes.byref(h))
        self.h = h
        if rc != 0:
            msg = lib().nspt_batch_last_error(h).decode() if h.value else "nspt_batch_create failed"
            self.close()
            raise NsptError("%s (%d)" % (msg, rc))
        self.replicas = []
        for r, L in enumerate(Ls):
            c = ctypes.c_void_p()
            lib().nspt_batch_replica(self.h, r, ctypes.byref(c))
            self.replicas.append(_ReplicaView(c, L, nc, ptord))

    def _check(self, rc, what):
        if rc != 0:
            raise NsptError("%s: %s (%d)" % (what, lib().nspt_batch_last_error(self.h).decode(), rc))

    def evolve(self, nsteps, want_obs=True):
        """Returns a list of (nsteps, obs_len) arrays, one per replica."""
        if not want_obs:
            self._check(lib().nspt_batch_evolve(self.h, nsteps, None), "batch_evolve")
            return None
        lens = [r.obs_len for r in self.replicas]
        obs = np.zeros(nsteps * sum(lens))
        self._check(lib().nspt_batch_evolve(self.h, nsteps, _ptr(obs)), "batch_evolve")
        out, o = [], 0
        for n in lens:
            out.append(obs[o:o + nsteps * n].reshape(nsteps, n))
            o += nsteps * n
        return out

    def set_form(self, table):
        """Step form: False (0) one graph branch per replica; True (1) one
        launch per stage over a table of the replicas; an int g > 1 the
        table form in g concurrent groups of replicas."""
        self._check(lib().nspt_batch_set_form(self.h, int(table)), "batch_set_form")

    @property
    def table(self):
        return lib().nspt_batch_form(self.h) >= 1

    @property
    def form(self):
        """0 branches, g >= 1 the table form in g groups."""
        return lib().nspt_batch_form(self.h)

    def tune(self, steps=3):
        """Returns (chosen tile edge, ms per step of each candidate: for each
        form -- branches, then the table in 1, 2 and 4 groups as the
        replicas allow -- the compiled tile edges, then the occupancy
        variants of the best edge in the best form; tune_candidates()
        describes them)."""
        q = ctypes.c_int()
        ms = np.zeros(16)
        self._check(lib().nspt_batch_tune(self.h, steps, ctypes.byref(q), _ptr(ms)), "batch_tune")
        return q.value, ms

    def tune_candidates(self):
        """The last tune()'s candidates as (form, tile edge, links per wave
        (0: the build's default), waves per SIMD) tuples, in its order."""
        d = (ctypes.c_int * 64)()
        n = lib().nspt_batch_tune_candidates(self.h, d, 16)
        self._check(0 if n >= 0 else n, "batch_tune_candidates")
        return [tuple(d[4 * i:4 * i + 4]) for i in range(n)]

    def kernel_shape(self):
        """(tile edge, links per wave, waves per SIMD) of the replicas' build."""
        v = [ctypes.c_int() for _ in range(3)]
        self._check(lib().nspt_batch_kernel_shape(self.h, *[ctypes.byref(x) for x in v]), "batch_kernel_shape")
        return tuple(x.value for x in v)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().nspt_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        if _lib is not None:
            self.close()


def adjoint_drift(L, nc, N, U, eps):
    """Reverse-mode gradient of sum Re Tr P_N and the drift -eps/(2nc) reH(U Sigma)
    of a generalised field U (4V, (N+1)*nc*nc), order 0 a matrix.  Returns (G, D)."""
    U = np.ascontiguousarray(U, np.complex128)
    g = _Geom((ctypes.c_int * 4)(*L), nc, N, N)
    G = np.zeros_like(U)
    D = np.zeros_like(U)
    rc = lib().nspt_adjoint_drift(ctypes.byref(g), _ptr(U), eps, _ptr(G), _ptr(D))
    if rc != 0:
        raise NsptError("nspt_adjoint_drift failed (%d)" % rc)
    return G, D


def measure_peaks():
    """(FP64 TFLOP/s of an FMA-chain kernel, HBM GB/s of a 1 GiB device copy)
    measured on the current GPU."""
    f, h = ctypes.c_double(), ctypes.c_double()
    rc = lib().nspt_measure_peaks(ctypes.byref(f), ctypes.byref(h))
    if rc != 0:
        raise NsptError("nspt_measure_peaks failed (%d)" % rc)
    return f.value, h.value


def split_obs(obs, N):
    """plaq (N+1 complex), norm (N), trU (4 x N complex); with an improved
    action also the rectangle plaquette (N+1 complex)."""
    plaq = obs[0:2 * (N + 1)].view(np.complex128)
    norm = obs[2 * (N + 1):2 * (N + 1) + N]
    o = 2 * (N + 1) + N
    tru = obs[o:o + 8 * N].view(np.complex128).reshape(4, N)
    if len(obs) > o + 8 * N:
        return plaq, norm, tru, obs[o + 8 * N:].view(np.complex128)
    return plaq, norm, tru
