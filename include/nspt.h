/* nspt.h -- C ABI of the MI355X NSPT Langevin engine (libnspt.so).
 *
 * Replaces the C++ function API of the reference's hot path (nspt.h:55-94):
 * NsptEvolve / gauge_wilson / zero_modes_subtraction /
 * FAstochastic_gauge_fixing / plaquette_measure operate there on a global
 * ptGluon_fld through process globals (Quenched.cc:42-67).  Here the same
 * operations act on an opaque context that owns the field in GPU memory.
 *
 * Conventions
 *  - Status: 0 = ok, negative = error; nspt_last_error() describes the last
 *    failure of a context.  No C++ exception crosses this boundary.
 *  - Host field image ("aos"): the byte layout of the reference's
 *    ptGluon[V] (QCDpt.h:47-48, 249-251; QCDenvNODEpt_test.h:338): for link
 *    4*site + mu, (1 + nc*nc*alloc_ord) complex doubles (re, im): the order-0
 *    flag, then order k = 1..alloc_ord as a row-major nc x nc matrix.  Orders
 *    above ptord are ignored on upload and written as zero on download.
 *    site = ((x0*L1 + x1)*L2 + x2)*L3 + x3 (lattice.h:37).
 *  - Noise: 4V matrices of nc*nc complex doubles in link order 4*site + mu,
 *    the consumption order of gauge_wilson (nspt.cc:365-424); the step uses
 *    -sqrt(eps)*eta at order 1 (nspt.cc:420-424).
 *  - Observables of a step ("obs", nspt_obs_len() doubles): plaquette per
 *    order k = 0..ptord (complex), norm per order 1..ptord (real), trU[mu]
 *    per order 1..ptord (complex), normalised as the reference writes its
 *    .txt/.nor/.trU files (nspt.cc:570-597).  They are measured on the field
 *    entering the step (nspt.cc:392-402).
 *  - One context per host thread; calls on a context are stream-ordered and
 *    return after the work is complete unless stated otherwise.
 */
#ifndef NSPT_MI355X_H
#define NSPT_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nspt_ctx nspt_ctx;

/* Lattice and series geometry (Quenched.cc:160-170, choices.h:30, 86-90). */
typedef struct {
  int sz[4];     /* taglia L0..L3 (periodic; see nspt_set_twist) */
  int nc;        /* 3 (SU(3)) or 2 (SU(2)) */
  int ptord;     /* PTORD: orders 1..ptord are evolved */
  int alloc_ord; /* allocORD: width of the host field image, >= ptord */
} nspt_geom;

/* NSPT_EULER: Jacobi Euler step (the reference built without UPDATE_ONTHEFLY).
 * NSPT_EULER_CHECKERBOARD: the in-place on-the-fly sweep of nspt.cc:365-441
 *   in checkerboard order (parity 0 then 1, mu = 0..3, each subset of links
 *   at once); Wilson action, even extents.
 * NSPT_GF_CHECKERBOARD: the sequential gauge fixing of nspt.cc:660-862 with
 *   the sites of parity 0, then 1; even extents. */
enum { NSPT_EULER = 0, NSPT_RK2 = 1, NSPT_EULER_CHECKERBOARD = 2 };
enum { NSPT_GF_NONE = 0, NSPT_GF_FA = 1, NSPT_GF_CHECKERBOARD = 2 };
enum { NSPT_NOISE_HOST = 0, NSPT_NOISE_DEVICE = 1, NSPT_NOISE_ZERO = 2 };

/* One Langevin step (nspt.cc:963-1033, Jacobi update of nspt.cc:428-429). */
typedef struct {
  double eps;        /* tau_g of Quench.cfg: Langevin time step */
  double alpha;      /* gauge-fixing alpha; FA is stable for alpha > 0 */
  int integrator;    /* NSPT_EULER, NSPT_RK2 or NSPT_EULER_CHECKERBOARD */
  int gauge_fix;     /* NSPT_GF_NONE, NSPT_GF_FA or NSPT_GF_CHECKERBOARD */
  int zero_modes;    /* 1: zero_modes_subtraction (REMOVE_GLUON_ZEROMOM) */
  int noise;         /* NSPT_NOISE_HOST (set by nspt_set_noise), _DEVICE, _ZERO */
  unsigned long long seed; /* device noise: Philox4x32-10 key */
  unsigned long long step; /* device noise: step counter (advanced by nspt_evolve) */
} nspt_step_params;

/* Context life cycle.  The field starts cold (flag 1, orders 0), which is
 * init_status 0 of Quenched.cc:596-604.  The context runs the default kernel
 * build of its order (the smallest compiled order >= ptord, tile edge 4,
 * the build's own links per wave, one wave per SIMD), or the build that the
 * environment names at creation when it is compiled: NSPT_TILE_EDGE,
 * NSPT_LINKS_PER_WAVE, NSPT_WAVES_PER_SIMD (A/B runs; results differ from
 * the default build's by rounding only). */
int nspt_create(const nspt_geom* geom, nspt_ctx** out);
int nspt_destroy(nspt_ctx* ctx);
const char* nspt_last_error(const nspt_ctx* ctx);
int nspt_obs_len(const nspt_ctx* ctx);
/* The kernel build a context runs: order-tile edge, links per wave (the
 * links one wave of a link kernel holds), waves per SIMD its register budget
 * allows.  No counterpart in the reference (a CPU code). */
int nspt_kernel_shape(const nspt_ctx* ctx, int* tile_edge, int* links_per_wave, int* waves_per_simd);

/* Field transfer in the ptGluon[V] image (Umu.load / Umu.save payload,
 * QCDenvNODEpt_test.h:342-374). */
int nspt_upload(nspt_ctx* ctx, const double* aos);
int nspt_download(nspt_ctx* ctx, double* aos);
int nspt_cold_start(nspt_ctx* ctx);

/* Host-generated noise for the next step (NSPT_NOISE_HOST), 4V x NE complex
 * in the reference's link order [4*site + mu][NE].  Copied before the call
 * returns: the caller may reuse eta at once. */
int nspt_set_noise(nspt_ctx* ctx, const double* eta);

/* NsptEvolve for one step; obs may be null.  nspt_step, nspt_evolve,
 * nspt_gauge_wilson and nspt_set_noise refuse a replica of a batch (its
 * steps are nspt_batch_evolve's). */
int nspt_step(nspt_ctx* ctx, const nspt_step_params* p, double* obs);

/* nsteps steps with device noise (p->noise must be NSPT_NOISE_DEVICE or
 * NSPT_NOISE_ZERO); obs (nsteps * nspt_obs_len doubles) may be null.
 * p->step is advanced by nsteps. */
int nspt_evolve(nspt_ctx* ctx, nspt_step_params* p, int nsteps, double* obs);

/* Gauge action of the drift (GAUGE_ACTION of choices.h:101-113, chosen at
 * compile time by the reference): c1 = 0 Wilson (default), -1/12 TLSYM,
 * -0.331 IWA, -1.4088 DBW2, c0 = 1 - 8 c1 (Quenched.cc:313-322).  With
 * c1 != 0 the drift is tau_g reH(c0 U Sigma + c1 U Sigma_2x1) (nspt.cc:385-412,
 * staple2x1 of QCDenvNODEpt_test.h:579-611) and the observable row gains the
 * rectangle plaquette (ptord+1 complex, normalised as the plaquette) after
 * trU: nspt_obs_len grows accordingly. */
int nspt_set_action(nspt_ctx* ctx, double c1);

/* Twisted boundary conditions (TWISTED_BC of choices.h:43-55, chosen at
 * compile time by the reference): bit d of mask twists direction d,
 * U(x + L_d e_d) = W_d U(x) W_d^dag, with W_d = omega + 2 nc^2 d (nc x nc,
 * row-major, interleaved complex; untwisted directions are not read).  The
 * reference keeps twisted copies of the boundary links in halo sites
 * (twist_boundary, nspt.cc:163-256; lattice.h:156-1053); here the links
 * read across a twisted boundary are conjugated as they are loaded by the
 * drift, rectangle, gauge-fixing and loop kernels.  Each W_d must be
 * monomial (one entry of modulus 1 per row and column), as all matrices of
 * twistmatrices.h are.  mask 0 restores periodic conditions.  The FA gauge
 * fixing needs periodic conditions and is refused on a twisted lattice. */
int nspt_set_twist(nspt_ctx* ctx, int mask, const double* omega);

/* plaquette_measure (nspt.cc:1316-1406): ptord+1 complex. */
int nspt_plaquette(nspt_ctx* ctx, double* out);

/* ComputeLoops (nspt.cc:1156-1312): Tr of the sums over sites and planes
 * mu < nu of the Wilson loops w21 (2x1 + 1x2), w22 (2x2), w42 (4x2 + 2x4)
 * and w44 (4x4), as the reference's l21, l22, l42, l44 before its output
 * normalisation: 4 x (ptord+1) complex, unnormalised. */
int nspt_loops(nspt_ctx* ctx, double* out);

/* U <- exp(reH(log U)) for every link: the re-unitarisation that
 * ptGluon_fld::loadlowerorder applies after reading a configuration of
 * fewer orders (QCDenvNODEpt_test.h:415-418). */
int nspt_unitarize(nspt_ctx* ctx);

/* The stages of NsptEvolve (nspt.cc:963-1033) one by one, in the order the
 * reference calls them: gauge_wilson, zero_modes_subtraction, then
 * stochastic_gauge_fixing or FAstochastic_gauge_fixing.  Together they are
 * nspt_step (which fuses them and is faster).
 *
 * gauge_wilson (nspt.cc:258-600): the Langevin update of every link with the
 * noise of p (integrator as in nspt_step; the zero_modes and gauge_fix
 * fields of p are ignored); obs as in nspt_step.  The Jacobi form writes the
 * new field into F and the reference then copies it back (Umu = Fmu,
 * nspt.cc:991-992): here the new field simply becomes the context's field. */
int nspt_gauge_wilson(nspt_ctx* ctx, const nspt_step_params* p, double* obs);

/* zero_modes_subtraction (nspt.cc:605-654): MM_mu = (1/V) sum_x log U_mu(x),
 * U <- exp(reH(log U - MM_mu)).  The reference accumulates MM over the links
 * gauge_wilson has just written; that is the current field, so the mean is
 * formed here from it (deterministic fixed-order sums). */
int nspt_zero_modes_subtraction(nspt_ctx* ctx);

/* stochastic_gauge_fixing (nspt.cc:660-862): per site W = sum_mu U_mu(x) -
 * U_mu(x-mu), W <- Trless(W - W^dag), U_mu(x) <- e^{alpha W} U_mu(x),
 * U_mu(x-mu) <- U_mu(x-mu) e^{-alpha W}.  The reference sweeps the sites
 * lexicographically in place; here the sites of parity 0, then 1 (sites of
 * one parity share no link), even extents only.  alpha in the reference's
 * convention (Quench.cfg alpha; the shipped value is -0.05). */
int nspt_stochastic_gauge_fixing(nspt_ctx* ctx, double alpha);

/* Stand-alone FAstochastic_gauge_fixing of the current field
 * (nspt.cc:866-948). */
int nspt_fa_gauge_fixing(nspt_ctx* ctx, double alpha);

/* Momentum norms of the __K_MOM_ANALYSIS__ build (nspt.cc:919-944, after the
 * gauge fixing of nspt.cc:833-857 and 919-944): the forward 4-D DFT of every
 * component of the field (planGluon[0], Quenched.cc:792-815: FFTW_FORWARD,
 * unnormalised), then per order k = 1..ptord and momentum p
 *   knorm[(k-1)*V + p] = sum_mu Re Tr(Ut_mu,k(p) Ut_mu,k(p)^dag),
 * p in the lattice's site order.  ptord * V doubles, the layout of the
 * reference's per-step record in knorm_*.dat (Quenched.cc:549-560), whose
 * fwrite reads knorm[0..allocORD-1] as one block.  The reference transforms
 * the field back and multiplies it by 1/V (a change by rounding only); here
 * the field is not touched.  Blocks until the values are on the host. */
int nspt_momentum_norm(nspt_ctx* ctx, double* knorm);

/* Langevin drift tau_g*reH(U*staples), tau_g = -eps/(2 nc), of the current
 * field for every link, in the host image layout with flag 0
 * (nspt.cc:383-417). */
int nspt_drift(nspt_ctx* ctx, double eps, double* aos_out);

/* Host noise sources, drawn in the reference's link order.
 * ranlxd (ranlxd/ranlxd.cc), the reference built without USE_RAND55: seeded
 * as the reference seeds Rand[0] (the first glibc rand() after srand(seed),
 * Quenched.cc:301-306) and mapped to su(nc) matrices by MyMath.cc:295-322;
 * SU(2) draws four uniforms per matrix.
 * RAND55 (MyRand, MyRand.cc:30-130), the generator of the shipped build
 * (Makefile:2 -DUSE_RAND55): seeded with Quench.cfg's seed as
 * Rand[0].init(seed) (Quenched.cc:296-299), Gaussian pairs of
 * MyRand::generate_gauss mapped by SU3rand / SU2rand (MyMath.cc:256-292,
 * 326-346). */
typedef struct nspt_rng nspt_rng;
int nspt_rng_create_ranlxd(int seed, int level, nspt_rng** out);
int nspt_rng_create_rand55(long seed, nspt_rng** out);
int nspt_rng_destroy(nspt_rng* rng);
int nspt_rng_uniform(nspt_rng* rng, int n, double* out);
int nspt_rng_noise(nspt_rng* rng, int nc, int nlinks, double* eta);

/* Device noise for inspection: 4V matrices (link order 4*site + mu) of
 * step `step` with key `seed`. */
int nspt_device_noise(nspt_ctx* ctx, unsigned long long seed, unsigned long long step, double* eta);

/* Per-kernel timing with HIP events recorded on the context stream around
 * every launch (no counterpart in the reference, whose MyTime.h timers are
 * host clocks).  nspt_profile(ctx, 1) resets and enables; the totals (ms)
 * and launch counts are indexed by the NSPT_PROF_* ids. */
enum {
  NSPT_PROF_DRIFT = 0,      /* staples + U*Sigma + reH (+ observables) */
  NSPT_PROF_UPDATE = 1,     /* exp(W) * U */
  NSPT_PROF_UPDATE_LOG = 2, /* exp(W) * U, then log for the zero mode */
  NSPT_PROF_ZERO_MODE = 3,  /* exp(reH(log U - MM)) */
  NSPT_PROF_FA_DIV = 4,     /* divergence of log U */
  NSPT_PROF_DFT = 5,        /* one 4-D DFT pass */
  NSPT_PROF_FA_GEXP = 6,    /* G = exp(reH(W)) per site */
  NSPT_PROF_TRANSFORM = 7,  /* U <- G U G^dag */
  NSPT_PROF_NOISE = 8,      /* device noise */
  NSPT_PROF_REDUCE = 9,     /* deterministic reductions */
  NSPT_PROF_RECT = 10,      /* rectangle staples of an improved action */
  NSPT_PROF_PLAQ_FIELD = 11, /* plaquette field of a Jacobi drift */
  NSPT_PROF_N = 12
};
int nspt_profile(nspt_ctx* ctx, int enable);
int nspt_profile_read(nspt_ctx* ctx, double* ms_total, long long* count);

/* Measured peaks of the current GPU (no counterpart in the reference; for
 * the roofline of bench.py): FP64 FMA throughput of a kernel of independent
 * FMA chains, in TFLOP/s, and the HBM rate of a 1 GiB device-to-device copy
 * (bytes read + written), in GB/s; best of three timed runs each.
 * Allocates 2 GiB for the duration of the call. */
int nspt_measure_peaks(double* fp64_tflops, double* hbm_gbs);

/* Independent replicas stepped together: the epsilon-sweep setting of
 * NSPT, where many runs of different volume and time step feed the
 * eps -> 0 extrapolation (no counterpart in the reference, which runs one
 * replica per process, Quenched.cc:262).  Replicas share nc and ptord; each
 * has its own geometry and step parameters and must draw device noise
 * (NSPT_NOISE_DEVICE) with its own seed.  One step of every replica is
 * captured once in a HIP graph and replayed, in one of two forms:
 *  - branches (the default): one graph branch per replica, each with the
 *    replica's own launches, running concurrently;
 *  - table: one launch per stage of the step for all replicas of a group,
 *    blockIdx.y selecting the replica whose fields, geometry, eps and seed
 *    the kernels read from a device table (replicas must share the
 *    integrator, the gauge fixing -- none or FA -- and the zero-mode
 *    choice); the replicas are dealt into 1 to 8 groups of about equal
 *    work, each group's launches a graph branch of its own, so that one
 *    group's stage fills the GPU where another's leaves its last round
 *    part-empty.
 * Either way the results equal nspt_evolve on each replica alone, bit for
 * bit. */
typedef struct nspt_batch nspt_batch;
int nspt_batch_create(const nspt_geom* geoms, const nspt_step_params* params, int nrep, nspt_batch** out);
int nspt_batch_destroy(nspt_batch* b);
const char* nspt_batch_last_error(const nspt_batch* b);
/* nsteps steps of every replica.  obs (may be null): replica-major, replica
 * r holding nsteps * nspt_obs_len(replica r) doubles. */
int nspt_batch_evolve(nspt_batch* b, int nsteps, double* obs);
/* Context of replica r (download, plaquette, ...), owned by the batch. */
int nspt_batch_replica(nspt_batch* b, int r, nspt_ctx** ctx);
/* Step form: 0 branches, g = 1..8 the table form in g groups (at most one
 * per replica; refused, -1, when the replicas do not allow the table form).
 * nspt_batch_form returns the form so set. */
int nspt_batch_set_form(nspt_batch* b, int table);
int nspt_batch_form(const nspt_batch* b);
/* Autotuner, in two stages.  Stage 1: for each step form (branches, then
 * when possible the table in 1, 2 and 4 groups, groups of at least two
 * replicas) and each compiled order-tile edge of the replicas' order, at the
 * default occupancy, time `steps` replays of the batch step.  Stage 2: in
 * the fastest form and edge so far, each compiled occupancy variant of that
 * edge (links per wave x waves per SIMD).  The fastest candidate is kept.
 * Fields and counters are restored afterwards, on failure too (and then the
 * form and kernel build too).  ms_per_step (may be null, up to 16 entries)
 * receives the measured times in candidate order, which
 * nspt_batch_tune_candidates describes. */
int nspt_batch_tune(nspt_batch* b, int steps, int* qt_chosen, double* ms_per_step);
/* The last nspt_batch_tune's candidates, 4 ints each: form (0 branches, g
 * the table in g groups), tile edge, links per wave (0: the build's
 * default), waves per SIMD.  Writes up to max; returns their number. */
int nspt_batch_tune_candidates(const nspt_batch* b, int* desc, int max);
/* The kernel build the batch's replicas run (as nspt_kernel_shape). */
int nspt_batch_kernel_shape(const nspt_batch* b, int* tile_edge, int* links_per_wave, int* waves_per_simd);

/* The Langevin drift as a reverse-mode gradient of the order-truncated
 * Wilson action (no counterpart in the reference, whose drift is the staple
 * sum of nspt.cc:383-417).  Generalised field: order 0 is a full NC x NC
 * matrix (e.g. Haar-random), so each link carries (ptord+1) matrices, order
 * 0 first, 4V links in link order 4*site + mu (no flag).  grad (may be
 * null, same layout): G = df/dRe U + i df/dIm U of
 * f = sum_{x, mu<nu} Re Tr P_{mu nu}(x)_{ptord}; one reverse pass seeded at
 * the top order yields every order of the staple sum, Sigma_m =
 * G_{ptord-m}^dag.  drift (may be null, same layout): -eps/(2 NC) reH(U Sigma)
 * per order 0..ptord.  Any ptord from 1 to 20: the instance compiled for
 * the next of 4, 6, 8, 12, 16, 20 runs with the orders above ptord zero
 * (-2 above 20); alloc_ord is ignored. */
int nspt_adjoint_drift(const nspt_geom* geom, const double* U, double eps, double* grad, double* drift);

/* nspt_adjoint_drift without any allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * U, grad and drift are DEVICE pointers in this layout (the adjoint's own): element e of
 * order k of link (site, mu) at complex index (k*nc*nc + e)*4V + mu*V + site,
 * k = 0..ptord (grad and drift may be null).  The six gradient slots per
 * link live in a static device array of 192 MiB held by each GPU until the
 * process ends: 6 * 4V * (ptord+1) * nc^2 * 16 bytes must fit (SU(3) 8^4 at
 * ptord 12 does), else -4 and nothing runs.  The kernels run on the null
 * stream and the call returns when they are done.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int nspt_adjoint_drift_dev(const nspt_geom* geom, const double* U, double eps, double* grad, double* drift);

#ifdef __cplusplus
}
#endif
#endif
