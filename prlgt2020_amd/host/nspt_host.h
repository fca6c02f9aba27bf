// nspt_host.h -- host C++ drop-in for the reference's NSPT API, running the
// Langevin step on the MI355X through libnspt.so (include/nspt.h).
//
// Kept from the reference (same names, same memory layout, same meaning):
//   Cplx, SU3 matrix storage      MyMath.h:65-159, 505-520 (whr[NC*NC] row-major)
//   ptSU3 {flag; ptU[allocORD]}   QCDpt.h:42-52 (order-0 flag, orders 1..allocORD)
//   ptGluon {ptSU3 U[dim]}        QCDpt.h:249-251
//   latt (sizes, site index)      lattice.h:37, 81-233
//   ptGluon_fld (W, handle, get, save/load)   QCDenvNODEpt_test.h:331-426
//   nspt_params_t, act_params_t   input.h:35-70
//   NsptEvolve (one-field and Jacobi two-field forms), gauge_wilson,
//   zero_modes_subtraction, stochastic_gauge_fixing,
//   FAstochastic_gauge_fixing, plaquette_measure, plaquette_check
//                                 nspt.h:55-94, nspt.cc:952-1058
//   ComputeLoops                  nspt.h:94, nspt.cc:1156-1312
// The series types are data containers; the arithmetic of the step runs on
// the GPU.  The field lives in GPU memory between NsptEvolve calls: the host
// image W is refreshed by mi355x_sync_to_host (done by plaquette_measure and
// before every save in the driver).
//
// Compile-time switches as in the reference's choices.h: allocORD (default 6,
// choices.h:30) and NC (3; 2 for the SU(2) build, _runSU2).
#pragma once
#include <cstdio>
#include <fstream>
#include <iostream>

#ifndef allocORD
#define allocORD 6
#endif
#ifndef NC
#define NC 3
#endif

constexpr int dim = 4;
extern int PTORD;

struct Cplx {
  double re, im;
  Cplx(double r = 0.0, double i = 0.0) : re(r), im(i) {}
  Cplx operator-(const Cplx& b) const { return Cplx(re - b.re, im - b.im); }
  double mod() const;
};

struct SU3 {
  Cplx whr[NC * NC];
};

struct ptSU3 {
  Cplx flag;
  SU3 ptU[allocORD];
  void zero();
  void id();
};

struct ptGluon {
  ptSU3 U[dim];
};

static_assert(sizeof(ptSU3) == sizeof(double) * 2 * (1 + NC * NC * allocORD), "ptSU3 layout");
static_assert(sizeof(ptGluon) == dim * sizeof(ptSU3), "ptGluon layout");

// Periodic 4-D lattice, site = ((x0*L1 + x1)*L2 + x2)*L3 + x3 (lattice.h:37).
class latt {
 public:
  int Sz[dim];
  int Size;
  int TSize;
  explicit latt(const int* sz);
  int get(const int* n) const;
  // neighbour of site n at distance step (+1/-1) along dir
  int get(int n, int step, int dir) const;
};

class ptGluon_fld {
 public:
  latt* Z;
  ptGluon* W;
  explicit ptGluon_fld(latt* z);
  ~ptGluon_fld();
  ptGluon_fld(const ptGluon_fld&) = delete;
  // Umu = Fmu of the Jacobi update (nspt.cc:991-992): a copy of the field
  // (fetched from the GPU first when the engine holds it); the engine then
  // holds this field, with the same content and no upload.
  ptGluon_fld& operator=(const ptGluon_fld& src);
  ptSU3* handle() { return W->U; }
  ptSU3& get(int n, int mu) { return W[n].U[mu]; }
  ptGluon& get(int* n) { return W[Z->get(n)]; }
  // Raw image of W (QCDenvNODEpt_test.h:342-348) / with the plaquette
  // appended (allocORD+1 Cplx, QCDenvNODEpt_test.h:360-368).
  int save(const char* filename);
  int save(const char* filename, const Cplx* placchetta);
  int load(const char* filename);
  int load(const char* filename, Cplx* placchetta);
  // A configuration saved with loworder <= allocORD orders (its plaquette,
  // loworder+1 Cplx, appended): orders above loworder set to zero, every
  // link re-unitarised as exp(reH(log U)) on the engine
  // (QCDenvNODEpt_test.h:376-426).
  int loadlowerorder(const char* filename, int loworder, Cplx* plaquette);
};

typedef struct {
  int Sweep;
  int Beat;
  int Kill;
  int Save;
  int Init;
  char* plaqn;
  char* confn;
  char* damon;
  char* normn;
  char* trun;
  char* logn;
  char* name_nxm;
  char* name2x2;
} nspt_params_t;

typedef struct {
  int iVol;
  int* sz;
  double rVol;
  double tau_g;  // eps as read; -eps/(2 NC) after QuenchedAllocate (Quenched.cc:583-588)
  double stau;   // sqrt(eps)
  double alpha;
  double c0, c1;
} act_params_t;

// Engine choices the reference makes with compile-time switches.
typedef struct {
  int integrator;    // 0 Euler (Jacobi), 1 RK2, 2 Euler on the fly in checkerboard order
  int fa;            // gauge fixing every step: 0 none, 1 Fourier accelerated,
                     // 2 sequential in checkerboard site order
  int zero_modes;    // 1: REMOVE_GLUON_ZEROMOM
  int device_noise;  // 0: ranlxd stream of the reference (Rand[0]); 1: device Philox;
                     // 2: RAND55 stream of the shipped build (Rand[0].init(seed))
  unsigned long long noise_seed;
  int action;        // GAUGE_ACTION (choices.h:101-113): 0 WIL, 1 TLSYM, 2 IWA, 3 DBW2
  int twist;         // TWISTED_BC (choices.h:43-55): bit d twists direction d; 0 periodic
  int wl2x2;         // __WILLOOP_2x2__ (choices.h:77): write the 2x2 loop file
  int knorm;         // __K_MOM_ANALYSIS__: momentum norms after every gauge fixing
} mi355x_opts_t;

// The twist matrices of twistmatrices.h (not IDENTITY_TWIST), 4 x NC x NC
// row-major interleaved complex, x first; x is the identity (the
// reference's SU(2) x twist is the identity, its SU(3) one is undefined).
void reference_twist(double* om);

// c1 of an action (Quenched.cc:313-322); c0 = 1 - 8 c1.
double action_c1(int action);

extern nspt_params_t nspt_pars;
extern act_params_t act_pars;
extern mi355x_opts_t mi355x_opts;
// Plaquette of plaquette_measure (PTORD+1, normalised), as the reference's global w.
extern Cplx* w;
// Output streams written once per Langevin step (nspt.cc:541-597), and the
// Wilson-loop stream written by ComputeLoops.
extern std::ofstream plaqfile, normfile, trufile, loop_nxm, loop2x2;
// With mi355x_opts.knorm, every gauge fixing (nspt.cc:833-857, 919-944)
// appends allocORD * V doubles to this binary stream: per order k =
// 1..allocORD (zero above PTORD) and momentum p in site order, sum_mu Re
// Tr(Ut Ut^dag) of the forward 4-D DFT of the field (nspt_momentum_norm).
// The reference's fwrite of knorm[0] (nspt.cc:857, 944) reads its first
// array as if the allocORD arrays were one block; this is that block.
extern std::ofstream knormfile;

// Seeds the host noise stream as the reference seeds Rand[0]: srand(seed);
// ranlxd(rand()) (Quenched.cc:301-306), or, with device_noise 2, RAND55
// Rand[0].init(seed) (Quenched.cc:296-299).
void mi355x_seed_noise(unsigned long seed);
// nspt_pars.Beat Langevin steps (nspt.cc:952-1039); observables of every
// step go to plaqfile / normfile / trufile.  The field stays on the GPU for
// the whole beat (the stages fused, nspt_step).
void NsptEvolve(ptGluon_fld& Umu);
// The Jacobi form (UPDATE_ONTHEFLY off): per step gauge_wilson(Umu, Fmu);
// Umu = Fmu; zero_modes_subtraction(Umu); the gauge fixing of
// mi355x_opts.fa (nspt.cc:952-1033), each stage a call of the engine.
// Euler or RK2.  The results equal NsptEvolve(Umu)'s up to rounding; it
// copies the field to the host twice per step (the two Jacobi copies).
void NsptEvolve(ptGluon_fld& Umu, ptGluon_fld& Fmu);
// Fmu = the Langevin update of Umu with this step's noise, Umu unchanged;
// observables of the step to the output streams (nspt.cc:258-600).
void gauge_wilson(ptGluon_fld& Umu, ptGluon_fld& Fmu);
// U <- exp(reH(log U - MM_mu)), MM_mu the mean of log U_mu (nspt.cc:605-654).
void zero_modes_subtraction(ptGluon_fld& Umu);
// The sequential gauge fixing with alpha = act_pars.alpha (nspt.cc:660-862),
// sites of parity 0, then 1.
void stochastic_gauge_fixing(ptGluon_fld& Umu);
// The Fourier-accelerated gauge fixing with alpha = act_pars.alpha
// (nspt.cc:866-948).
void FAstochastic_gauge_fixing(ptGluon_fld& Umu);
// Plaquette per order into w (nspt.cc:1316-1406); refreshes the host image.
void plaquette_measure(ptGluon_fld& Umu, act_params_t& act);
// 0 if equal to 1e-7 per order (nspt.cc:1048-1058).
int plaquette_check(const Cplx* w1, const Cplx* w2);
// 1x2, 2x2, 2x4 and 4x4 Wilson loops of the current field (nspt.cc:1156-1312):
// one line on loop_nxm with the even orders 0, 2, .. <= PTORD of each, in
// the reference's normalisation.
void ComputeLoops(ptGluon_fld& Umu);
// 2x2 Wilson loops (nspt.cc:1064-1145, __WILLOOP_2x2__): one line on loop2x2
// with Re Tr of their sum over sites and planes at orders 2, 4, .. (PTORD+1
// when PTORD is odd, which the truncated products leave zero), times
// 1/(18 V).
void WL2x2(ptGluon_fld& Umu);
void mi355x_sync_to_host(ptGluon_fld& Umu);
void mi355x_sync_to_device(ptGluon_fld& Umu);
void mi355x_release();
