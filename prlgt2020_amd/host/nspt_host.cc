// nspt_host.cc -- the host drop-in's types and entry points over libnspt.so.
// Every Langevin step runs on the GPU; nothing here computes a series.
#include "nspt_host.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/nspt.h"

int PTORD = allocORD;
nspt_params_t nspt_pars;
act_params_t act_pars;
mi355x_opts_t mi355x_opts = {0, 1, 1, 0, 1};
Cplx* w = nullptr;
std::ofstream plaqfile, normfile, trufile, loop_nxm, loop2x2;
std::ofstream knormfile;

double Cplx::mod() const { return std::sqrt(re * re + im * im); }

void ptSU3::zero() {
  flag = Cplx();
  for (int k = 0; k < allocORD; k++)
    for (int e = 0; e < NC * NC; e++) ptU[k].whr[e] = Cplx();
}
void ptSU3::id() {
  zero();
  flag = Cplx(1.0, 0.0);
}

latt::latt(const int* sz) {
  Size = 1;
  for (int d = 0; d < dim; d++) {
    Sz[d] = sz[d];
    Size *= sz[d];
  }
  TSize = Size;
}
int latt::get(const int* n) const { return ((n[0] * Sz[1] + n[1]) * Sz[2] + n[2]) * Sz[3] + n[3]; }
int latt::get(int n, int step, int dir) const {
  int x[dim];
  for (int d = dim - 1; d >= 0; d--) {
    x[d] = n % Sz[d];
    n /= Sz[d];
  }
  x[dir] = ((x[dir] + step) % Sz[dir] + Sz[dir]) % Sz[dir];
  return get(x);
}

ptGluon_fld::ptGluon_fld(latt* z) : Z(z), W(new ptGluon[z->TSize]()) {}

int ptGluon_fld::save(const char* filename) {
  std::ofstream os(filename, std::ofstream::binary);
  os.write(reinterpret_cast<const char*>(W), sizeof(ptGluon) * Z->Size);
  if (os.rdstate() != 0) std::cerr << "Error writing " << filename << "\n";
  return 0;
}
int ptGluon_fld::save(const char* filename, const Cplx* placchetta) {
  std::ofstream os(filename, std::ofstream::binary);
  os.write(reinterpret_cast<const char*>(W), sizeof(ptGluon) * Z->Size);
  os.write(reinterpret_cast<const char*>(placchetta), sizeof(Cplx) * (allocORD + 1));
  if (os.rdstate() != 0) std::cerr << "Error writing " << filename << "\n";
  return 0;
}
int ptGluon_fld::load(const char* filename) {
  std::ifstream is(filename, std::ifstream::binary);
  if (!is) return -1;
  is.read(reinterpret_cast<char*>(W), sizeof(ptGluon) * Z->Size);
  return is ? 0 : -1;
}
int ptGluon_fld::load(const char* filename, Cplx* placchetta) {
  std::ifstream is(filename, std::ifstream::binary);
  if (!is) return -1;
  is.read(reinterpret_cast<char*>(W), sizeof(ptGluon) * Z->Size);
  is.read(reinterpret_cast<char*>(placchetta), sizeof(Cplx) * (allocORD + 1));
  return is ? 0 : -1;
}

namespace {

struct Engine {
  nspt_ctx* ctx = nullptr;
  nspt_rng* rng = nullptr;
  ptGluon_fld* field = nullptr;
  std::vector<double> eta, obs;
  unsigned long long step = 0;
};
Engine E;

[[noreturn]] void fail(const char* what) {
  std::fprintf(stderr, "MI355X engine: %s: %s\n", what, E.ctx ? nspt_last_error(E.ctx) : "no context");
  std::exit(1);
}

void attach(ptGluon_fld& U) {
  if (E.ctx && E.field == &U) return;
  mi355x_release();
  nspt_geom g = {{U.Z->Sz[0], U.Z->Sz[1], U.Z->Sz[2], U.Z->Sz[3]}, NC, PTORD, allocORD};
  if (nspt_create(&g, &E.ctx) != 0) fail("nspt_create");
  if (nspt_set_action(E.ctx, action_c1(mi355x_opts.action)) != 0) fail("nspt_set_action");
  if (mi355x_opts.twist) {
    double om[4 * 2 * NC * NC];
    reference_twist(om);
    if (nspt_set_twist(E.ctx, mi355x_opts.twist, om) != 0) fail("nspt_set_twist");
  }
  E.field = &U;
  if (nspt_upload(E.ctx, reinterpret_cast<const double*>(U.W)) != 0) fail("nspt_upload");
}

void write_step(const double* o) {
  const int N = PTORD;
  const double* plaq = o;                  // N+1 complex
  const double* nrm = o + 2 * (N + 1);     // N
  const double* tru = nrm + N;             // [mu][order] complex
  const double* w2 = tru + 8 * N;          // improved actions: rectangle plaquette, N+1 complex
  for (int i = 0; i <= N; i++) {
    plaqfile << plaq[2 * i] << "\t" << plaq[2 * i + 1];
    if (mi355x_opts.action != 0) plaqfile << "\t" << w2[2 * i] << "\t" << w2[2 * i + 1];  // nspt.cc:575-578
    plaqfile << std::endl;
  }
  for (int i = 0; i < N; i++) {
    normfile << nrm[i] << std::endl;
    for (int mu = 0; mu < dim; mu++)
      trufile << tru[2 * (mu * N + i)] << "\t" << tru[2 * (mu * N + i) + 1] << "\t";
    trufile << std::endl;
  }
}

}  // namespace

void reference_twist(double* om) {
  const double c = std::cos(2.0 * M_PI / 3.0), s = std::sin(2.0 * M_PI / 3.0);
  // (re, im) of the nonzero entry of each row and its column, per direction
  struct Row {
    int col;
    double re, im;
  };
#if NC == 3
  const Row rows[4][3] = {{{0, 1, 0}, {1, 1, 0}, {2, 1, 0}},
                          {{0, c, -s}, {1, 1, 0}, {2, c, s}},
                          {{1, 1, 0}, {2, 1, 0}, {0, 1, 0}},
                          {{1, c, -s}, {2, 1, 0}, {0, c, s}}};
#else
  (void)c;
  (void)s;
  const Row rows[4][2] = {{{0, 1, 0}, {1, 1, 0}},
                          {{0, 0, -1}, {1, 0, 1}},
                          {{1, 1, 0}, {0, -1, 0}},
                          {{1, 0, -1}, {0, 0, -1}}};
#endif
  for (int q = 0; q < 4 * 2 * NC * NC; q++) om[q] = 0.0;
  for (int d = 0; d < 4; d++)
    for (int i = 0; i < NC; i++) {
      const Row& r = rows[d][i];
      om[2 * ((d * NC + i) * NC + r.col)] = r.re;
      om[2 * ((d * NC + i) * NC + r.col) + 1] = r.im;
    }
}

double action_c1(int action) {
  switch (action) {
    case 1: return -1.0 / 12.0;  // TLSYM
    case 2: return -0.331;       // IWA
    case 3: return -1.4088;      // DBW2
    default: return 0.0;         // WIL
  }
}

void mi355x_seed_noise(unsigned long seed) {
  if (E.rng) nspt_rng_destroy(E.rng);
  E.rng = nullptr;
  if (mi355x_opts.device_noise == 2) {
    if (nspt_rng_create_rand55((long)seed, &E.rng) != 0) fail("nspt_rng_create_rand55");
  } else {
    std::srand((unsigned)seed);
    if (nspt_rng_create_ranlxd(std::rand(), 1, &E.rng) != 0) fail("nspt_rng_create_ranlxd");
  }
  mi355x_opts.noise_seed = seed;
}

namespace {

// Step parameters of the current options; noise drawn on the host (and
// uploaded) unless device noise is chosen.
nspt_step_params step_params() {
  const int integ[3] = {NSPT_EULER, NSPT_RK2, NSPT_EULER_CHECKERBOARD};
  const int gfix[3] = {NSPT_GF_NONE, NSPT_GF_FA, NSPT_GF_CHECKERBOARD};
  const bool dev = mi355x_opts.device_noise == 1;
  return {act_pars.stau * act_pars.stau, act_pars.alpha, integ[mi355x_opts.integrator], gfix[mi355x_opts.fa],
          mi355x_opts.zero_modes, dev ? NSPT_NOISE_DEVICE : NSPT_NOISE_HOST, mi355x_opts.noise_seed, E.step};
}

// __K_MOM_ANALYSIS__ record of the current field (nspt.cc:919-944)
void write_knorm(int V) {
  if (!mi355x_opts.knorm || !knormfile.is_open()) return;
  std::vector<double> rec((size_t)allocORD * V, 0.0);
  if (nspt_momentum_norm(E.ctx, rec.data()) != 0) fail("nspt_momentum_norm");
  knormfile.write(reinterpret_cast<const char*>(rec.data()), (std::streamsize)(sizeof(double) * rec.size()));
}

void host_noise(int V) {
  if (!E.rng) mi355x_seed_noise(mi355x_opts.noise_seed);
  E.eta.resize((size_t)4 * V * NC * NC * 2);
  nspt_rng_noise(E.rng, NC, 4 * V, E.eta.data());
  if (nspt_set_noise(E.ctx, E.eta.data()) != 0) fail("nspt_set_noise");
}

}  // namespace

void gauge_wilson(ptGluon_fld& Umu, ptGluon_fld& Fmu) {
  attach(Umu);
  if (mi355x_opts.integrator == 2) {
    std::fprintf(stderr, "MI355X engine: gauge_wilson(U, F) is the Jacobi update (-I 0 or 1)\n");
    std::exit(1);
  }
  // Umu keeps its value: its host image is brought up to date first, then
  // the engine's field becomes Fmu's
  mi355x_sync_to_host(Umu);
  nspt_step_params p = step_params();
  if (p.noise == NSPT_NOISE_HOST) host_noise(Umu.Z->Size);
  E.obs.resize(nspt_obs_len(E.ctx));
  if (nspt_gauge_wilson(E.ctx, &p, E.obs.data()) != 0) fail("nspt_gauge_wilson");
  E.step++;
  E.field = &Fmu;
  write_step(E.obs.data());
}

ptGluon_fld& ptGluon_fld::operator=(const ptGluon_fld& src) {
  if (this == &src) return *this;
  ptGluon_fld& s = const_cast<ptGluon_fld&>(src);
  mi355x_sync_to_host(s);
  std::memcpy(W, src.W, sizeof(ptGluon) * Z->Size);
  if (E.ctx && E.field == &s)
    E.field = this;  // the engine's field is this content
  else if (E.ctx && E.field == this)
    mi355x_sync_to_device(*this);
  return *this;
}

void zero_modes_subtraction(ptGluon_fld& Umu) {
  attach(Umu);
  if (nspt_zero_modes_subtraction(E.ctx) != 0) fail("nspt_zero_modes_subtraction");
}

void stochastic_gauge_fixing(ptGluon_fld& Umu) {
  attach(Umu);
  if (nspt_stochastic_gauge_fixing(E.ctx, act_pars.alpha) != 0) fail("nspt_stochastic_gauge_fixing");
  write_knorm(Umu.Z->Size);
}

void FAstochastic_gauge_fixing(ptGluon_fld& Umu) {
  attach(Umu);
  if (nspt_fa_gauge_fixing(E.ctx, act_pars.alpha) != 0) fail("nspt_fa_gauge_fixing");
  write_knorm(Umu.Z->Size);
}

void NsptEvolve(ptGluon_fld& Umu, ptGluon_fld& Fmu) {
  for (int t = 0; t < nspt_pars.Beat; t++) {
    gauge_wilson(Umu, Fmu);
    Umu = Fmu;  // nspt.cc:991-992
    if (mi355x_opts.zero_modes) zero_modes_subtraction(Umu);
    if (mi355x_opts.fa == 1) FAstochastic_gauge_fixing(Umu);
    if (mi355x_opts.fa == 2) stochastic_gauge_fixing(Umu);
  }
}

void NsptEvolve(ptGluon_fld& Umu) {
  attach(Umu);
  const int ol = nspt_obs_len(E.ctx);
  nspt_step_params p = step_params();
  const int beat = nspt_pars.Beat;
  E.obs.resize((size_t)beat * ol);
  // momentum norms after each step's gauge fixing: one step per call
  const bool kn = mi355x_opts.knorm && mi355x_opts.fa != 0;
  if (p.noise == NSPT_NOISE_DEVICE) {
    for (int t = 0; t < beat; t += kn ? 1 : beat) {
      if (nspt_evolve(E.ctx, &p, kn ? 1 : beat, E.obs.data() + (size_t)t * ol) != 0) fail("nspt_evolve");
      if (kn) write_knorm(Umu.Z->Size);
    }
    E.step = p.step;
  } else {
    for (int t = 0; t < beat; t++) {
      host_noise(Umu.Z->Size);
      if (nspt_step(E.ctx, &p, E.obs.data() + (size_t)t * ol) != 0) fail("nspt_step");
      E.step++;
      if (kn) write_knorm(Umu.Z->Size);
    }
  }
  for (int t = 0; t < beat; t++) write_step(E.obs.data() + (size_t)t * ol);
}

void mi355x_sync_to_host(ptGluon_fld& Umu) {
  if (!E.ctx || E.field != &Umu) return;
  if (nspt_download(E.ctx, reinterpret_cast<double*>(Umu.W)) != 0) fail("nspt_download");
}

void mi355x_sync_to_device(ptGluon_fld& Umu) {
  if (!E.ctx || E.field != &Umu) {
    attach(Umu);
    return;
  }
  if (nspt_upload(E.ctx, reinterpret_cast<const double*>(Umu.W)) != 0) fail("nspt_upload");
}

void plaquette_measure(ptGluon_fld& Umu, act_params_t& /*act*/) {
  attach(Umu);
  std::vector<double> out(2 * (PTORD + 1));
  if (nspt_plaquette(E.ctx, out.data()) != 0) fail("nspt_plaquette");
  for (int i = 0; i <= allocORD; i++) w[i] = (i <= PTORD) ? Cplx(out[2 * i], out[2 * i + 1]) : Cplx();
  mi355x_sync_to_host(Umu);
}

int ptGluon_fld::loadlowerorder(const char* filename, int loworder, Cplx* plaquette) {
  if (loworder > allocORD) {
    std::cout << "ERROR: trying to load more orders than the ones allocated." << std::endl;
    return -1;
  }
  std::ifstream handle(filename, std::ios::binary | std::ios::ate);
  if (!handle.is_open()) {
    std::cout << "ERROR: file " << filename << " not found." << std::endl;
    return -1;
  }
  const std::streamoff want =
      (std::streamoff)sizeof(double) * (2 * (NC * NC * loworder + 1) * dim * (std::streamoff)Z->Size + 2 * (loworder + 1));
  if (handle.tellg() != want) {
    std::cout << "ERROR: not possible to read the configuration, unexpected file dimension." << std::endl;
    return -1;
  }
  handle.seekg(0);
  for (int x = 0; x < Z->Size; x++)
    for (int mu = 0; mu < dim; mu++) {
      ptSU3& u = W[x].U[mu];
      handle.read(reinterpret_cast<char*>(&u.flag), sizeof(Cplx));
      for (int k = 0; k < loworder; k++) handle.read(reinterpret_cast<char*>(u.ptU[k].whr), sizeof(Cplx) * NC * NC);
      // every order above loworder is zero (the reference's loop starts one
      // order higher and relies on fresh, zero memory for order loworder+1)
      for (int k = loworder; k < allocORD; k++)
        for (int a = 0; a < NC * NC; a++) u.ptU[k].whr[a] = Cplx();
    }
  // the plaquette stored with it: loworder+1 Cplx
  handle.read(reinterpret_cast<char*>(plaquette), sizeof(Cplx) * (loworder + 1));
  if (!handle) {
    std::cout << "ERROR: reading " << filename << " failed." << std::endl;
    return -1;
  }
  // the field has to be unitary
  mi355x_sync_to_device(*this);
  if (nspt_unitarize(E.ctx) != 0) fail("nspt_unitarize");
  mi355x_sync_to_host(*this);
  return 0;
}

void ComputeLoops(ptGluon_fld& Umu) {
  attach(Umu);
  const int n1 = PTORD + 1;
  std::vector<double> l((size_t)4 * 2 * n1);  // l21, l22, l42, l44: Tr per order
  if (nspt_loops(E.ctx, l.data()) != 0) fail("nspt_loops");
  // normalisation of nspt.cc:1287-1311 (3*2*3 for either group, as there)
  const double loopNorm = 1. / (3 * 2 * 3 * (Umu.Z->Size));
  const double half[4] = {.5, 1., .5, 1.};
  for (int w = 0; w < 4; w++)
    for (int i1 = 0; i1 <= PTORD; i1 += 2) loop_nxm << l[2 * ((size_t)w * n1 + i1)] * half[w] * loopNorm << "\t";
  loop_nxm << std::endl;
}

void WL2x2(ptGluon_fld& Umu) {
  attach(Umu);
  const int n1 = PTORD + 1;
  std::vector<double> l((size_t)4 * 2 * n1);
  if (nspt_loops(E.ctx, l.data()) != 0) fail("nspt_loops");
  // the reference walks each 2x2 loop the other way round (mu > nu, along mu
  // first), which conjugates its trace: the real parts are those of l22.
  // Its loop prints ptU[i1] (order i1 + 1) for odd i1 <= PTORD (OpenMP form,
  // the default build).
  const double D18 = 1.0 / 18.0;  // MyMath.h:48
  for (int i1 = 1; i1 <= PTORD; i1 += 2) {
    const int k = i1 + 1;
    const double v = k <= PTORD ? l[2 * ((size_t)n1 + k)] : 0.0;
    loop2x2 << v / Umu.Z->Size * D18 << "\t";
  }
  loop2x2 << std::endl;
}

int plaquette_check(const Cplx* w1, const Cplx* w2) {
  for (int i = 0; i <= PTORD; i++)
    if ((w1[i] - w2[i]).mod() > 1e-7) return 1;
  return 0;
}

void mi355x_release() {
  if (E.ctx) nspt_destroy(E.ctx);
  E.ctx = nullptr;
  E.field = nullptr;
}

ptGluon_fld::~ptGluon_fld() {
  if (E.field == this) mi355x_release();
  delete[] W;
}
