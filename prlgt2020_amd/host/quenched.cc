// quenched.cc -- the Quenched driver (Quenched.cc:126-871) over the MI355X
// engine: the same Quench.cfg and damocle.dag inputs, command-line options,
// output file names and formats, checkpoint files and sweep/beat structure.
//
//   quenched [options]     (reads ./Quench.cfg, ./damocle.dag by default)
//
// Options of the reference (Quenched.cc:193-256): -l L (all four sizes),
// -s SWEEP, -b BEAT, -r seed, -g tau_g, -a v (alpha = -v), -t init_status,
// -p plaquette file stem, -c configuration stem, -d damocle file, -n PTORD,
// -h.  Engine options (compile-time switches in the reference): -I 1 (RK2
// integrator) or 2 (the on-the-fly sweep in checkerboard order), -F 0 (no
// gauge fixing) or 2 (the sequential gauge fixing in checkerboard order),
// -Z 0 (keep zero modes), -D 1 (device noise instead of the reference's
// ranlxd stream) or 2 (the RAND55 stream of the shipped build, USE_RAND55),
// -J 1 (the Jacobi two-field NsptEvolve(Umu, Fmu), stage by stage),
// -A (improved gauge action), -L (restart from a
// configuration of fewer orders), -T (twisted boundary conditions), -W 1
// (the 2x2 Wilson loop file).  -I 2 -F 2 with the shipped Quench.cfg
// (alpha -0.05) is the reference's default build, reordered.
//
// The default gauge fixing is the Fourier-accelerated one (nspt.cc:866-948);
// it is stable for alpha > 0 only, so with it the shipped Quench.cfg (alpha
// -0.05, written for the sequential gauge fixing) needs "-a -0.05".
#include <algorithm>
#include <cmath>
#include <ctime>
#include <cstdlib>
#include <cstring>
#include <getopt.h>
#include <string>
#include <vector>

#include "nspt_host.h"

namespace {

std::ofstream logfile;
int nInt = 0;
int low_order = 0;      // -L: start from a configuration saved with this many orders
std::string low_confn;  // its file name
bool two_field = false;  // -J 1: NsptEvolve(Umu, Fmu), the Jacobi stage sequence
int config_counter = 0;
Cplx* w1 = nullptr;

// "KEY value" with the key checked (Quenched.cc:106-127)
int get_val(FILE* fp, const char* str, const char* fmt, void* val) {
  char c[128];
  if (1 != std::fscanf(fp, "%127s", c)) {
    std::fprintf(stderr, "Error reading input file at %s\n", str);
    std::exit(1);
  }
  if (std::strcmp(str, c) != 0) {
    std::fprintf(stderr, "Error reading input file expected %s found %s\n", str, c);
    std::exit(1);
  }
  if (1 != std::fscanf(fp, fmt, val)) {
    std::fprintf(stderr, "Error reading input file at %s\nCannot read value format %s\n", str, fmt);
    std::exit(1);
  }
  return 0;
}

// file-name tag of the group and action (Quenched.cc:325-347)
const char* type_g() {
  static const char* t3[4] = {"SU3_wil", "SU3_tls", "SU3_iwa", "SU3_dbw"};
  static const char* t2[4] = {"SU2_wil", "SU2_tls", "SU2_iwa", "SU2_dbw"};
  const int a = (mi355x_opts.action >= 0 && mi355x_opts.action < 4) ? mi355x_opts.action : 0;
  return NC == 3 ? t3[a] : t2[a];
}

void stamp(const char* what) {
  std::time_t t;
  std::time(&t);
  logfile << what << std::ctime(&t) << std::endl;
}

int initialize(int argc, char** argv) {
  unsigned long seed0 = 0;
  act_pars.sz = new int[dim];
  nspt_pars.plaqn = new char[256];
  nspt_pars.confn = new char[256];
  nspt_pars.damon = new char[256];
  nspt_pars.normn = new char[256];
  nspt_pars.logn = new char[256];
  nspt_pars.trun = new char[256];
  nspt_pars.name_nxm = new char[256];
  nspt_pars.name2x2 = new char[256];

  FILE* fp = std::fopen("Quench.cfg", "r");
  if (!fp) {
    std::cout << "Error: cannot read the configuration file Quench.cfg." << std::endl;
    return 1;
  }
  if (std::fscanf(fp, "taglia %d %d %d %d\n", &act_pars.sz[0], &act_pars.sz[1], &act_pars.sz[2], &act_pars.sz[3]) != 4) {
    std::fprintf(stderr, "Error reading input file at taglia\n");
    std::exit(1);
  }
  get_val(fp, "SWEEP", "%d", &nspt_pars.Sweep);
  get_val(fp, "BEAT", "%d", &nspt_pars.Beat);
  get_val(fp, "tau_g", "%lf", &act_pars.tau_g);
  get_val(fp, "alpha", "%lf", &act_pars.alpha);
  get_val(fp, "init_status", "%d", &nspt_pars.Init);
  get_val(fp, "plaq_out", "%255s", nspt_pars.plaqn);
  get_val(fp, "last_conf", "%255s", nspt_pars.confn);
  get_val(fp, "PTORD", "%d", &PTORD);
  get_val(fp, "seed", "%lu", &seed0);  // one thread: the first seed line (NTHR = 1)
  std::fclose(fp);
  std::strcpy(nspt_pars.damon, "damocle.dag");

  int opt;
  while ((opt = getopt(argc, argv, "l:s:b:d:r:g:a:t:p:c:n:I:F:Z:D:J:L:A:T:W:K:h")) != -1) {
    switch (opt) {
      case 'l': act_pars.sz[0] = act_pars.sz[1] = act_pars.sz[2] = act_pars.sz[3] = std::atoi(optarg); break;
      case 's': nspt_pars.Sweep = std::atoi(optarg); break;
      case 'b': nspt_pars.Beat = std::atoi(optarg); break;
      case 'r': seed0 = std::strtoul(optarg, nullptr, 10); break;
      case 'n': PTORD = std::atoi(optarg); break;
      case 'g': act_pars.tau_g = std::atof(optarg); break;
      case 'a': act_pars.alpha = -std::atof(optarg); break;
      case 't': nspt_pars.Init = std::atoi(optarg); break;
      case 'p': std::strncpy(nspt_pars.plaqn, optarg, 255); break;
      case 'c': std::strncpy(nspt_pars.confn, optarg, 255); break;
      case 'd': std::strncpy(nspt_pars.damon, optarg, 255); break;
      case 'I': mi355x_opts.integrator = std::atoi(optarg); break;
      case 'F': mi355x_opts.fa = std::atoi(optarg); break;
      case 'Z': mi355x_opts.zero_modes = std::atoi(optarg); break;
      case 'D': mi355x_opts.device_noise = std::atoi(optarg); break;
      case 'J': two_field = std::atoi(optarg) != 0; break;
      case 'L': low_order = std::atoi(optarg); break;
      case 'A': mi355x_opts.action = std::atoi(optarg); break;
      case 'T': mi355x_opts.twist = std::atoi(optarg); break;
      case 'W': mi355x_opts.wl2x2 = std::atoi(optarg); break;
      case 'K': mi355x_opts.knorm = std::atoi(optarg); break;
      case 'h':
        std::printf("\n-l\t\tlattice size (L0 = L1 = L2 = L3)\n-s\t\tnumber of iterations (SWEEP)\n"
                    "-b\t\titerations between damocle reads (BEAT)\n-r\t\trandom seed\n-g\t\ttau gluons\n"
                    "-a\t\talpha (alpha = -value)\n-t\t\tinitial state: 0 cold, otherwise from configuration\n"
                    "-p\t\tplaquette file name\n-c\t\tsaved configuration name\n-d\t\tdamocle file name\n"
                    "-n\t\tperturbative order\n-I\t\tintegrator: 0 Euler, 1 RK2, 2 Euler on the fly (checkerboard)\n"
                    "-F\t\tgauge fixing: 0 off, 1 Fourier accelerated, 2 sequential (checkerboard)\n"
                    "-Z\t\tzero-mode subtraction: 1 on, 0 off\n-D\t\tnoise: 0 host ranlxd, 1 device Philox, 2 host RAND55\n"
                    "-J\t\t1: Jacobi two-field NsptEvolve(Umu, Fmu), stage by stage\n"
                    "-L\t\twith -t 1: start from the configuration saved with this lower allocORD\n"
                    "-A\t\tgauge action: 0 Wilson, 1 TLSYM, 2 IWA, 3 DBW2\n"
                    "-T\t\ttwisted boundary conditions: bit 0 x, 1 y, 2 z, 3 t (6: y and z)\n"
                    "-W\t\t1: write the 2x2 Wilson loop file (__WILLOOP_2x2__)\n"
                    "-K\t\t1: write the momentum norms after every gauge fixing (__K_MOM_ANALYSIS__)\n\n");
        std::exit(0);
    }
  }
  if (PTORD < 1 || PTORD > allocORD) {
    std::fprintf(stderr, "PTORD = %d outside 1..allocORD = %d\n", PTORD, allocORD);
    return 1;
  }
  {
    FILE* df = std::fopen(nspt_pars.damon, "r");
    if (!df) {
      std::cout << "Error: damocle file " << nspt_pars.damon << " not found." << std::endl;
      std::exit(-1);
    }
    std::fclose(df);
  }
  mi355x_opts.noise_seed = seed0;
  if (mi355x_opts.device_noise < 0 || mi355x_opts.device_noise > 2) {
    std::fprintf(stderr, "-D takes 0, 1 or 2\n");
    return 1;
  }
  if (two_field && mi355x_opts.integrator == 2) {
    std::fprintf(stderr, "-J 1 is the Jacobi update: -I 0 or 1\n");
    return 1;
  }
  if (mi355x_opts.device_noise != 1) mi355x_seed_noise(seed0);
  if (mi355x_opts.integrator < 0 || mi355x_opts.integrator > 2 || mi355x_opts.fa < 0 || mi355x_opts.fa > 2) {
    std::fprintf(stderr, "-I takes 0, 1 or 2 and -F 0, 1 or 2\n");
    return 1;
  }
  if (mi355x_opts.twist < 0 || mi355x_opts.twist > 15 || (mi355x_opts.twist && mi355x_opts.fa == 1)) {
    std::fprintf(stderr, "-T takes 0..15, and twisted lattices need -F 0 or -F 2\n");
    return 1;
  }
  if (mi355x_opts.fa == 1 && act_pars.alpha <= 0.0)
    std::fprintf(stderr, "warning: Fourier-accelerated gauge fixing needs alpha > 0 (alpha = %g)\n", act_pars.alpha);

  // file names, non-MPI form (Quenched.cc:388-423)
  const double a = std::fabs(act_pars.alpha), tg = act_pars.tau_g;
  const int* z = act_pars.sz;
  std::string plaq_stem = nspt_pars.plaqn, conf_stem = nspt_pars.confn;
  std::snprintf(nspt_pars.name_nxm, 256, "%s_nxm_%s_nf0_%dx%dx%dx%d_a%2.2f_tg%3.3f_o%d.txt", plaq_stem.c_str(),
                type_g(), z[0], z[1], z[2], z[3], a, tg, PTORD);
  std::snprintf(nspt_pars.name2x2, 256, "%s2x2_%s_nf0_%dx%dx%dx%d_a%2.2f_tg%3.3f_o%d.txt", plaq_stem.c_str(),
                type_g(), z[0], z[1], z[2], z[3], a, tg, PTORD);
  std::snprintf(nspt_pars.plaqn, 256, "%s_%s_nf0_%dx%dx%dx%d_a%2.2f_tg%3.3f_o%d.txt", plaq_stem.c_str(), type_g(),
                z[0], z[1], z[2], z[3], a, tg, PTORD);
  std::snprintf(nspt_pars.confn, 256, "%s%s_nf0_%dx%dx%dx%d_a%2.2f_tg%3.3f_o%d", conf_stem.c_str(), type_g(), z[0],
                z[1], z[2], z[3], a, tg, allocORD);
  if (low_order > 0) {
    char nm[300];
    std::snprintf(nm, sizeof(nm), "%s%s_nf0_%dx%dx%dx%d_a%2.2f_tg%3.3f_o%d", conf_stem.c_str(), type_g(), z[0], z[1],
                  z[2], z[3], a, tg, low_order);
    low_confn = nm;
  }
  if (mi355x_opts.knorm && mi355x_opts.fa) {  // __K_MOM_ANALYSIS__ (Quenched.cc:549-560)
    char kname[300];
    std::snprintf(kname, sizeof(kname), "knorm_%s%s_nf0_%dx%dx%dx%d_a%2.2f_tg%3.3f_o%d.dat",
                  mi355x_opts.fa == 1 ? "acc_" : "", type_g(), z[0], z[1], z[2], z[3], a, tg, PTORD);
    knormfile.open(kname, std::ios::binary | std::ios::trunc);
    if (!knormfile.is_open()) {
      std::cout << "Error: cannot open " << kname << std::endl;
      return 1;
    }
  }
  std::snprintf(nspt_pars.logn, 256, "%s_nf0_%dx%dx%dx%d_a%2.2f_tg%3.3f_o%d.log", type_g(), z[0], z[1], z[2], z[3],
                a, tg, PTORD);
  std::snprintf(nspt_pars.normn, 256, "%s_nf0_%dx%dx%dx%d_a%2.2f_tg%3.3f_o%d.nor", type_g(), z[0], z[1], z[2],
                z[3], a, tg, PTORD);
  std::snprintf(nspt_pars.trun, 256, "%s_nf0_%dx%dx%dx%d_a%2.2f_tg%3.3f_o%d.trU", type_g(), z[0], z[1], z[2], z[3],
                a, tg, PTORD);

  plaqfile.open(nspt_pars.plaqn, std::ios_base::app);
  plaqfile.precision(30);
  plaqfile << std::scientific;
  if (mi355x_opts.wl2x2) {  // __WILLOOP_2x2__ (Quenched.cc:434-437)
    loop2x2.open(nspt_pars.name2x2, std::ios_base::app);
    loop2x2.precision(30);
    loop2x2 << std::scientific;
  }
  loop_nxm.open(nspt_pars.name_nxm, std::ios_base::app);  // __MANY_LOOP__ (choices.h:76)
  loop_nxm.precision(30);
  loop_nxm << std::scientific;
  normfile.open(nspt_pars.normn);
  logfile.open(nspt_pars.logn);
  trufile.open(nspt_pars.trun);
  if (!plaqfile.is_open()) {
    std::cout << "Error: cannot open " << nspt_pars.plaqn << std::endl;
    return 1;
  }

  logfile << "Gauge group\t\t" << (NC == 3 ? "SU3" : "SU2") << std::endl
          << (mi355x_opts.integrator == 2 ? "\tgauge field update \"on the fly\"" : "\tgauge field update not \"on the fly\"")
          << std::endl
          << "Gauge field boundary conditions\t\t" << (mi355x_opts.twist ? "TWISTED" : "PERIODIC") << std::endl;
  for (int d = 0; d < 4; d++)
    if ((mi355x_opts.twist >> d) & 1) logfile << "\t\t\t\t\t- " << "XYZT"[d] << " direction is twisted" << std::endl;
  logfile
          << "Taglia\t\t" << z[0] << " " << z[1] << " " << z[2] << " " << z[3] << std::endl
          << "Sweep\t\t" << nspt_pars.Sweep << std::endl
          << "Beat\t\t" << nspt_pars.Beat << std::endl
          << "Tau_g\t\t" << act_pars.tau_g << std::endl
          << "Alpha\t\t" << act_pars.alpha << std::endl
          << "Init status\t" << nspt_pars.Init << std::endl
          << "Ordine\t\t" << PTORD << std::endl
          << std::endl
          << "Placchetta\t" << nspt_pars.plaqn << std::endl
          << (mi355x_opts.wl2x2 ? std::string("Altri loops\t") + nspt_pars.name2x2 + "\n" : std::string())
          << "Altri loops\t" << nspt_pars.name_nxm << std::endl
          << "Configurazione\t" << nspt_pars.confn << std::endl
          << "Damocle\t\t" << nspt_pars.damon << std::endl
          << "Norma\t\t" << nspt_pars.normn << std::endl
          << "Tr(U)\t\t" << nspt_pars.trun << std::endl
          << "Gauge action\t" << type_g() + 4 << " (c1 = " << action_c1(mi355x_opts.action) << ")" << std::endl
          << "Integrator\t"
          << (mi355x_opts.integrator == 1 ? "RK2" : mi355x_opts.integrator == 2 ? "Euler, on the fly (checkerboard)"
                                                                                 : "Euler")
          << std::endl
          << "Gauge fixing\t"
          << (mi355x_opts.fa == 1 ? "Fourier accelerated" : mi355x_opts.fa == 2 ? "sequential (checkerboard)" : "none")
          << std::endl
          << "Noise\t\t"
          << (mi355x_opts.device_noise == 1 ? "device Philox4x32-10"
                                             : mi355x_opts.device_noise == 2 ? "RAND55 (Rand[0])" : "ranlxd (Rand[0])")
          << std::endl
          << std::endl;
  return 0;
}

void transform_tau() {
  act_pars.stau = std::sqrt(act_pars.tau_g);
  act_pars.tau_g /= -(2.0 * NC);  // _SQRT_BETA_EXPANSION_ (Quenched.cc:583-588)
}

int allocate(ptGluon_fld& Umu) {
  act_pars.iVol = act_pars.sz[0] * act_pars.sz[1] * act_pars.sz[2] * act_pars.sz[3];
  act_pars.rVol = 1.0 / act_pars.iVol;
  transform_tau();
  if (!nspt_pars.Init) {
    for (int i = 0; i < Umu.Z->TSize; i++)
      for (int mu = 0; mu < dim; mu++) Umu.W[i].U[mu].id();
  } else {
    const bool low = low_order > 0;
    if (low ? Umu.loadlowerorder(low_confn.c_str(), low_order, w1) : Umu.load(nspt_pars.confn, w1)) {
      std::cout << "Error reading the configuration " << (low ? low_confn.c_str() : nspt_pars.confn) << "."
                << std::endl;
      return 1;
    }
    plaquette_measure(Umu, act_pars);
    // a lower-order configuration carries its plaquette up to its own order only
    const int nchk = low ? std::min(PTORD, low_order) : PTORD;
    bool differ = false;
    for (int i = 0; i <= nchk; i++) differ = differ || (w[i] - w1[i]).mod() > 1e-7;
    if (differ) {
      std::cout << "The plaquette does not match the one stored with the configuration." << std::endl;
      for (int i = 0; i <= PTORD; i++) std::printf("%e %e\t%e %e\n", w[i].re, w[i].im, w1[i].re, w1[i].im);
    }
    std::cout << "\nPlaquette:" << std::endl;
    for (int i = 0; i <= PTORD; i++) std::printf("%e %e\n", w[i].re, w[i].im);
  }
  stamp("Inizio run\t\t");
  return 0;
}

// After every BEAT steps: the Wilson loops, then damocle.dag re-read
// (Quenched.cc:656-725)
int update_parameters(ptGluon_fld& Umu) {
  if (mi355x_opts.wl2x2) WL2x2(Umu);
  ComputeLoops(Umu);
  nInt += nspt_pars.Beat;
  logfile << "______________________________________" << std::endl << std::endl;
  stamp("Damocle Update\t\t");
  logfile << "# Iterazioni\t\t" << nInt << std::endl;
  FILE* df = std::fopen(nspt_pars.damon, "r");
  if (!df) {
    std::cout << "Error: damocle file " << nspt_pars.damon << " not found." << std::endl;
    std::exit(-1);
  }
  get_val(df, "SWEEP", "%d", &nspt_pars.Sweep);
  get_val(df, "BEAT", "%d", &nspt_pars.Beat);
  get_val(df, "tau_g", "%lf", &act_pars.tau_g);
  get_val(df, "alpha", "%lf", &act_pars.alpha);
  get_val(df, "save_config", "%d", &nspt_pars.Save);
  get_val(df, "kill_run", "%d", &nspt_pars.Kill);
  std::fclose(df);
  logfile << "Tau_g\t\t\t" << act_pars.tau_g << std::endl << std::endl;
  logfile.flush();
  transform_tau();
  if (nspt_pars.Save == 1) {
    plaquette_measure(Umu, act_pars);
    char name[300];
    std::snprintf(name, sizeof(name), "%s_%d.dat", nspt_pars.confn, ++config_counter);
    Umu.save(name, w);
  }
  if (nspt_pars.Kill == 1) nspt_pars.Sweep = 0;
  return 0;
}

int finalize(ptGluon_fld& Umu, int t) {
  logfile << "__________________________________" << std::endl;
  stamp("Termine simulazione\t\t");
  logfile << "# Iterazioni\t\t" << t << std::endl;
  plaquette_measure(Umu, act_pars);
  std::printf("Plaquette:\n");
  for (int i = 0; i <= PTORD; i++) std::printf("%e\t%e\n", w[i].re, w[i].im);
  Umu.save(nspt_pars.confn, w);
  if (mi355x_opts.wl2x2) {
    WL2x2(Umu);
    loop2x2.close();
  }
  if (knormfile.is_open()) knormfile.close();  // NsptFinalize (Quenched.cc:719-721)
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  w = new Cplx[allocORD + 1];
  w1 = new Cplx[allocORD + 1];
  if (initialize(argc, argv)) return 1;
  latt LL(act_pars.sz);
  ptGluon_fld Umu(&LL);
  ptGluon_fld Fmu(&LL);  // the Jacobi form's second field (Quenched.cc:633-636)
  if (allocate(Umu)) return 1;
  int t1 = 0;
  while (t1 < nspt_pars.Sweep) {
    t1 += nspt_pars.Beat;
    if (two_field)
      NsptEvolve(Umu, Fmu);
    else
      NsptEvolve(Umu);
    if (update_parameters(Umu)) return 0;
  }
  finalize(Umu, t1);
  mi355x_release();
  return 0;
}
