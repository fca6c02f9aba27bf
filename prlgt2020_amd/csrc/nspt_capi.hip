// nspt_capi.hip -- host side of libnspt.so: context, device buffers, the
// order-independent kernels (layout conversion, noise, reductions, the 4-D
// DFT of the Fourier-accelerated gauge fixing) and the C ABI of
// include/nspt.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nspt.h"
#include "nspt_device.hpp"
#include "nspt_internal.hpp"

using nspt::Geo;
using nspt::KernelSet;
using nspt::SerField;

namespace {

// ------------------------------------------------------------------ kernels

// Host image (ptGluon[V], width A) -> device SoA field (orders < nrt).
__global__ void k_aos_to_soa(const double2* aos, int A, int NE, int nrt, int V, SerField U) {
  const long P = 4L * V;
  const long n = (long)nrt * NE * P;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long idx = t % P;
    const long ke = t / P;  // k*NE + e
    const long site = idx % V, mu = idx / V;
    U.data[t] = aos[(4 * site + mu) * (1 + (long)NE * A) + 1 + ke];
  }
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < P; idx += (long)gridDim.x * blockDim.x) {
    const long site = idx % V, mu = idx / V;
    U.flag[idx] = aos[(4 * site + mu) * (1 + (long)NE * A)];
  }
}

// Device SoA field -> host image; orders >= nrt and a missing flag array
// are written as zero.
__global__ void k_soa_to_aos(SerField U, int A, int NE, int nrt, int V, double2* aos) {
  const long P = 4L * V;
  const long S = 1 + (long)NE * A;
  const long n = P * S;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long link = t / S, j = t % S;
    const long site = link / 4, mu = link % 4;
    const long idx = mu * V + site;
    double2 v = make_double2(0.0, 0.0);
    if (j == 0) {
      if (U.flag) v = U.flag[idx];
    } else if ((j - 1) / NE < nrt) {
      v = U.data[(j - 1) * P + idx];
    }
    aos[t] = v;
  }
}

// eta: reference link order [4*site+mu][NE] <-> device [mu*V+site][NE]
__global__ void k_eta_permute(const double2* in, double2* out, int NE, int V, int to_device) {
  const long n = 4L * V * NE;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long e = t % NE, l = t / NE;
    if (to_device) {
      const long site = l / 4, mu = l % 4;
      out[(mu * V + site) * NE + e] = in[t];
    } else {
      const long mu = l / V, site = l % V;
      out[(4 * site + mu) * NE + e] = in[t];
    }
  }
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; round++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
    c[1] = (uint32_t)p1;
    c[3] = (uint32_t)p0;
    c[0] = n0;
    c[2] = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Device noise: link (site, mu) of step `step` uses counters
// (4*site+mu, step_lo, step_hi, q), q = 0 (and 1 for SU(3)), key = seed;
// uniforms (u + 0.5) 2^-32; Box-Muller with variance 1/2 and the su(N)
// mapping of MyMath.cc:295-322.
// step_dev (may be null): a device counter added to step, so that a step
// replayed from a captured graph draws fresh noise.
__global__ void k_noise_philox(int nc, int V, unsigned long long seed, unsigned long long step0,
                               const unsigned long long* step_dev, double2* eta) {
  const unsigned long long step = step0 + (step_dev ? *step_dev : 0ull);
  for (long idx = blockIdx.x * (long)blockDim.x + threadIdx.x; idx < 4L * V; idx += (long)gridDim.x * blockDim.x) {
  const long mu = idx / V, site = idx % V;
  const uint32_t link = (uint32_t)(4 * site + mu);
  double g[8];
  const int nq = (nc == 3) ? 2 : 1;
  for (int q = 0; q < nq; q++) {
    uint32_t c[4] = {link, (uint32_t)step, (uint32_t)(step >> 32), (uint32_t)q};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    for (int w = 0; w < 4; w++) g[4 * q + w] = ((double)c[w] + 0.5) * (1.0 / 4294967296.0);
  }
  const double twopi = 6.283185307179586476925286766559;
  const int ng = (nc == 3) ? 8 : 4;
  for (int i = 0; i < ng; i += 2) {
    const double t = twopi * g[i];
    const double r = sqrt(-log(1.0 - g[i + 1]));
    g[i] = r * cos(t);
    g[i + 1] = r * sin(t);
  }
  if (nc == 3) {
    const double g7 = g[7] * sqrt(1.0 / 3.0);
    double2* o = eta + idx * 9;
    o[0] = make_double2(0, g7 + g[6]);
    o[1] = make_double2(g[1], g[0]);
    o[2] = make_double2(g[3], g[2]);
    o[3] = make_double2(-g[1], g[0]);
    o[4] = make_double2(0, g7 - g[6]);
    o[5] = make_double2(g[5], g[4]);
    o[6] = make_double2(-g[3], g[2]);
    o[7] = make_double2(-g[5], g[4]);
    o[8] = make_double2(0, -g7 * 2);
  } else {
    double2* o = eta + idx * 4;
    o[0] = make_double2(0, g[2]);
    o[1] = make_double2(g[1], g[0]);
    o[2] = make_double2(-g[1], g[0]);
    o[3] = make_double2(0, -g[2]);
  }
  }
}

// Deterministic column sums: out[chunk][c] = sum of rows
// [row0 + chunk*rows_per, ...) of in (row stride ld), optional scale.
__global__ void k_reduce_rows(const double* in, long ld, long row0, long nrows, int ncols, long rows_per,
                              double* out, double scale) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncols) return;
  const long r0 = row0 + blockIdx.y * rows_per;
  const long r1 = min(row0 + nrows, r0 + rows_per);
  double s = 0.0;
  for (long r = r0; r < r1; r++) s += in[r * ld + col];
  out[(long)blockIdx.y * ncols + col] = s * scale;
}

// Per-step observables from the per-mu block sums (layout per row:
// plaq (N+1) cplx, norm N, trU N cplx), normalised as nspt.cc:570-597.
// Advances the noise step and observable row counters of a captured step.
__global__ void k_count(unsigned long long* ctr) {
  if (threadIdx.x == 0) {
    ctr[0] += 1;
    ctr[1] += 1;
  }
}

// slot_dev (may be null): a device counter selecting the output row, for
// steps replayed from a captured graph.
__global__ void k_obs_finalize(const double* obs_mu, int nrt, int nc, int V, double* out,
                               const unsigned long long* slot_dev, int obs_len) {
  if (slot_dev) out += (long)(*slot_dev) * obs_len;
  const int nob = 2 * (nrt + 1) + 3 * nrt;
  const double pn = (nc == 3 ? 72.0 : 48.0) * V, nn = (nc == 3 ? 12.0 : 8.0) * V, tn = (nc == 3 ? 3.0 : 2.0) * V;
  for (int t = threadIdx.x; t < 2 * (nrt + 1) + nrt; t += blockDim.x) {
    double s = 0.0;
    for (int mu = 0; mu < 4; mu++) s += obs_mu[mu * nob + t];
    out[t] = s / (t < 2 * (nrt + 1) ? pn : nn);
  }
  for (int t = threadIdx.x; t < 8 * nrt; t += blockDim.x) {
    const int mu = t / (2 * nrt), q = t % (2 * nrt);
    out[2 * (nrt + 1) + nrt + t] = obs_mu[mu * nob + 2 * (nrt + 1) + nrt + q] / tn;
  }
}

// W(x) = sum_mu X_mu(x) - X_mu(x-mu) per component (nspt.cc:871-877);
// fb[comp*V + site].
__global__ void k_fa_div(Geo g, int ncomp, const double2* X, double2* fb) {
  const long V = g.V;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < ncomp * V; t += (long)gridDim.x * blockDim.x) {
    const long comp = t / V;
    const int site = (int)(t % V);
    const double2* xc = X + comp * 4 * V;
    double2 w = make_double2(0.0, 0.0);
    for (int mu = 0; mu < 4; mu++) {
      const int x = (site / g.st[mu]) % g.L[mu];
      const int xm = (x == 0) ? site + (g.L[mu] - 1) * g.st[mu] : site - g.st[mu];
      const double2 a = xc[mu * V + site], b = xc[mu * V + xm];
      w = nspt::c_add(w, nspt::c_sub(a, b));
    }
    fb[t] = w;
  }
}

// One pass of the unnormalised 4-D DFT along direction d (sign -1 forward,
// +1 backward, FFTW convention of Quenched.cc:827-838) on an LDS tile of
// L_d x T elements (T consecutive sites of the fastest other direction):
// a mixed-radix Stockham FFT of each line, one output per thread per stage,
// radices the prime factors of L_d (smallest first), ping-ponging between
// two tile buffers so that the output comes out in natural order.  With
// scale != 0 the output is multiplied by -alpha/(V * p^2(k)),
// p^2 = sum_mu 4 sin^2(pi k_mu / L_mu), and the k = 0 mode is set to zero
// (nspt.cc:881-894, 902).
__global__ void k_dft(Geo g, double2* fb, int d, int sign, int scale, double alpha) {
  extern __shared__ double2 sm[];
  const int n = g.L[d];
  const int bdim = (d == 3) ? 2 : 3;
  const int T = g.L[bdim];
  const long V = g.V;
  const long tiles = V / ((long)n * T);
  const long comp = blockIdx.x / tiles;
  long tt = blockIdx.x % tiles;
  // the two remaining directions, slower one first
  int f[2], q = 0;
  for (int dd = 0; dd < 4; dd++)
    if (dd != d && dd != bdim) f[q++] = dd;
  const int x1 = (int)(tt % g.L[f[1]]);
  const int x0 = (int)(tt / g.L[f[1]]);
  const long base = (long)x0 * g.st[f[0]] + (long)x1 * g.st[f[1]];
  const int tid = threadIdx.x;
  double2* X = sm;
  double2* Y = sm + n * T;
  double2* tw = sm + 2 * n * T;  // tw[k] = exp(sign 2 pi i k / n)
  // per stage and output index: first input, input stride, twiddle step of
  // the r-point DFT and the output twiddle, built once per block
  int* tab = reinterpret_cast<int*>(tw + n);  // [stage][n][4]
  int a = 0, b = 0;
  const bool act = tid < n * T;
  if (act) {
    if (d < 3) {
      a = tid / T;
      b = tid % T;
    } else {
      a = tid % n;
      b = tid / n;
    }
  }
  const long site = base + (long)a * g.st[d] + (long)b * g.st[bdim];
  double2* line = fb + comp * V;
  if (act) X[a * T + b] = line[site];
  if (tid < n) {
    double s, c;
    sincospi(2.0 * (double)tid / (double)n, &s, &c);
    tw[tid] = make_double2(c, sign * s);
    // stages: length len = r * m, stride st; output o = q + st*(r*p + t) is
    // the r-point DFT of x[q + st*(p + u*m)], u = 0..r-1, times w_len^{p t}
    int len = n, st = 1, stage = 0;
    while (len > 1) {
      int r = 2;
      while (len % r) r++;
      const int m = len / r;
      const int qq = tid % st, rest = tid / st, t = rest % r, p = rest / r;
      int* e = tab + (stage * n + tid) * 4;
      e[0] = qq + st * p;
      e[1] = st * m;
      e[2] = (t * (n / r)) % n;
      e[3] = (p * t * (n / len)) % n;
      len = m;
      st *= r;
      stage++;
    }
  }
  int len = n, stage = 0;
  while (len > 1) {
    int r = 2;
    while (len % r) r++;
    __syncthreads();
    if (act) {
      const int* e = tab + (stage * n + a) * 4;
      int in = e[0], wi = 0;
      const int istep = e[1], wstep = e[2];
      double2 acc = make_double2(0.0, 0.0);
      for (int u = 0; u < r; u++) {
        const double2 v = X[in * T + b];
        const double2 w = tw[wi];
        acc.x = fma(v.x, w.x, acc.x);
        acc.x = fma(-v.y, w.y, acc.x);
        acc.y = fma(v.x, w.y, acc.y);
        acc.y = fma(v.y, w.x, acc.y);
        in += istep;
        wi += wstep;
        wi -= (wi >= n) ? n : 0;
      }
      Y[a * T + b] = nspt::c_mul(acc, tw[e[3]]);
    }
    double2* sw = X;
    X = Y;
    Y = sw;
    len /= r;
    stage++;
  }
  __syncthreads();
  if (!act) return;
  double2 acc = X[a * T + b];
  if (scale) {
    double p2 = 0.0;
    for (int mu = 0; mu < 4; mu++) {
      const int k = (int)((site / g.st[mu]) % g.L[mu]);
      const double h = sinpi((double)k / (double)g.L[mu]);
      p2 += 4.0 * h * h;
    }
    const double fac = (site == 0) ? 0.0 : -alpha / (p2 * (double)V);
    acc = nspt::c_scale(fac, acc);
  }
  line[site] = acc;
}

// Smallest compiled N >= n; among those the tile edge qt if compiled, else
// the default edge 4.
KernelSet find_set(int nc, int n, int qt = 4) {
  KernelSet best;
  auto consider = [&](KernelSet s) {
    if (s.nc != nc || s.nmax < n) return;
    if (best.nmax == 0 || s.nmax < best.nmax) {
      best = s;
    } else if (s.nmax == best.nmax && s.qt == qt && best.qt != qt) {
      best = s;
    }
  };
#define NSPT_CONSIDER(a, b, q) consider(nspt::kernel_set_##a##_##b##_q##q());
  NSPT_INSTANCES(NSPT_CONSIDER)
#undef NSPT_CONSIDER
  if (best.nmax && best.qt != qt) {
    KernelSet dflt = best;
    best = KernelSet();
    auto prefer4 = [&](KernelSet s) {
      if (s.nc == nc && s.nmax == dflt.nmax && s.qt == 4) best = s;
    };
#define NSPT_CONSIDER4(a, b, q) prefer4(nspt::kernel_set_##a##_##b##_q##q());
    NSPT_INSTANCES(NSPT_CONSIDER4)
#undef NSPT_CONSIDER4
    if (!best.nmax) best = dflt;
  }
  return best;
}

int grid_for(long n, int block) {
  long g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (int)g;
}

}  // namespace

// ------------------------------------------------------------------ context
struct nspt_ctx {
  nspt_geom geom{};
  int nc = 3, NE = 9, nrt = 0, A = 0, V = 0;
  KernelSet ks;
  Geo g{};
  int nblk_links = 0, nblk_sites = 0, nob = 0, obs_len = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  SerField U{}, F{}, D{}, D1{}, L{}, X{}, G{};
  double2* fb = nullptr;
  double2* eta = nullptr;
  double2* eta_tmp = nullptr;
  double2* mm_part = nullptr;
  double2* MM = nullptr;
  double* obs_part = nullptr;
  double* obs_mu = nullptr;
  double* red_tmp = nullptr;
  double* obs_dev = nullptr;
  int obs_cap = 0;
  double2* aos_buf = nullptr;
  bool noise_ready = false;
  // Captured-graph mode: noise step and observable row come from ctr_dev
  // ({step offset, row}), advanced by k_count at the end of each step.
  bool graph = false;
  unsigned long long* ctr_dev = nullptr;
  std::string err;
  // per-kernel HIP-event timing (nspt_profile)
  bool prof = false;
  struct Mark {
    int id;
    hipEvent_t a, b;
  };
  std::vector<Mark> marks;
  std::vector<hipEvent_t> pool;
  double prof_ms[NSPT_PROF_N] = {};
  long long prof_cnt[NSPT_PROF_N] = {};
  hipEvent_t get_event() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
  }

  int fail(const std::string& m) {
    err = m;
    return -1;
  }
  bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
  }
  template <class T>
  bool alloc(T** p, size_t n) {
    void* q = nullptr;
    if (!hip_ok(hipMalloc(&q, n * sizeof(T) + 16), "hipMalloc")) return false;
    allocs.push_back(q);
    *p = static_cast<T*>(q);
    return true;
  }
  bool alloc_field(SerField& f, long P, bool with_flag) {
    f.P = P;
    f.flag = nullptr;
    if (!alloc(&f.data, (size_t)nrt * NE * P)) return false;
    if (with_flag && !alloc(&f.flag, (size_t)P)) return false;
    return true;
  }
};

namespace {

bool launch_ok(nspt_ctx* c, const char* what) { return c->hip_ok(hipGetLastError(), what); }

// Brackets a kernel launch with HIP events on the context stream when
// profiling is on.
struct Prof {
  nspt_ctx* c;
  nspt_ctx::Mark m;
  Prof(nspt_ctx* c_, int id) : c(c_) {
    m.id = id;
    if (c->prof) {
      m.a = c->get_event();
      m.b = c->get_event();
      (void)hipEventRecord(m.a, c->stream);
    }
  }
  ~Prof() {
    if (c->prof) {
      (void)hipEventRecord(m.b, c->stream);
      c->marks.push_back(m);
    }
  }
};

void prof_collect(nspt_ctx* c) {
  if (c->marks.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto& m : c->marks) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, m.a, m.b) == hipSuccess) {
      c->prof_ms[m.id] += ms;
      c->prof_cnt[m.id] += 1;
    }
    c->pool.push_back(m.a);
    c->pool.push_back(m.b);
  }
  c->marks.clear();
}

// deterministic reduction of rows [row0, row0+nrows) of a (rows x ncols)
// matrix into out[ncols] (times scale)
bool reduce(nspt_ctx* c, const double* in, long ld, long row0, long nrows, int ncols, double* out, double scale) {
  Prof pr(c, NSPT_PROF_REDUCE);
  const long per = 256;
  const long nch = (nrows + per - 1) / per;
  dim3 g1((ncols + 63) / 64, (unsigned)nch);
  hipLaunchKernelGGL(k_reduce_rows, g1, dim3(64), 0, c->stream, in, ld, row0, nrows, ncols, per, c->red_tmp, 1.0);
  dim3 g2((ncols + 63) / 64, 1);
  hipLaunchKernelGGL(k_reduce_rows, g2, dim3(64), 0, c->stream, (const double*)c->red_tmp, (long)ncols, 0L, nch,
                     ncols, nch, out, scale);
  return launch_ok(c, "reduce");
}

bool obs_reduce(nspt_ctx* c, double* out_dev) {
  for (int mu = 0; mu < 4; mu++)
    if (!reduce(c, c->obs_part, c->nob, (long)mu * c->g.nb_mu, c->g.nb_mu, c->nob, c->obs_mu + mu * c->nob, 1.0))
      return false;
  Prof pr(c, NSPT_PROF_REDUCE);
  hipLaunchKernelGGL(k_obs_finalize, dim3(1), dim3(256), 0, c->stream, (const double*)c->obs_mu, c->nrt, c->nc, c->V,
                     out_dev, c->graph ? (const unsigned long long*)(c->ctr_dev + 1) : nullptr, c->obs_len);
  return launch_ok(c, "obs_finalize");
}

bool fa_apply(nspt_ctx* c, const SerField& Xsrc, double alpha) {
  const long ncomp = (long)c->nrt * c->NE;
  {
    Prof pr(c, NSPT_PROF_FA_DIV);
    hipLaunchKernelGGL(k_fa_div, dim3(grid_for(ncomp * c->V, 256)), dim3(256), 0, c->stream, c->g, (int)ncomp,
                     (const double2*)Xsrc.data, c->fb);
  }
  if (!launch_ok(c, "fa_div")) return false;
  const int order_fwd[4] = {3, 2, 1, 0};
  for (int pass = 0; pass < 8; pass++) {
    const int d = pass < 4 ? order_fwd[pass] : order_fwd[7 - pass];
    const int sign = pass < 4 ? -1 : +1;
    const int n = c->g.L[d], T = c->g.L[d == 3 ? 2 : 3];
    const long tiles = (long)c->V / ((long)n * T) * ncomp;
    const int block = ((n * T + 63) / 64) * 64;
    int nstage = 0;
    for (int len = n; len > 1; nstage++) {
      int r = 2;
      while (len % r) r++;
      len /= r;
    }
    const size_t shm = sizeof(double2) * (2 * n * T + n) + sizeof(int) * 4 * (size_t)nstage * n;
    Prof pr(c, NSPT_PROF_DFT);
    hipLaunchKernelGGL(k_dft, dim3((unsigned)tiles), dim3(block), shm, c->stream, c->g, c->fb, d, sign,
                       pass == 3 ? 1 : 0, alpha);
    if (!launch_ok(c, "dft")) return false;
  }
  {
    Prof pr(c, NSPT_PROF_FA_GEXP);
    c->ks.gexp(c->nblk_sites, c->stream, c->g, c->fb, c->G);
  }
  {
    Prof pr(c, NSPT_PROF_TRANSFORM);
    c->ks.transform(c->nblk_links, c->stream, c->g, c->G, c->U);
  }
  return launch_ok(c, "fa gexp/transform");
}

bool make_noise(nspt_ctx* c, const nspt_step_params* p, const double2** eta) {
  if (p->noise == NSPT_NOISE_ZERO) {
    *eta = nullptr;
    return true;
  }
  if (p->noise == NSPT_NOISE_DEVICE) {
    Prof pr(c, NSPT_PROF_NOISE);
    hipLaunchKernelGGL(k_noise_philox, dim3(grid_for(4L * c->V, 256)), dim3(256), 0, c->stream, c->nc, c->V,
                       p->seed, p->step, c->graph ? (const unsigned long long*)c->ctr_dev : nullptr, c->eta);
    *eta = c->eta;
    return launch_ok(c, "noise");
  }
  if (!c->noise_ready) {
    c->err = "nspt_step: NSPT_NOISE_HOST requires nspt_set_noise before each step";
    return false;
  }
  c->noise_ready = false;
  *eta = c->eta;
  return true;
}

// One step; observables to obs_dev (device) if non-null.
bool do_step(nspt_ctx* c, const nspt_step_params* p, double* obs_dev) {
  const double stau = std::sqrt(p->eps);
  const double tau_g = -p->eps / (2.0 * c->nc);
  const double rk2c = c->nc * p->eps / 6.0;
  const double2* eta = nullptr;
  if (!make_noise(c, p, &eta)) return false;
  const bool need_log = p->zero_modes || p->gauge_fix == NSPT_GF_FA;
  SerField none{nullptr, nullptr, 0};
  double* obs_part = obs_dev ? c->obs_part : nullptr;
  if (p->integrator == NSPT_RK2) {
    {
      Prof pr(c, NSPT_PROF_DRIFT);
      c->ks.drift(c->nblk_links, c->stream, c->g, c->U, c->D1, none, tau_g, 0.0, 0, obs_part);
    }
    {
      Prof pr(c, NSPT_PROF_UPDATE);
      c->ks.update(c->nblk_links, c->stream, c->g, c->D1, eta, stau, c->U, c->F, none, nullptr);
    }
    {
      Prof pr(c, NSPT_PROF_DRIFT);
      c->ks.drift(c->nblk_links, c->stream, c->g, c->F, c->D, c->D1, tau_g, rk2c, 1, nullptr);
    }
  } else {
    Prof pr(c, NSPT_PROF_DRIFT);
    c->ks.drift(c->nblk_links, c->stream, c->g, c->U, c->D, none, tau_g, 0.0, 0, obs_part);
  }
  if (!launch_ok(c, "drift")) return false;
  if (obs_dev && !obs_reduce(c, obs_dev)) return false;
  {
    Prof pr(c, need_log ? NSPT_PROF_UPDATE_LOG : NSPT_PROF_UPDATE);
    c->ks.update(c->nblk_links, c->stream, c->g, c->D, eta, stau, c->U, c->F, need_log ? c->L : none,
                 p->zero_modes ? c->mm_part : nullptr);
  }
  if (!launch_ok(c, "update")) return false;
  std::swap(c->U, c->F);  // Umu = Fmu (nspt.cc:992)
  const SerField* xsrc = &c->L;
  if (p->zero_modes) {
    const int ncol = 2 * c->nrt * c->NE;
    for (int mu = 0; mu < 4; mu++)
      if (!reduce(c, reinterpret_cast<const double*>(c->mm_part), ncol, (long)mu * c->g.nb_mu, c->g.nb_mu, ncol,
                  reinterpret_cast<double*>(c->MM) + (long)mu * ncol, 1.0 / (double)c->V))
        return false;
    {
      Prof pr(c, NSPT_PROF_ZERO_MODE);
      c->ks.zero_mode(c->nblk_links, c->stream, c->g, c->L, c->MM, c->U, c->X);
    }
    if (!launch_ok(c, "zero_mode")) return false;
    xsrc = &c->X;
  }
  if (p->gauge_fix == NSPT_GF_FA && !fa_apply(c, *xsrc, p->alpha)) return false;
  return true;
}

bool ensure_obs(nspt_ctx* c, int n) {
  if (c->obs_cap >= n) return true;
  if (!c->alloc(&c->obs_dev, (size_t)n * c->obs_len)) return false;
  c->obs_cap = n;
  return true;
}

}  // namespace

// ================================================================== C ABI
extern "C" {

int nspt_create(const nspt_geom* geom, nspt_ctx** out) {
  if (!geom || !out) return -1;
  *out = nullptr;
  nspt_ctx* c = new nspt_ctx();
  c->geom = *geom;
  c->nc = geom->nc;
  c->NE = geom->nc * geom->nc;
  c->nrt = geom->ptord;
  c->A = geom->alloc_ord;
  auto bad = [&](const char* m) {
    c->err = m;
    *out = c;
    return -2;
  };
  if (c->nc != 2 && c->nc != 3) return bad("nspt_create: nc must be 2 or 3");
  if (c->nrt < 1) return bad("nspt_create: ptord must be >= 1");
  if (c->A < c->nrt) return bad("nspt_create: alloc_ord must be >= ptord");
  long V = 1;
  for (int d = 0; d < 4; d++) {
    if (geom->sz[d] < 2 || geom->sz[d] > 32) return bad("nspt_create: each lattice extent must be in [2, 32]");
    V *= geom->sz[d];
  }
  c->V = (int)V;
  c->ks = find_set(c->nc, c->nrt);
  if (!c->ks.drift) return bad("nspt_create: no compiled kernel instance for this order (SU(3) up to 32, SU(2) up to 24)");
  for (int d = 0; d < 4; d++) c->g.L[d] = geom->sz[d];
  c->g.st[3] = 1;
  c->g.st[2] = c->g.L[3];
  c->g.st[1] = c->g.L[2] * c->g.L[3];
  c->g.st[0] = c->g.L[1] * c->g.L[2] * c->g.L[3];
  c->g.V = c->V;
  c->g.nrt = c->nrt;
  c->g.nb_mu = (c->V + c->ks.lpb - 1) / c->ks.lpb;
  c->nblk_links = 4 * c->g.nb_mu;
  c->nblk_sites = c->g.nb_mu;
  c->nob = 2 * (c->nrt + 1) + 3 * c->nrt;
  c->obs_len = 2 * (c->nrt + 1) + c->nrt + 8 * c->nrt;
  *out = c;
  if (!c->hip_ok(hipStreamCreate(&c->stream), "hipStreamCreate")) return -3;
  const long P = 4L * V;
  bool ok = c->alloc_field(c->U, P, true) && c->alloc_field(c->F, P, true) && c->alloc_field(c->D, P, false) &&
            c->alloc_field(c->D1, P, false) && c->alloc_field(c->L, P, true) && c->alloc_field(c->X, P, false) &&
            c->alloc_field(c->G, V, true) &&
            c->alloc(&c->fb, (size_t)c->nrt * c->NE * V) && c->alloc(&c->eta, (size_t)P * c->NE) &&
            c->alloc(&c->eta_tmp, (size_t)P * c->NE) &&
            c->alloc(&c->mm_part, (size_t)c->nblk_links * c->nrt * c->NE) &&
            c->alloc(&c->MM, (size_t)4 * c->nrt * c->NE) && c->alloc(&c->obs_part, (size_t)c->nblk_links * c->nob) &&
            c->alloc(&c->obs_mu, (size_t)4 * c->nob) &&
            c->alloc(&c->red_tmp, (size_t)(c->nblk_links / 256 + 2) * 2 * c->nrt * c->NE + 4096) &&
            c->alloc(&c->aos_buf, (size_t)P * (1 + (size_t)c->NE * c->A)) && c->alloc(&c->ctr_dev, 2) &&
            ensure_obs(c, 1);
  if (!ok) return -3;
  return nspt_cold_start(c);
}

int nspt_destroy(nspt_ctx* c) {
  if (!c) return 0;
  prof_collect(c);
  for (hipEvent_t e : c->pool) (void)hipEventDestroy(e);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return 0;
}

const char* nspt_last_error(const nspt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int nspt_obs_len(const nspt_ctx* c) { return c ? c->obs_len : -1; }

int nspt_cold_start(nspt_ctx* c) {
  if (!c) return -1;
  const long P = 4L * c->V;
  if (!c->hip_ok(hipMemsetAsync(c->U.data, 0, sizeof(double2) * (size_t)c->nrt * c->NE * P, c->stream), "memset"))
    return -3;
  std::vector<double2> ones(P, make_double2(1.0, 0.0));
  if (!c->hip_ok(hipMemcpyAsync(c->U.flag, ones.data(), sizeof(double2) * P, hipMemcpyHostToDevice, c->stream),
                 "memcpy"))
    return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_upload(nspt_ctx* c, const double* aos) {
  if (!c || !aos) return -1;
  const size_t n = 4UL * c->V * (1 + (size_t)c->NE * c->A);
  if (!c->hip_ok(hipMemcpyAsync(c->aos_buf, aos, n * sizeof(double2), hipMemcpyHostToDevice, c->stream), "upload"))
    return -3;
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_for((long)n, 256)), dim3(256), 0, c->stream, (const double2*)c->aos_buf,
                     c->A, c->NE, c->nrt, c->V, c->U);
  if (!launch_ok(c, "aos_to_soa")) return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

static int download_field(nspt_ctx* c, const SerField& f, double* aos) {
  const size_t n = 4UL * c->V * (1 + (size_t)c->NE * c->A);
  hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_for((long)n, 256)), dim3(256), 0, c->stream, f, c->A, c->NE, c->nrt,
                     c->V, c->aos_buf);
  if (!launch_ok(c, "soa_to_aos")) return -3;
  if (!c->hip_ok(hipMemcpyAsync(aos, c->aos_buf, n * sizeof(double2), hipMemcpyDeviceToHost, c->stream), "download"))
    return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_download(nspt_ctx* c, double* aos) {
  if (!c || !aos) return -1;
  return download_field(c, c->U, aos);
}

int nspt_set_noise(nspt_ctx* c, const double* eta) {
  if (!c || !eta) return -1;
  const size_t n = 4UL * c->V * c->NE;
  if (!c->hip_ok(hipMemcpyAsync(c->eta_tmp, eta, n * sizeof(double2), hipMemcpyHostToDevice, c->stream), "noise"))
    return -3;
  hipLaunchKernelGGL(k_eta_permute, dim3(grid_for((long)n, 256)), dim3(256), 0, c->stream,
                     (const double2*)c->eta_tmp, c->eta, c->NE, c->V, 1);
  if (!launch_ok(c, "eta_permute")) return -3;
  c->noise_ready = true;
  return 0;
}

static int check_params(nspt_ctx* c, const nspt_step_params* p) {
  if (!p) return c->fail("null step parameters");
  if (!(p->eps > 0.0)) return c->fail("eps must be > 0");
  if (p->integrator != NSPT_EULER && p->integrator != NSPT_RK2) return c->fail("unknown integrator");
  if (p->gauge_fix != NSPT_GF_NONE && p->gauge_fix != NSPT_GF_FA) return c->fail("unknown gauge fixing");
  if (p->noise < NSPT_NOISE_HOST || p->noise > NSPT_NOISE_ZERO) return c->fail("unknown noise source");
  return 0;
}

int nspt_step(nspt_ctx* c, const nspt_step_params* p, double* obs) {
  if (!c) return -1;
  if (check_params(c, p)) return -2;
  if (!do_step(c, p, obs ? c->obs_dev : nullptr)) return -3;
  if (obs && !c->hip_ok(hipMemcpyAsync(obs, c->obs_dev, sizeof(double) * c->obs_len, hipMemcpyDeviceToHost,
                                       c->stream),
                        "obs"))
    return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_evolve(nspt_ctx* c, nspt_step_params* p, int nsteps, double* obs) {
  if (!c) return -1;
  if (check_params(c, p)) return -2;
  if (p->noise == NSPT_NOISE_HOST) return c->fail("nspt_evolve needs device or zero noise") - 1;
  if (nsteps < 0) return c->fail("nsteps < 0") - 1;
  if (obs && !ensure_obs(c, nsteps)) return -3;
  for (int s = 0; s < nsteps; s++) {
    if (!do_step(c, p, obs ? c->obs_dev + (size_t)s * c->obs_len : nullptr)) return -3;
    p->step++;
  }
  if (obs && nsteps > 0 &&
      !c->hip_ok(hipMemcpyAsync(obs, c->obs_dev, sizeof(double) * c->obs_len * nsteps, hipMemcpyDeviceToHost,
                                c->stream),
                 "obs"))
    return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_plaquette(nspt_ctx* c, double* out) {
  if (!c || !out) return -1;
  SerField none{nullptr, nullptr, 0};
  c->ks.drift(c->nblk_links, c->stream, c->g, c->U, none, none, 0.0, 0.0, 2, c->obs_part);
  if (!launch_ok(c, "plaquette") || !obs_reduce(c, c->obs_dev)) return -3;
  if (!c->hip_ok(hipMemcpyAsync(out, c->obs_dev, sizeof(double) * 2 * (c->nrt + 1), hipMemcpyDeviceToHost,
                                c->stream),
                 "plaq"))
    return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_fa_gauge_fixing(nspt_ctx* c, double alpha) {
  if (!c) return -1;
  c->ks.logf(c->nblk_links, c->stream, c->g, c->U, c->L);
  if (!launch_ok(c, "log") || !fa_apply(c, c->L, alpha)) return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_drift(nspt_ctx* c, double eps, double* aos_out) {
  if (!c || !aos_out) return -1;
  SerField none{nullptr, nullptr, 0};
  c->ks.drift(c->nblk_links, c->stream, c->g, c->U, c->D, none, -eps / (2.0 * c->nc), 0.0, 0, nullptr);
  if (!launch_ok(c, "drift")) return -3;
  return download_field(c, c->D, aos_out);
}

int nspt_device_noise(nspt_ctx* c, unsigned long long seed, unsigned long long step, double* eta) {
  if (!c || !eta) return -1;
  const size_t n = 4UL * c->V * c->NE;
  hipLaunchKernelGGL(k_noise_philox, dim3(grid_for(4L * c->V, 256)), dim3(256), 0, c->stream, c->nc, c->V, seed,
                     step, (const unsigned long long*)nullptr, c->eta_tmp);
  hipLaunchKernelGGL(k_eta_permute, dim3(grid_for((long)n, 256)), dim3(256), 0, c->stream,
                     (const double2*)c->eta_tmp, c->eta, c->NE, c->V, 0);
  if (!launch_ok(c, "device_noise")) return -3;
  if (!c->hip_ok(hipMemcpyAsync(eta, c->eta, n * sizeof(double2), hipMemcpyDeviceToHost, c->stream), "noise"))
    return -3;
  c->noise_ready = false;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_profile(nspt_ctx* c, int enable) {
  if (!c) return -1;
  prof_collect(c);
  for (int i = 0; i < NSPT_PROF_N; i++) {
    c->prof_ms[i] = 0.0;
    c->prof_cnt[i] = 0;
  }
  c->prof = enable != 0;
  return 0;
}

int nspt_profile_read(nspt_ctx* c, double* ms_total, long long* count) {
  if (!c || !ms_total || !count) return -1;
  prof_collect(c);
  for (int i = 0; i < NSPT_PROF_N; i++) {
    ms_total[i] = c->prof_ms[i];
    count[i] = c->prof_cnt[i];
  }
  return 0;
}

}  // extern "C"

// ================================================================== batches
// Independent replicas stepped together.  A step of every replica is
// captured once into a HIP graph -- one branch per replica, on the
// replica's own stream, so the branches run concurrently -- and replayed.
// Two graphs are kept because a step swaps the roles of the U and F
// buffers: graph 0 is captured for the current buffer roles, graph 1 for the
// swapped ones; after each replay the host-side roles are swapped to match.
struct nspt_batch {
  std::vector<nspt_ctx*> reps;
  std::vector<nspt_step_params> params;
  hipStream_t master = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> join;
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  int parity = 0;
  bool with_obs = false;
  int cap = 0;  // observable rows per replica held on the device
  std::string err;
  int fail(const std::string& m) {
    err = m;
    return -1;
  }
};

namespace {

void batch_drop_graphs(nspt_batch* b) {
  for (auto& e : b->exec) {
    if (e) (void)hipGraphExecDestroy(e);
    e = nullptr;
  }
}

// Capture the two step graphs (with or without observables).
bool batch_build(nspt_batch* b, bool with_obs) {
  batch_drop_graphs(b);
  for (auto* c : b->reps) {
    c->graph = true;
    c->prof = false;
  }
  bool ok = true;
  for (int g = 0; g < 2 && ok; g++) {
    if (hipStreamBeginCapture(b->master, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      b->err = "hipStreamBeginCapture failed";
      ok = false;
      break;
    }
    (void)hipEventRecord(b->fork, b->master);
    for (size_t r = 0; r < b->reps.size() && ok; r++) {
      nspt_ctx* c = b->reps[r];
      (void)hipStreamWaitEvent(c->stream, b->fork, 0);
      if (!do_step(c, &b->params[r], with_obs ? c->obs_dev : nullptr)) {
        b->err = "replica " + std::to_string(r) + ": " + c->err;
        ok = false;
      }
      hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, c->stream, c->ctr_dev);
      (void)hipEventRecord(b->join[r], c->stream);
      (void)hipStreamWaitEvent(b->master, b->join[r], 0);
    }
    hipGraph_t graph = nullptr;
    if (hipStreamEndCapture(b->master, &graph) != hipSuccess || !graph) {
      if (ok) b->err = "hipStreamEndCapture failed";
      ok = false;
    }
    if (ok && hipGraphInstantiate(&b->exec[g], graph, nullptr, nullptr, 0) != hipSuccess) {
      b->err = "hipGraphInstantiate failed";
      ok = false;
    }
    if (graph) (void)hipGraphDestroy(graph);
  }
  // each capture performed one U/F swap per replica: two captures restore
  // the roles that graph 0 expects
  if (!ok) {
    batch_drop_graphs(b);
  }
  for (auto* c : b->reps) c->graph = false;
  b->parity = 0;
  b->with_obs = with_obs;
  return ok;
}

bool batch_sync(nspt_batch* b) {
  if (hipStreamSynchronize(b->master) != hipSuccess) {
    b->err = "hipStreamSynchronize failed";
    return false;
  }
  return true;
}

// n replays; the host-side U/F roles follow the device state
bool batch_replay(nspt_batch* b, int n) {
  for (int s = 0; s < n; s++) {
    if (hipGraphLaunch(b->exec[b->parity], b->master) != hipSuccess) {
      b->err = "hipGraphLaunch failed";
      return false;
    }
    for (auto* c : b->reps) std::swap(c->U, c->F);
    b->parity ^= 1;
  }
  return true;
}

bool batch_set_counters(nspt_batch* b, unsigned long long step, unsigned long long row) {
  for (auto* c : b->reps) {
    const unsigned long long v[2] = {step, row};
    if (!c->hip_ok(hipMemcpyAsync(c->ctr_dev, v, sizeof(v), hipMemcpyHostToDevice, b->master), "counters")) {
      b->err = c->err;
      return false;
    }
  }
  return batch_sync(b);
}

}  // namespace

extern "C" {

int nspt_batch_create(const nspt_geom* geoms, const nspt_step_params* params, int nrep, nspt_batch** out) {
  if (!geoms || !params || !out || nrep < 1) return -1;
  nspt_batch* b = new nspt_batch();
  *out = b;
  for (int r = 0; r < nrep; r++) {
    if (geoms[r].nc != geoms[0].nc || geoms[r].ptord != geoms[0].ptord)
      return b->fail("nspt_batch_create: replicas must share nc and ptord");
    if (params[r].noise != NSPT_NOISE_DEVICE)
      return b->fail("nspt_batch_create: replicas draw device noise (NSPT_NOISE_DEVICE)");
    if (params[r].integrator != NSPT_EULER && params[r].integrator != NSPT_RK2)
      return b->fail("nspt_batch_create: unknown integrator");
    nspt_ctx* c = nullptr;
    const int rc = nspt_create(&geoms[r], &c);
    if (rc != 0) {
      b->err = std::string("nspt_batch_create: replica ") + std::to_string(r) + ": " + (c ? c->err : "create failed");
      if (c) nspt_destroy(c);
      return -2;
    }
    b->reps.push_back(c);
    b->params.push_back(params[r]);
  }
  if (hipStreamCreate(&b->master) != hipSuccess || hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    return b->fail("nspt_batch_create: stream/event creation failed");
  b->join.resize(nrep);
  for (auto& e : b->join)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return b->fail("event creation failed");
  return batch_set_counters(b, 0, 0) ? 0 : -3;
}

int nspt_batch_destroy(nspt_batch* b) {
  if (!b) return -1;
  batch_drop_graphs(b);
  for (auto* c : b->reps) nspt_destroy(c);
  for (auto e : b->join) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->master) (void)hipStreamDestroy(b->master);
  delete b;
  return 0;
}

const char* nspt_batch_last_error(const nspt_batch* b) { return b ? b->err.c_str() : "null batch"; }

int nspt_batch_replica(nspt_batch* b, int r, nspt_ctx** ctx) {
  if (!b || !ctx || r < 0 || r >= (int)b->reps.size()) return -1;
  *ctx = b->reps[r];
  return 0;
}

int nspt_batch_evolve(nspt_batch* b, int nsteps, double* obs) {
  if (!b || nsteps < 0) return -1;
  const bool want = obs != nullptr;
  const int chunk = 256;
  if (want && b->cap < chunk) {
    for (auto* c : b->reps)
      if (!ensure_obs(c, chunk)) return b->fail(c->err);
    b->cap = chunk;
    batch_drop_graphs(b);
  }
  if (!b->exec[0] || b->with_obs != want) {
    if (!batch_build(b, want)) return -3;
  }
  // replica r's rows: obs + r * nsteps * obs_len(r)
  std::vector<size_t> base(b->reps.size(), 0);
  for (size_t r = 1; r < b->reps.size(); r++) base[r] = base[r - 1] + (size_t)nsteps * b->reps[r - 1]->obs_len;
  int done = 0;
  while (done < nsteps) {
    const int n = want ? std::min(chunk, nsteps - done) : nsteps;
    if (want) {
      // the row counter restarts for each chunk; the noise step counter runs on
      for (auto* c : b->reps) {
        const unsigned long long zero = 0;
        if (hipMemcpyAsync(c->ctr_dev + 1, &zero, sizeof(zero), hipMemcpyHostToDevice, b->master) != hipSuccess)
          return b->fail("counter reset failed");
      }
    }
    if (!batch_replay(b, n)) return -3;
    if (want) {
      for (size_t r = 0; r < b->reps.size(); r++) {
        nspt_ctx* c = b->reps[r];
        if (hipMemcpyAsync(obs + base[r] + (size_t)done * c->obs_len, c->obs_dev, sizeof(double) * n * c->obs_len,
                           hipMemcpyDeviceToHost, b->master) != hipSuccess)
          return b->fail("observable copy failed");
      }
    }
    if (!batch_sync(b)) return -3;
    done += n;
  }
  return 0;
}

int nspt_batch_tune(nspt_batch* b, int steps, int* qt_chosen, double* ms_per_step) {
  if (!b || steps < 1) return -1;
  nspt_ctx* c0 = b->reps[0];
  // candidate tile edges compiled for the replicas' instance
  std::vector<int> cand;
  for (int q : {4, 3, 5, 2}) {
    const KernelSet ks = find_set(c0->nc, c0->nrt, q);
    if (ks.qt == q && ks.nmax == c0->ks.nmax) cand.push_back(q);
  }
  // snapshot fields and counters: tuning steps are not part of the run
  std::vector<double2*> snap(b->reps.size(), nullptr);
  std::vector<unsigned long long> ctr(2 * b->reps.size());
  for (size_t r = 0; r < b->reps.size(); r++) {
    nspt_ctx* c = b->reps[r];
    const size_t n = (size_t)c->nrt * c->NE * 4 * c->V + 4 * (size_t)c->V;
    if (hipMalloc(&snap[r], n * sizeof(double2)) != hipSuccess) return b->fail("tune: snapshot allocation failed");
    (void)hipMemcpyAsync(snap[r], c->U.data, (n - 4 * (size_t)c->V) * sizeof(double2), hipMemcpyDeviceToDevice,
                         b->master);
    (void)hipMemcpyAsync(snap[r] + (n - 4 * (size_t)c->V), c->U.flag, 4 * (size_t)c->V * sizeof(double2),
                         hipMemcpyDeviceToDevice, b->master);
    (void)hipMemcpyAsync(&ctr[2 * r], c->ctr_dev, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, b->master);
  }
  if (!batch_sync(b)) return -3;
  double best = 1e30;
  int best_q = c0->ks.qt;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  for (size_t i = 0; i < cand.size(); i++) {
    for (auto* c : b->reps) c->ks = find_set(c->nc, c->nrt, cand[i]);
    if (!batch_build(b, false) || !batch_replay(b, 1)) return -3;
    (void)hipEventRecord(e0, b->master);
    if (!batch_replay(b, steps)) return -3;
    (void)hipEventRecord(e1, b->master);
    if (!batch_sync(b)) return -3;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double per = ms / steps;
    if (ms_per_step && i < 4) ms_per_step[i] = per;
    if (per < best) {
      best = per;
      best_q = cand[i];
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  for (auto* c : b->reps) c->ks = find_set(c->nc, c->nrt, best_q);
  batch_drop_graphs(b);
  for (size_t r = 0; r < b->reps.size(); r++) {
    nspt_ctx* c = b->reps[r];
    const size_t n = (size_t)c->nrt * c->NE * 4 * c->V;
    (void)hipMemcpyAsync(c->U.data, snap[r], n * sizeof(double2), hipMemcpyDeviceToDevice, b->master);
    (void)hipMemcpyAsync(c->U.flag, snap[r] + n, 4 * (size_t)c->V * sizeof(double2), hipMemcpyDeviceToDevice,
                         b->master);
    (void)hipMemcpyAsync(c->ctr_dev, &ctr[2 * r], 2 * sizeof(unsigned long long), hipMemcpyHostToDevice, b->master);
  }
  if (!batch_sync(b)) return -3;
  for (auto* p : snap) (void)hipFree(p);
  if (qt_chosen) *qt_chosen = best_q;
  return 0;
}

}  // extern "C"
