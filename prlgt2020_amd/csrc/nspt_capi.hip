// nspt_capi.hip -- host side of libnspt.so: context, device buffers, the
// order-independent kernels (layout conversion, noise, reductions, the 4-D
// DFT of the Fourier-accelerated gauge fixing) and the C ABI of
// include/nspt.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nspt.h"
#include "nspt_device.hpp"
#include "nspt_internal.hpp"
#include "nspt_table.hpp"

using nspt::DriftArgs;
using nspt::DriftPArgs;
using nspt::GexpArgs;
using nspt::Geo;
using nspt::KernelSet;
using nspt::PlaqArgs;
using nspt::TransformArgs;
using nspt::UpdateArgs;
using nspt::ZeroModeArgs;
using nspt::SerField;
using nspt::TwistTab;

namespace {

// ------------------------------------------------------------------ kernels

// Host image (ptGluon[V], width A) -> device field (orders < nrt, tiled
// series as in nspt::ser_index).
__global__ void k_aos_to_soa(const double2* aos, int A, int NE, int nrt, int V, SerField U) {
  const long P = 4L * V;
  const long n = (long)nrt * NE * P;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const long idx = t % P;
    const long ke = t / P;  // k*NE + e
    const long site = idx % V, mu = idx / V;
    U.data[nspt::ser_index(idx, (int)ke, nrt, NE)] = aos[(4 * site + mu) * (1 + (long)NE * A) + 1 + ke];
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
      v = U.data[nspt::ser_index(idx, (int)(j - 1), nrt, NE)];
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
__device__ __forceinline__ void noise_philox_body(int nc, int V, unsigned long long seed, unsigned long long step0,
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

__global__ void k_noise_philox(int nc, int V, unsigned long long seed, unsigned long long step0,
                               const unsigned long long* step_dev, double2* eta) {
  noise_philox_body(nc, V, seed, step0, step_dev, eta);
}

// One replica per blockIdx.y (the grid-stride loop covers each replica's
// links with the common x extent).
struct NoiseArgs {
  int nc, V;
  unsigned long long seed, step0;
  const unsigned long long* step_dev;
  double2* eta;
};
__global__ void k_noise_philox_tab(const NoiseArgs* __restrict__ t) {
  const NoiseArgs& a = t[blockIdx.y];
  noise_philox_body(a.nc, a.V, a.seed, a.step0, a.step_dev, a.eta);
}

// Deterministic column sums, one level of a tree over the four directions
// at once (blockIdx.z = mu): rows [mu*in_mu + chunk*per, ...) of the rows
// `rows` of direction mu (row stride ld) are summed into row
// mu*out_mu + chunk of out (row stride out_ld), times scale.
__global__ void k_reduce_rows(const double* in, long ld, long in_mu, long rows, int ncols, long per, double* out,
                              long out_ld, long out_mu, double scale) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= ncols) return;
  const long mu = blockIdx.z;
  const long r0 = mu * in_mu + blockIdx.y * per;
  const long r1 = min(mu * in_mu + rows, r0 + per);
  double s = 0.0;
  for (long r = r0; r < r1; r++) s += in[r * ld + col];
  out[(mu * out_mu + blockIdx.y) * out_ld + col] = s * scale;
}

// Mean over the sites of every (direction, order, element) column of a
// link field: out[(mu*nrt + k)*NE + e] = (1/V) sum_x (order k, element e
// of series mu*V + x) (the layout of the zero-mode means MM).  One block per
// column; each thread sums a fixed stride of sites, then a fixed tree:
// deterministic.
constexpr int MEAN_THREADS = 256;
__global__ __launch_bounds__(MEAN_THREADS) void k_field_mean(const double2* f, int V, int nrt, int NE,
                                                             double2* out) {
  const int col = blockIdx.x;
  const int e = col % NE, k = (col / NE) % nrt, mu = col / (NE * nrt);
  const int ke = k * NE + e;
  double2 s = make_double2(0.0, 0.0);
  for (int x = threadIdx.x; x < V; x += MEAN_THREADS)
    s = nspt::c_add(s, f[nspt::ser_index((long)mu * V + x, ke, nrt, NE)]);
  __shared__ double2 red[MEAN_THREADS];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = MEAN_THREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = nspt::c_add(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[col] = nspt::c_scale(1.0 / (double)V, red[0]);
}

// k_reduce_rows for the replicas of a batch: blockIdx.z = 4 replica + mu;
// a replica whose tree has fewer levels, or fewer chunks at this level, has
// nch below blockIdx.y + 1 (0: inactive at this level).
struct ReduceArgs {
  const double* in;
  long ld, in_mu, rows;
  int ncols;
  long per;
  double* out;
  long out_ld, out_mu;
  double scale;
  int nch;
};
__global__ void k_reduce_rows_tab(const ReduceArgs* __restrict__ t) {
  const ReduceArgs& a = t[blockIdx.z >> 2];
  if ((int)blockIdx.y >= a.nch) return;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.ncols) return;
  const long mu = blockIdx.z & 3;
  const long r0 = mu * a.in_mu + blockIdx.y * a.per;
  const long r1 = min(mu * a.in_mu + a.rows, r0 + a.per);
  double s = 0.0;
  for (long r = r0; r < r1; r++) s += a.in[r * a.ld + col];
  a.out[(mu * a.out_mu + blockIdx.y) * a.out_ld + col] = s * a.scale;
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
// the counters of nrep replicas
__global__ void k_count_tab(unsigned long long* const* __restrict__ ctrs, int nrep) {
  for (int r = threadIdx.x; r < nrep; r += blockDim.x) {
    ctrs[r][0] += 1;
    ctrs[r][1] += 1;
  }
}

// slot_dev (may be null): a device counter selecting the output row, for
// steps replayed from a captured graph.
// obs2_mu (may be null): per-direction sums of the rectangle plaquette of an
// improved action, appended after trU (nspt.cc:575-578).
__device__ __forceinline__ void obs_finalize_body(const double* obs_mu, const double* obs2_mu, int nrt, int nc, int V,
                                                  double* out, const unsigned long long* slot_dev, int obs_len) {
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
  if (obs2_mu)
    for (int t = threadIdx.x; t < 2 * (nrt + 1); t += blockDim.x) {
      double s = 0.0;
      for (int mu = 0; mu < 4; mu++) s += obs2_mu[mu * 2 * (nrt + 1) + t];
      out[2 * (nrt + 1) + nrt + 8 * nrt + t] = s / pn;
    }
}
__global__ void k_obs_finalize(const double* obs_mu, const double* obs2_mu, int nrt, int nc, int V, double* out,
                               const unsigned long long* slot_dev, int obs_len) {
  obs_finalize_body(obs_mu, obs2_mu, nrt, nc, V, out, slot_dev, obs_len);
}
struct ObsFinArgs {
  const double* obs_mu;
  int nrt, nc, V;
  double* out;
  const unsigned long long* slot_dev;
  int obs_len;
};
// one block per replica
__global__ void k_obs_finalize_tab(const ObsFinArgs* __restrict__ t) {
  const ObsFinArgs& a = t[blockIdx.x];
  obs_finalize_body(a.obs_mu, nullptr, a.nrt, a.nc, a.V, a.out, a.slot_dev, a.obs_len);
}

// ------------------------------------------------------------ FA transform
// The unnormalised 4-D DFT of the divergence of X (FFTW's sign convention,
// Quenched.cc:827-838), the momentum-space factor and the inverse DFT
// (nspt.cc:871-902), in three passes over fb[comp*V + site]:
//   pass 0: W = sum_mu X_mu(x) - X_mu(x-mu), then forward FFT along 2 and 3;
//   pass 1: forward FFT along 0 and 1, times -alpha/(V p^2(k)) with
//           p^2 = sum_mu 4 sin^2(pi k_mu / L_mu) and the k = 0 mode zeroed,
//           then backward FFT along 1 and 0;
//   pass 2: backward FFT along 2 and 3.
// The momentum norms of the whole field (nspt_momentum_norm) take two more
// passes over the planes of one link direction:
//   pass 3: the plane of the field itself (Xsrc + comp 4V), forward FFT
//           along 2 and 3;
//   pass 4: forward FFT along 0 and 1.
// A block (one wave) holds one tile of L_da x L_db x C sites in LDS, C
// consecutive values of a third direction: direction 3 in passes 1 and 4
// (sites (x0, x1) x C consecutive x3), direction 1 in passes 0, 2 and 3 (C
// consecutive (x2, x3) planes), so every load and store is a run of consecutive sites;
// C makes about 64 lines per axis, one per lane.  Each line is transformed by
// one lane in registers: a mixed-radix Stockham FFT unrolled at compile time
// for the extents in NSPT_FFT_SIZES, a direct DFT read from LDS for any
// other extent.
constexpr int FFT_THREADS = 64;
constexpr int FFT_MAXE = 2048;  // tile sites: L_da * L_db * C
constexpr int FFT_MAXN = 32;    // lattice extents are at most 32
constexpr int FFT_LOAD_BATCH = 8;  // tile loads in flight per thread

#define NSPT_FFT_SIZES(X) X(2) X(3) X(4) X(5) X(6) X(8) X(10) X(12) X(16) X(20) X(24) X(32)

__host__ __device__ constexpr int smallest_factor(int n) {
  int r = 2;
  while (n % r) r++;
  return r;
}

__host__ __device__ constexpr int fft_nstages(int n) {
  int s = 0;
  for (int len = n; len > 1; s++) len /= smallest_factor(len);
  return s;
}

__host__ inline bool fft_unrolled(int n, int maxn) {
  bool hit = false;
#define NSPT_FFT_HIT(N) hit = hit || (n == N && N <= maxn);
  NSPT_FFT_SIZES(NSPT_FFT_HIT)
#undef NSPT_FFT_HIT
  return hit;
}

// acc + v w^k with w = exp(-2 pi i / n) (FWD) or its conjugate; w holds the
// forward twiddles.  k is known at compile time once unrolled, so the
// trivial twiddles 1, -1 and -+i cost no multiplication.
template <int n, bool FWD>
__device__ __forceinline__ double2 tw_mac(double2 acc, double2 v, const double2* w, int k) {
  if (k == 0) return nspt::c_add(acc, v);
  if (2 * k == n) return nspt::c_sub(acc, v);
  if (4 * k == n || 4 * k == 3 * n) {
    // w^k = -i (forward, k = n/4) or +i; the sign flips for 3n/4 and backward
    const bool minus_i = (4 * k == n) == FWD;
    return minus_i ? make_double2(acc.x + v.y, acc.y - v.x) : make_double2(acc.x - v.y, acc.y + v.x);
  }
  const double tx = w[k].x, ty = FWD ? w[k].y : -w[k].y;
  return make_double2(fma(v.x, tx, fma(-v.y, ty, acc.x)), fma(v.x, ty, fma(v.y, tx, acc.y)));
}

// One Stockham stage (length len = r*m, stride st) from x into y: output
// q + st*(r*p + t) is the r-point DFT of x[q + st*(p + u*m)], u = 0..r-1,
// times w_len^{p t}; then the remaining stages from y into x.
template <int n, int len, int st, bool FWD>
__device__ __forceinline__ void fft_stages(double2 (&x)[n], double2 (&y)[n], const double2* w) {
  if constexpr (len > 1) {
    constexpr int r = smallest_factor(len), m = len / r;
#pragma unroll
    for (int q = 0; q < st; q++)
#pragma unroll
      for (int p = 0; p < m; p++)
#pragma unroll
        for (int t = 0; t < r; t++) {
          double2 acc = x[q + st * p];
#pragma unroll
          for (int u = 1; u < r; u++) acc = tw_mac<n, FWD>(acc, x[q + st * (p + u * m)], w, (u * t * (n / r)) % n);
          y[q + st * (r * p + t)] = tw_mac<n, FWD>(make_double2(0.0, 0.0), acc, w, (p * t * (n / len)) % n);
        }
    fft_stages<n, m, st * r, FWD>(y, x, w);
  }
}

// w: the n forward twiddles in LDS (every lane reads the same entry)
template <int n, bool FWD>
__device__ __forceinline__ void fft_line(double2 (&x)[n], const double2* w) {
  double2 y[n];
  fft_stages<n, n, 1, FWD>(x, y, w);
  if constexpr (fft_nstages(n) % 2 == 1) {
#pragma unroll
    for (int i = 0; i < n; i++) x[i] = y[i];
  }
}

// Per-block tables in LDS: forward twiddles of both axes, 4 sin^2(pi k/L) of
// every direction, the tile and (direct DFT only) a second tile buffer.
struct FftShared {
  double2* twa;  // exp(-2 pi i k / L_da)
  double2* twb;
  double* s2[4];
  double2* tile;
  double2* tmp;
};

// Line transform along one tile axis: nl = nh * C lines, line t = (hi, ic)
// (ic = t % C when c_inner, else hi = t % nh) starting at LDS offset
// hi * pitch_hi + ic * pitch_c, element stride ls.  mode 0 forward, 1
// backward, 2 forward, the momentum factor, backward (axis b of pass 1,
// where hi = k0 and the element index is k1).
struct FftAxis {
  int n, nh, C, ls, pitch_hi, pitch_c;
  bool c_inner;
};

__device__ __forceinline__ void line_of(const FftAxis& ax, int t, int& hi, int& ic) {
  if (ax.c_inner) {
    hi = t / ax.C;
    ic = t % ax.C;
  } else {
    hi = t % ax.nh;
    ic = t / ax.nh;
  }
}

struct FftScale {
  double alpha, invV;
  int k2, k3base;  // momentum along 2 and the first momentum along 3 of the tile (pass 1)
};

__device__ __forceinline__ double fa_factor(const FftShared& sh, const FftScale& sc, int k0, int k1, int k3off) {
  const int k3 = sc.k3base + k3off;
  const double p2 = sh.s2[0][k0] + sh.s2[1][k1] + sh.s2[2][sc.k2] + sh.s2[3][k3];
  return (k0 == 0 && k1 == 0 && sc.k2 == 0 && k3 == 0) ? 0.0 : -sc.alpha * sc.invV / p2;
}

template <int n>
__device__ void axis_fft_fixed(const FftShared& sh, const double2* tw, const FftAxis& ax, int mode,
                               const FftScale& sc) {
  for (int t = threadIdx.x; t < ax.nh * ax.C; t += blockDim.x) {
    int hi, ic;
    line_of(ax, t, hi, ic);
    double2* base = sh.tile + hi * ax.pitch_hi + ic * ax.pitch_c;
    double2 x[n];
#pragma unroll
    for (int j = 0; j < n; j++) x[j] = base[j * ax.ls];
    if (mode == 1)
      fft_line<n, false>(x, tw);
    else
      fft_line<n, true>(x, tw);
    if (mode == 2) {
#pragma unroll
      for (int j = 0; j < n; j++) x[j] = nspt::c_scale(fa_factor(sh, sc, hi, j, ic), x[j]);
      fft_line<n, false>(x, tw);
    }
#pragma unroll
    for (int j = 0; j < n; j++) base[j * ax.ls] = x[j];
  }
}

// Direct DFT of every line for extents without an unrolled FFT: each output
// read from the tile, written to the second buffer, copied back.
__device__ void axis_dft_direct(const FftShared& sh, const double2* tw, const FftAxis& ax, int mode,
                                const FftScale& sc) {
  const int n = ax.n;
  const int passes = mode == 2 ? 2 : 1;
  for (int ps = 0; ps < passes; ps++) {
    const double cj = (mode == 1 || ps == 1) ? -1.0 : 1.0;
    for (int o = threadIdx.x; o < ax.nh * ax.C * n; o += blockDim.x) {
      int hi, ic;
      line_of(ax, o / n, hi, ic);
      const int k = o % n;
      const int off = hi * ax.pitch_hi + ic * ax.pitch_c;
      double2 acc = make_double2(0.0, 0.0);
      int wi = 0;
      for (int j = 0; j < n; j++) {
        const double2 v = sh.tile[off + j * ax.ls];
        const double2 w = tw[wi];
        const double wy = cj * w.y;
        acc.x = fma(v.x, w.x, fma(-v.y, wy, acc.x));
        acc.y = fma(v.x, wy, fma(v.y, w.x, acc.y));
        wi += k;
        wi -= (wi >= n) ? n : 0;
      }
      if (mode == 2 && ps == 0) acc = nspt::c_scale(fa_factor(sh, sc, hi, k, ic), acc);
      sh.tmp[off + k * ax.ls] = acc;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < ax.nh * ax.C * n; o += blockDim.x) {
      int hi, ic;
      line_of(ax, o / n, hi, ic);
      const int off = hi * ax.pitch_hi + ic * ax.pitch_c + (o % n) * ax.ls;
      sh.tile[off] = sh.tmp[off];
    }
    __syncthreads();
  }
}

// MAXN bounds the unrolled sizes of one kernel variant, and so its registers.
template <int MAXN>
__device__ void axis_transform(const FftShared& sh, const double2* tw, const FftAxis& ax, int mode,
                               const FftScale& sc) {
  switch (ax.n) {
#define NSPT_FFT_CASE(N)                                  \
  case N:                                                 \
    if constexpr (N <= MAXN) {                            \
      axis_fft_fixed<N>(sh, tw, ax, mode, sc);            \
      break;                                              \
    }                                                     \
    [[fallthrough]];
    NSPT_FFT_SIZES(NSPT_FFT_CASE)
#undef NSPT_FFT_CASE
    default:
      axis_dft_direct(sh, tw, ax, mode, sc);
  }
  __syncthreads();
}

// Tile geometry of one pass.  Pass 1: LDS offset ia*ra + ib*C + ic with
// ra = L_1*C + 1; passes 0, 2: ic*cs + ia*ra + ib with ra = L_3 + 1,
// cs = L_2*ra.  The extra site per row spreads the rows over the LDS banks.
struct FftTile {
  int da, db, na, nb, C, E, tiles, ra, cs, words;
};

// C: the largest divisor of the batch direction's extent that keeps the
// tile within FFT_MAXE sites and the lines per axis within one per lane.
__host__ __device__ inline bool fft_axes01(int pass) { return pass == 1 || pass == 4; }
__host__ __device__ inline FftTile fft_tile(const Geo& g, int pass) {
  FftTile t;
  const bool p01 = fft_axes01(pass);
  t.da = p01 ? 0 : 2;
  t.db = p01 ? 1 : 3;
  t.na = g.L[t.da];
  t.nb = g.L[t.db];
  const int Lc = g.L[p01 ? 3 : 1];
  const int mn = t.na < t.nb ? t.na : t.nb;
  t.C = 1;
  for (int c = 2; c <= Lc; c++)
    if (Lc % c == 0 && t.na * t.nb * c <= FFT_MAXE && mn * c <= FFT_THREADS) t.C = c;
  t.E = t.na * t.nb * t.C;
  t.tiles = g.V / t.E;
  if (p01) {
    t.ra = t.nb * t.C + 1;
    t.cs = 1;
    t.words = t.na * t.ra;  // the last offset is na*ra - 2
  } else {
    t.ra = t.nb + 1;
    t.cs = t.na * t.ra;
    t.words = t.C * t.cs;  // the last offset is C*na*ra - 2
  }
  return t;
}

__host__ __device__ inline int fft_tile_words(const FftTile& t) {  // double2 per tile buffer
  return t.words;
}

__host__ inline size_t fft_shm(const FftTile& t, const Geo& g, bool direct) {
  const int nsin = (g.L[0] + g.L[1] + g.L[2] + g.L[3] + 1) / 2;  // in double2 units
  return sizeof(double2) * ((size_t)t.na + t.nb + nsin + (direct ? 2 : 1) * (size_t)fft_tile_words(t));
}

template <int MAXN>
__device__ __forceinline__ void fa_fft_body(int blk, const Geo& g, int pass, const double2* Xsrc, int ne, int xmu,
                                            double2* fb, double alpha) {
  extern __shared__ double2 sm[];
  const FftTile t = fft_tile(g, pass);
  const bool p01 = fft_axes01(pass);
  const long V = g.V;
  const long comp = blk / t.tiles;
  const int tt = (int)(blk % t.tiles);
  const int L0 = g.L[0], L1 = g.L[1], L2 = g.L[2], L3 = g.L[3];
  // tile origin: (x2, first x3) in passes 1 and 4, (x0, first x1) otherwise
  const int nch = (p01 ? L3 : L1) / t.C;
  const int o0 = tt / nch;
  const int o1 = (tt % nch) * t.C;
  const long base = p01 ? (long)o0 * g.st[2] + o1 : (long)o0 * g.st[0] + (long)o1 * g.st[1];
  FftShared sh;
  sh.twa = sm;
  sh.twb = sm + t.na;
  double* s2 = reinterpret_cast<double*>(sh.twb + t.nb);
  sh.s2[0] = s2;
  sh.s2[1] = s2 + L0;
  sh.s2[2] = s2 + L0 + L1;
  sh.s2[3] = s2 + L0 + L1 + L2;
  sh.tile = sh.twb + t.nb + (L0 + L1 + L2 + L3 + 1) / 2;
  sh.tmp = sh.tile + fft_tile_words(t);
  for (int k = threadIdx.x; k < t.na + t.nb; k += blockDim.x) {
    const bool on_a = k < t.na;
    const int n = on_a ? t.na : t.nb;
    const int i = on_a ? k : k - t.na;
    double sn, cs;
    sincospi(2.0 * (double)i / (double)n, &sn, &cs);
    sh.twa[k] = make_double2(cs, -sn);  // twb follows twa
  }
  if (pass == 1) {
    // s2 holds 4 sin^2(pi i / L_mu) for mu = 0..3 back to back
    for (int k = threadIdx.x; k < L0 + L1 + L2 + L3; k += blockDim.x) {
      int i = k, n = L0;
      if (i >= n) {
        i -= n;
        n = L1;
        if (i >= n) {
          i -= n;
          n = L2;
          if (i >= n) {
            i -= n;
            n = L3;
          }
        }
      }
      const double h = sinpi((double)i / (double)n);
      s2[k] = 4.0 * h * h;
    }
  }

  double2* line = fb + comp * V;
  // component comp (order and element) of link series idx of the source
  auto xc = [&](long idx) { return Xsrc[nspt::ser_index(idx, (int)comp, g.nrt, ne)]; };
  // FFT_LOAD_BATCH sites per thread are loaded before any is stored to LDS,
  // so that their global-load latencies overlap (the pass is a constant
  // inside, so no branch separates the loads)
  auto load_tile = [&](auto pass_c) {
    constexpr int PASS = decltype(pass_c)::value;
    // the divergence already has eight loads per site
    constexpr int NB = PASS == 0 ? 4 : FFT_LOAD_BATCH;
    for (int l0 = threadIdx.x; l0 < t.E; l0 += NB * FFT_THREADS) {
      double wx[NB], wy[NB];  // plain doubles: an array of double2 is kept in scratch
      int offs[NB];
#pragma unroll
      for (int u = 0; u < NB; u++) {
        // past the tile's end a lane reloads the tile's last site and stores
        // the same value there as the lane that owns it (no branch, so the
        // compiler keeps the loads together)
        const int l = min(l0 + u * FFT_THREADS, t.E - 1);
        int ia, ib, ic, site, off;
        if constexpr (PASS == 1 || PASS == 4) {
          ic = l % t.C;
          ib = (l / t.C) % t.nb;
          ia = l / (t.C * t.nb);
          site = (int)(base + (long)ia * g.st[0] + (long)ib * g.st[1] + ic);
          off = ia * t.ra + ib * t.C + ic;
        } else {
          ib = l % t.nb;
          ia = (l / t.nb) % t.na;
          ic = l / (t.na * t.nb);
          site = (int)(base + (long)ic * g.st[1] + (long)ia * g.st[2] + ib);
          off = ic * t.cs + ia * t.ra + ib;
        }
        offs[u] = off;
        if constexpr (PASS == 0) {
          // divergence at x = (o0, o1 + ic, ia, ib)
          const int x[4] = {o0, o1 + ic, ia, ib};
          double2 w = make_double2(0.0, 0.0);
#pragma unroll
          for (int mu = 0; mu < 4; mu++) {
            const int xm = (x[mu] == 0) ? site + (g.L[mu] - 1) * g.st[mu] : site - g.st[mu];
            w = nspt::c_add(w, nspt::c_sub(xc(mu * V + site), xc(mu * V + xm)));
          }
          wx[u] = w.x;
          wy[u] = w.y;
        } else {
          // pass 3 reads direction xmu of the source field itself
          const double2 w = PASS == 3 ? xc((long)xmu * V + site) : line[site];
          wx[u] = w.x;
          wy[u] = w.y;
        }
      }
#pragma unroll
      for (int u = 0; u < NB; u++) sh.tile[offs[u]] = make_double2(wx[u], wy[u]);
    }
  };
  if (pass == 0)
    load_tile(std::integral_constant<int, 0>());
  else if (pass == 1)
    load_tile(std::integral_constant<int, 1>());
  else if (pass == 2)
    load_tile(std::integral_constant<int, 2>());
  else if (pass == 3)
    load_tile(std::integral_constant<int, 3>());
  else
    load_tile(std::integral_constant<int, 4>());
  __syncthreads();
  FftAxis axa, axb;  // lines (ib, ic) along a, (ia, ic) along b
  if (p01) {
    axa = FftAxis{t.na, t.nb, t.C, t.ra, t.C, 1, true};
    axb = FftAxis{t.nb, t.na, t.C, t.C, t.ra, 1, true};
  } else {
    axa = FftAxis{t.na, t.nb, t.C, t.ra, 1, t.cs, false};
    axb = FftAxis{t.nb, t.na, t.C, 1, t.ra, t.cs, false};
  }
  const FftScale sc{alpha, 1.0 / (double)V, o0, o1};
  if (pass == 2) {
    axis_transform<MAXN>(sh, sh.twa, axa, 1, sc);
    axis_transform<MAXN>(sh, sh.twb, axb, 1, sc);
  } else {
    axis_transform<MAXN>(sh, sh.twa, axa, 0, sc);
    axis_transform<MAXN>(sh, sh.twb, axb, pass == 1 ? 2 : 0, sc);
    if (pass == 1) axis_transform<MAXN>(sh, sh.twa, axa, 1, sc);
  }
  for (int l = threadIdx.x; l < t.E; l += FFT_THREADS) {
    int site, off;
    if (p01) {
      const int ic = l % t.C, ib = (l / t.C) % t.nb, ia = l / (t.C * t.nb);
      site = (int)(base + (long)ia * g.st[0] + (long)ib * g.st[1] + ic);
      off = ia * t.ra + ib * t.C + ic;
    } else {
      const int ib = l % t.nb, ia = (l / t.nb) % t.na, ic = l / (t.na * t.nb);
      site = (int)(base + (long)ic * g.st[1] + (long)ia * g.st[2] + ib);
      off = ic * t.cs + ia * t.ra + ib;
    }
    line[site] = sh.tile[off];
  }
}
template <int MAXN>
__global__ __launch_bounds__(FFT_THREADS) void k_fa_fft(Geo g, int pass, const double2* Xsrc, int ne, int xmu,
                                                       double2* fb, double alpha) {
  fa_fft_body<MAXN>(blockIdx.x, g, pass, Xsrc, ne, xmu, fb, alpha);
}
struct FftArgs {
  Geo g;
  const double2* Xsrc;
  double2* fb;
  double alpha;
  int nblk;  // tiles x components of this replica
  int ne;    // elements per matrix of the source field
};
template <int MAXN>
__global__ __launch_bounds__(FFT_THREADS) void k_fa_fft_tab(int pass, const FftArgs* __restrict__ t) {
  const FftArgs& a = t[blockIdx.y];
  if ((int)blockIdx.x >= a.nblk) return;
  fa_fft_body<MAXN>(blockIdx.x, a.g, pass, a.Xsrc, a.ne, 0, a.fb, a.alpha);
}

// Momentum norms of one link direction (nspt.cc:925-934): out[k V + p] is
// set (first) or increased by sum_e |fb[(k NE + e) V + p]|^2, fb holding the
// 4-D transforms of the direction's planes; the directions are added in
// order 0..3, the elements of a matrix in row-major order.
__global__ void k_knorm_acc(const double2* __restrict__ fb, long V, int NE, int nrt, int first, double* out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)nrt * V) return;
  const long k = i / V, p = i - k * V;
  double s = 0.0;
  for (int e = 0; e < NE; e++) {
    const double2 v = fb[(k * NE + e) * V + p];
    s += v.x * v.x + v.y * v.y;
  }
  out[i] = first ? s : out[i] + s;
}

// Smallest compiled N >= n; among those the shape (tile edge qt, links per
// wave lpw, waves per SIMD wpe) if compiled, else the default shape (edge 4,
// lpw 0, one wave per SIMD).
KernelSet find_set(int nc, int n, int qt = 4, bool big = false, int lpw = 0, int wpe = 1) {
  int nmax = 0;
  auto order = [&](KernelSet s) {
    if (s.nc == nc && s.nmax >= n && s.big == big && (nmax == 0 || s.nmax < nmax)) nmax = s.nmax;
  };
  KernelSet want, dflt;
  auto pick = [&](KernelSet s) {
    if (s.nc != nc || s.nmax != nmax || s.big != big) return;
    if (s.qt == qt && s.lpw == lpw && s.wpe == wpe) want = s;
    if (s.qt == 4 && s.lpw == 0 && s.wpe == 1) dflt = s;
  };
#define NSPT_ORDER(a, b, q, l, w)                        \
  order(nspt::kernel_set_##a##_##b##_q##q##_l##l##_w##w()); \
  order(nspt::kernel_set_##a##_##b##_q##q##_l##l##_w##w##_big());
  NSPT_INSTANCES(NSPT_ORDER)
#undef NSPT_ORDER
  if (!nmax) return KernelSet();
#define NSPT_PICK(a, b, q, l, w)                        \
  pick(nspt::kernel_set_##a##_##b##_q##q##_l##l##_w##w()); \
  pick(nspt::kernel_set_##a##_##b##_q##q##_l##l##_w##w##_big());
  NSPT_INSTANCES(NSPT_PICK)
#undef NSPT_PICK
  return want.nmax ? want : dflt;
}

// Fewest links per wave among the compiled builds of (nc, nmax): the
// per-block partial-sum buffers are sized for it, so that the autotuner can
// move a context to any of them.
int min_links_per_wave(int nc, int nmax) {
  int m = 64;
  auto see = [&](KernelSet s) {
    if (s.nc == nc && s.nmax == nmax) m = std::min(m, s.lpb);
  };
#define NSPT_SEE(a, b, q, l, w)                        \
  see(nspt::kernel_set_##a##_##b##_q##q##_l##l##_w##w()); \
  see(nspt::kernel_set_##a##_##b##_q##q##_l##l##_w##w##_big());
  NSPT_INSTANCES(NSPT_SEE)
#undef NSPT_SEE
  return m;
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
  bool big = false;  // kernels built for fields past 2 GiB
  KernelSet ks;
  Geo g{};
  int nblk_links = 0, nblk_sites = 0, nob = 0, obs_len = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  SerField U{}, F{}, D{}, D1{}, L{}, X{}, G{};
  // Jacobi drift form: from the plaquette field Pf (6 per site, plane-major)
  // or from the staple sum as the reference forms it
  bool plaq_form = false;
  SerField Pf{};
  // Wilson-loop measurement, allocated on first use: 2-link and 4-link
  // fields (the 2-link field also feeds the rectangles of an improved
  // action), per-block rectangle sums, reduction scratch, the six sums
  SerField L2{}, L4{};
  double *loop_part = nullptr, *loop_tmp = nullptr, *loop_res = nullptr;
  // improved gauge action (nspt_set_action): rectangle coefficient, the
  // rectangle-staple sum of every link, per-block and per-direction sums of
  // the rectangle plaquette
  double c1 = 0.0;
  SerField R{};
  // twisted boundary conditions (nspt_set_twist): the twist matrices
  TwistTab* twist_tab = nullptr;
  double *obs2_part = nullptr, *obs2_mu = nullptr;
  double2* fb = nullptr;
  double* knorm_dev = nullptr;  // nspt_momentum_norm, allocated on first use
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
  // a replica of a batch: its step is replayed from captured graphs, so the
  // action and the boundary conditions are fixed at creation
  bool in_batch = false;
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
    if (!alloc(&f.data, (size_t)nspt::ser_alloc(P, nrt, NE))) return false;
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

// Deterministic reduction, for each of `groups` groups g, of the rows
// [g*rows, (g+1)*rows) of a (groups*rows x ncols) matrix (row stride ld)
// into out[g*ncols + col] (times scale): a tree of launches summing
// REDUCE_PER rows per thread, all groups in each launch.  tmp holds
// reduce_tmp_doubles(rows, groups, ncols) doubles.
constexpr long REDUCE_PER = 32;

size_t reduce_tmp_doubles(long rows, int groups, int ncols) {
  return 2 * (size_t)groups * ((rows + REDUCE_PER - 1) / REDUCE_PER) * ncols;
}

bool reduce_groups(nspt_ctx* c, const double* in, long ld, long rows, int groups, int ncols, double* out,
                   double scale, double* tmp) {
  Prof pr(c, NSPT_PROF_REDUCE);
  const size_t half = reduce_tmp_doubles(rows, groups, ncols) / 2;
  const double* src = in;
  long src_ld = ld;
  for (int level = 0;; level++) {
    const long nch = (rows + REDUCE_PER - 1) / REDUCE_PER;
    const dim3 grid((ncols + 63) / 64, (unsigned)nch, (unsigned)groups);
    if (nch == 1) {
      hipLaunchKernelGGL(k_reduce_rows, grid, dim3(64), 0, c->stream, src, src_ld, rows, rows, ncols, REDUCE_PER, out,
                         (long)ncols, 1L, scale);
      break;
    }
    double* dst = tmp + (level % 2) * half;
    hipLaunchKernelGGL(k_reduce_rows, grid, dim3(64), 0, c->stream, src, src_ld, rows, rows, ncols, REDUCE_PER, dst,
                       (long)ncols, nch, 1.0);
    src = dst;
    src_ld = ncols;
    rows = nch;
  }
  return launch_ok(c, "reduce");
}

// the four link directions of the per-block partial sums
bool reduce4(nspt_ctx* c, const double* in, long ld, int ncols, double* out, double scale) {
  return reduce_groups(c, in, ld, c->g.nb_mu, 4, ncols, out, scale, c->red_tmp);
}

// with_w2: the rectangle plaquette of an improved action is in obs2_part;
// rows: per-direction rows of obs_part (nb_mu, or 2 nb_half after a
// checkerboard sweep)
bool obs_reduce(nspt_ctx* c, double* out_dev, bool with_w2 = false, long rows = 0) {
  if (!reduce_groups(c, c->obs_part, c->nob, rows ? rows : c->g.nb_mu, 4, c->nob, c->obs_mu, 1.0, c->red_tmp))
    return false;
  if (with_w2 && !reduce4(c, c->obs2_part, 2 * (c->nrt + 1), 2 * (c->nrt + 1), c->obs2_mu, 1.0)) return false;
  Prof pr(c, NSPT_PROF_REDUCE);
  hipLaunchKernelGGL(k_obs_finalize, dim3(1), dim3(256), 0, c->stream, (const double*)c->obs_mu,
                     with_w2 ? (const double*)c->obs2_mu : nullptr, c->nrt, c->nc, c->V, out_dev,
                     c->graph ? (const unsigned long long*)(c->ctr_dev + 1) : nullptr, c->obs_len);
  return launch_ok(c, "obs_finalize");
}

// FA gauge fixing of U with Xsrc = log U up to a constant per direction.
// u_from_log: U is not in c->U but is exp(Xsrc - MM) (MM may be null),
// formed by the transform.
// One pass of k_fa_fft over the NE nrt planes of the FFT buffer (and of
// the link field Xsrc, read in pass 0, and in pass 3 for direction xmu).
bool fft_pass(nspt_ctx* c, int pass, const double2* Xsrc, double alpha, int xmu = 0) {
  const long ncomp = (long)c->nrt * c->NE;
  const FftTile t = fft_tile(c->g, pass);
  const int maxn = std::max(t.na, t.nb) <= 16 ? 16 : FFT_MAXN;
  const bool direct = !fft_unrolled(t.na, maxn) || !fft_unrolled(t.nb, maxn);
  const dim3 grid((unsigned)((long)t.tiles * ncomp));
  const size_t shm = fft_shm(t, c->g, direct);
  Prof pr(c, pass == 0 ? NSPT_PROF_FA_DIV : NSPT_PROF_DFT);
  if (maxn == 16)
    hipLaunchKernelGGL(k_fa_fft<16>, grid, dim3(FFT_THREADS), shm, c->stream, c->g, pass, Xsrc, c->NE, xmu, c->fb,
                       alpha);
  else
    hipLaunchKernelGGL(k_fa_fft<FFT_MAXN>, grid, dim3(FFT_THREADS), shm, c->stream, c->g, pass, Xsrc, c->NE, xmu,
                       c->fb, alpha);
  return launch_ok(c, pass == 0 ? "fa_div" : "fa_fft");
}

bool fa_apply(nspt_ctx* c, const SerField& Xsrc, double alpha, const double2* MM, bool u_from_log) {
  for (int pass = 0; pass < 3; pass++)
    if (!fft_pass(c, pass, (const double2*)Xsrc.data, alpha)) return false;
  {
    Prof pr(c, NSPT_PROF_FA_GEXP);
    c->ks.gexp(c->nblk_sites, c->stream, c->g, c->fb, c->G);
  }
  {
    Prof pr(c, NSPT_PROF_TRANSFORM);
    c->ks.transform(c->nblk_links, c->stream, c->g, c->G, c->U, u_from_log ? Xsrc : SerField{nullptr, nullptr, 0},
                    MM);
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

// Drift of every link of Uin (Jacobi passes), from its plaquette field or
// from the staple sums.
void drift_p(nspt_ctx* c, const SerField& Uin, const SerField& D, const SerField& D1, double tau_g, double rk2c,
             int mode, double* obs, const SerField& R) {
  if (!c->plaq_form) {
    Prof pr(c, NSPT_PROF_DRIFT);
    c->ks.drift(c->nblk_links, c->stream, c->g, Uin, D, D1, tau_g, rk2c, mode, obs, R, c->c1);
    return;
  }
  {
    Prof pr(c, NSPT_PROF_PLAQ_FIELD);
    c->ks.plaq_field(6 * c->g.nb_mu, c->stream, c->g, Uin, c->Pf);
  }
  Prof pr(c, NSPT_PROF_DRIFT);
  c->ks.drift_p(c->nblk_links, c->stream, c->g, Uin, c->Pf, D, D1, tau_g, rk2c, mode, obs, R, c->c1);
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
  // an improved action first forms the rectangle staples of the field the
  // drift reads (and, for the observables, their plaquette)
  const bool impr = c->c1 != 0.0;
  const SerField R = impr ? c->R : none;
  auto rect = [&](const SerField& Uin, bool with_obs) {
    if (!impr) return;
    Prof pr(c, NSPT_PROF_RECT);
    c->ks.link_pair(c->nblk_links, c->stream, c->g, Uin, c->L2, 1);
    c->ks.rect(c->nblk_links, c->stream, c->g, Uin, c->L2, c->R, with_obs && obs_part ? c->obs2_part : nullptr);
  };
  const bool cb = p->integrator == NSPT_EULER_CHECKERBOARD;
  if (cb) {
    // on the fly in checkerboard order (the reference's in-place sweep of
    // nspt.cc:365-441 reordered): for parity 0 then 1, for mu = 0..3, every
    // link (x, mu) of that parity at once -- their staples use no link of
    // the same direction and parity -- drift, then the update in place
    Geo gs = c->g;
    for (int par = 0; par < 2; par++)
      for (int mu = 0; mu < 4; mu++) {
        gs.sub = 1 + 4 * par + mu;
        {
          Prof pr(c, NSPT_PROF_DRIFT);
          c->ks.drift(c->g.nb_half, c->stream, gs, c->U, c->D, none, tau_g, 0.0, 0, obs_part, none, 0.0);
        }
        Prof pr(c, need_log ? NSPT_PROF_UPDATE_LOG : NSPT_PROF_UPDATE);
        c->ks.update(c->g.nb_half, c->stream, gs, c->D, eta, stau, c->U, c->U, need_log ? c->L : none,
                     p->zero_modes ? c->mm_part : nullptr);
      }
    if (!launch_ok(c, "checkerboard sweep")) return false;
    if (obs_dev && !obs_reduce(c, obs_dev, false, 2L * c->g.nb_half)) return false;
  } else if (p->integrator == NSPT_RK2) {
    rect(c->U, true);
    drift_p(c, c->U, c->D1, none, tau_g, 0.0, 0, obs_part, R);
    {
      Prof pr(c, NSPT_PROF_UPDATE);
      c->ks.update(c->nblk_links, c->stream, c->g, c->D1, eta, stau, c->U, c->F, none, nullptr);
    }
    rect(c->F, false);
    drift_p(c, c->F, c->D, c->D1, tau_g, rk2c, 1, nullptr, R);
  } else {
    rect(c->U, true);
    drift_p(c, c->U, c->D, none, tau_g, 0.0, 0, obs_part, R);
  }
  if (!cb) {
    if (!launch_ok(c, "drift")) return false;
    if (obs_dev && !obs_reduce(c, obs_dev, impr)) return false;
    {
      Prof pr(c, need_log ? NSPT_PROF_UPDATE_LOG : NSPT_PROF_UPDATE);
      // with zero modes the updated links are re-formed from exp(L - MM) by
      // the zero-mode kernel or the FA gauge transform (series and flags):
      // the update writes only L
      c->ks.update(c->nblk_links, c->stream, c->g, c->D, eta, stau, c->U, p->zero_modes ? none : c->F,
                   need_log ? c->L : none, p->zero_modes ? c->mm_part : nullptr);
    }
    if (!launch_ok(c, "update")) return false;
    std::swap(c->U, c->F);  // Umu = Fmu (nspt.cc:992)
  }
  const SerField* xsrc = &c->L;
  if (p->zero_modes) {
    const int ncol = 2 * c->nrt * c->NE;
    if (!reduce_groups(c, reinterpret_cast<const double*>(c->mm_part), ncol, cb ? 2L * c->g.nb_half : c->g.nb_mu,
                       4, ncol, reinterpret_cast<double*>(c->MM), 1.0 / (double)c->V, c->red_tmp))
      return false;
    if (p->gauge_fix != NSPT_GF_FA) {
      Prof pr(c, NSPT_PROF_ZERO_MODE);
      c->ks.zero_mode(c->nblk_links, c->stream, c->g, c->L, c->MM, c->U, c->X);
      if (!launch_ok(c, "zero_mode")) return false;
    }
    // with FA the divergence of L - MM is that of L, and the gauge
    // transform forms U = exp(L - MM) itself
  }
  if (p->gauge_fix == NSPT_GF_FA &&
      !fa_apply(c, *xsrc, p->alpha, p->zero_modes ? reinterpret_cast<const double2*>(c->MM) : nullptr,
                p->zero_modes != 0))
    return false;
  if (p->gauge_fix == NSPT_GF_CHECKERBOARD) {
    // the sequential gauge fixing of nspt.cc:660-862 in checkerboard site order
    Geo gs = c->g;
    for (int par = 0; par < 2; par++) {
      gs.sub = 1 + 4 * par;
      Prof pr(c, NSPT_PROF_TRANSFORM);
      c->ks.sgf(c->g.nb_half, c->stream, gs, c->U, p->alpha);
    }
    if (!launch_ok(c, "checkerboard gauge fixing")) return false;
  }
  return true;
}

bool ensure_obs(nspt_ctx* c, int n) {
  if (c->obs_cap >= n) return true;
  if (!c->alloc(&c->obs_dev, (size_t)n * c->obs_len)) return false;
  c->obs_cap = n;
  return true;
}

// Moves a context to kernel build ks: the link and site blocks follow its
// links per wave (the partial-sum buffers are sized for the fewest, see
// min_links_per_wave).
void set_kernels(nspt_ctx* c, const KernelSet& ks) {
  c->ks = ks;
  c->g.nb_mu = (c->V + ks.lpb - 1) / ks.lpb;
  c->g.nb_half = (c->V / 2 + ks.lpb - 1) / ks.lpb;
  c->nblk_links = 4 * c->g.nb_mu;
  c->nblk_sites = c->g.nb_mu;
}

// A kernel-build choice from the environment (NSPT_TILE_EDGE,
// NSPT_LINKS_PER_WAVE, NSPT_WAVES_PER_SIMD; A/B runs and tests), else dflt.
int env_choice(const char* name, int dflt) {
  const char* v = std::getenv(name);
  return (v && *v) ? std::atoi(v) : dflt;
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
  // the default build of the order, or the shape the environment names if
  // it is compiled
  const int qt = env_choice("NSPT_TILE_EDGE", 4), lpw = env_choice("NSPT_LINKS_PER_WAVE", 0),
            wpe = env_choice("NSPT_WAVES_PER_SIMD", 1);
  c->ks = find_set(c->nc, c->nrt, qt, false, lpw, wpe);
  if (!c->ks.drift) return bad("nspt_create: no compiled kernel instance for this order (SU(3) up to 32, SU(2) up to 24)");
  // fields past 2 GiB (the plaquette field, 6V series, is the widest) take
  // the instance built for them
  c->big = nspt::ser_alloc(6L * V, c->ks.nmax, c->NE) * 16 >= 0x80000000L;
  if (c->big) c->ks = find_set(c->nc, c->nrt, qt, true, lpw, wpe);
  for (int d = 0; d < 4; d++) c->g.L[d] = geom->sz[d];
  c->g.st[3] = 1;
  c->g.st[2] = c->g.L[3];
  c->g.st[1] = c->g.L[2] * c->g.L[3];
  c->g.st[0] = c->g.L[1] * c->g.L[2] * c->g.L[3];
  c->g.V = c->V;
  c->g.nrt = c->nrt;
  c->g.sub = 0;
  set_kernels(c, c->ks);
  c->nob = 2 * (c->nrt + 1) + 3 * c->nrt;
  c->obs_len = 2 * (c->nrt + 1) + c->nrt + 8 * c->nrt;
  // The plaquette form of the drift does fewer products but sums them in
  // another association than the reference; above 16 orders a step is
  // ill-conditioned in FP64 anyway (the spread of the step under a rounding
  // of its input exceeds 1e-12), up to 16 the staple form keeps the
  // reference's association.  NSPT_DRIFT_FORM=staples|plaquettes forces one.
  c->plaq_form = c->nrt > 16;
  if (const char* f = std::getenv("NSPT_DRIFT_FORM")) {
    if (!std::strcmp(f, "staples")) c->plaq_form = false;
    if (!std::strcmp(f, "plaquettes")) c->plaq_form = true;
  }
  *out = c;
  if (!c->hip_ok(hipStreamCreate(&c->stream), "hipStreamCreate")) return -3;
  const long P = 4L * V;
  // per-block partial sums: 4 nb_mu rows, or 8 nb_half for a checkerboard
  // sweep, for the build of this order with the fewest links per wave
  const int lpb_min = std::min(c->ks.lpb, min_links_per_wave(c->nc, c->ks.nmax));
  const long part_rows = std::max(4L * ((c->V + lpb_min - 1) / lpb_min), 8L * ((c->V / 2 + lpb_min - 1) / lpb_min));
  bool ok = c->alloc_field(c->U, P, true) && c->alloc_field(c->F, P, true) && c->alloc_field(c->D, P, false) &&
            c->alloc_field(c->D1, P, false) && c->alloc_field(c->L, P, true) && c->alloc_field(c->X, P, false) &&
            c->alloc_field(c->G, V, true) && (!c->plaq_form || c->alloc_field(c->Pf, 6L * V, true)) &&
            c->alloc(&c->fb, (size_t)c->nrt * c->NE * V) && c->alloc(&c->eta, (size_t)P * c->NE) &&
            c->alloc(&c->eta_tmp, (size_t)P * c->NE) &&
            c->alloc(&c->mm_part, (size_t)part_rows * c->nrt * c->NE) &&
            c->alloc(&c->MM, (size_t)4 * c->nrt * c->NE) && c->alloc(&c->obs_part, (size_t)part_rows * c->nob) &&
            c->alloc(&c->obs_mu, (size_t)4 * c->nob) &&
            c->alloc(&c->red_tmp, reduce_tmp_doubles(part_rows / 4, 4, std::max(c->nob, 2 * c->nrt * c->NE))) &&
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

int nspt_kernel_shape(const nspt_ctx* c, int* qt, int* links_per_wave, int* waves_per_simd) {
  if (!c || !c->ks.drift) return -1;
  if (qt) *qt = c->ks.qt;
  if (links_per_wave) *links_per_wave = c->ks.lpb;
  if (waves_per_simd) *waves_per_simd = c->ks.wpe;
  return 0;
}

int nspt_cold_start(nspt_ctx* c) {
  if (!c) return -1;
  const long P = 4L * c->V;
  if (!c->hip_ok(hipMemsetAsync(c->U.data, 0, sizeof(double2) * (size_t)nspt::ser_alloc(P, c->nrt, c->NE), c->stream),
                 "memset"))
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

// A replica's step is replayed from the batch's captured graphs, which
// track its U/F buffer roles and its observable rows: stepping it on its own
// would desynchronise them.
static int refuse_batch_step(nspt_ctx* c, const char* what) {
  return c->fail(std::string(what) + ": the context is a replica of a batch (step it with nspt_batch_evolve)") - 1;
}

int nspt_set_noise(nspt_ctx* c, const double* eta) {
  if (!c || !eta) return -1;
  if (c->in_batch) return refuse_batch_step(c, "nspt_set_noise");
  const size_t n = 4UL * c->V * c->NE;
  if (!c->hip_ok(hipMemcpyAsync(c->eta_tmp, eta, n * sizeof(double2), hipMemcpyHostToDevice, c->stream), "noise"))
    return -3;
  hipLaunchKernelGGL(k_eta_permute, dim3(grid_for((long)n, 256)), dim3(256), 0, c->stream,
                     (const double2*)c->eta_tmp, c->eta, c->NE, c->V, 1);
  if (!launch_ok(c, "eta_permute")) return -3;
  // the caller may reuse eta as soon as this returns
  if (!c->hip_ok(hipStreamSynchronize(c->stream), "sync")) return -3;
  c->noise_ready = true;
  return 0;
}

static int check_params(nspt_ctx* c, const nspt_step_params* p) {
  if (!p) return c->fail("null step parameters");
  if (!(p->eps > 0.0)) return c->fail("eps must be > 0");
  if (p->integrator != NSPT_EULER && p->integrator != NSPT_RK2 && p->integrator != NSPT_EULER_CHECKERBOARD)
    return c->fail("unknown integrator");
  if (p->gauge_fix != NSPT_GF_NONE && p->gauge_fix != NSPT_GF_FA && p->gauge_fix != NSPT_GF_CHECKERBOARD)
    return c->fail("unknown gauge fixing");
  if (p->integrator == NSPT_EULER_CHECKERBOARD || p->gauge_fix == NSPT_GF_CHECKERBOARD) {
    for (int d = 0; d < 4; d++)
      if (c->g.L[d] % 2) return c->fail("checkerboard passes need even lattice extents");
  }
  if (c->g.twist && p->gauge_fix == NSPT_GF_FA)
    return c->fail("the FA gauge fixing needs periodic boundary conditions (its Fourier modes are periodic)");
  if (p->integrator == NSPT_EULER_CHECKERBOARD && c->c1 != 0.0)
    return c->fail("the checkerboard sweep is for the Wilson action (rectangles couple links of one parity)");
  if (p->noise < NSPT_NOISE_HOST || p->noise > NSPT_NOISE_ZERO) return c->fail("unknown noise source");
  return 0;
}

int nspt_step(nspt_ctx* c, const nspt_step_params* p, double* obs) {
  if (!c) return -1;
  if (c->in_batch) return refuse_batch_step(c, "nspt_step");
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
  if (c->in_batch) return refuse_batch_step(c, "nspt_evolve");
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
  c->ks.drift(c->nblk_links, c->stream, c->g, c->U, none, none, 0.0, 0.0, 2, c->obs_part, none, 0.0);
  if (!launch_ok(c, "plaquette") || !obs_reduce(c, c->obs_dev)) return -3;
  if (!c->hip_ok(hipMemcpyAsync(out, c->obs_dev, sizeof(double) * 2 * (c->nrt + 1), hipMemcpyDeviceToHost,
                                c->stream),
                 "plaq"))
    return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_loops(nspt_ctx* c, double* out) {
  if (!c || !out) return -1;
  const int nrt = c->nrt, ncols = 2 * (nrt + 1);
  const long rows = 6L * c->g.nb_mu;  // per rectangle type: 6 planes x site blocks
  if (!c->loop_part) {
    const long P = 4L * c->V;
    // sized for the build of this order with the fewest links per wave (the
    // batch autotuner may move a replica to it after this first call)
    const int lpb_min = std::min(c->ks.lpb, min_links_per_wave(c->nc, c->ks.nmax));
    const long rows_max = 6L * ((c->V + lpb_min - 1) / lpb_min);
    if (!((c->L2.data || c->alloc_field(c->L2, P, true)) && c->alloc_field(c->L4, P, true) &&
          c->alloc(&c->loop_part, (size_t)6 * rows_max * ncols) &&
          c->alloc(&c->loop_tmp, reduce_tmp_doubles(rows_max, 6, ncols)) &&
          c->alloc(&c->loop_res, (size_t)6 * ncols))) {
      c->loop_part = nullptr;
      return -2;
    }
  }
  c->ks.link_pair(c->nblk_links, c->stream, c->g, c->U, c->L2, 1);
  c->ks.link_pair(c->nblk_links, c->stream, c->g, c->L2, c->L4, 2);
  if (!launch_ok(c, "long links")) return -3;
  c->ks.loops((int)(6 * rows), c->stream, c->g, c->U, c->L2, c->L4, c->loop_part);
  if (!launch_ok(c, "loops") || !reduce_groups(c, c->loop_part, ncols, rows, 6, ncols, c->loop_res, 1.0, c->loop_tmp))
    return -3;
  std::vector<double> r((size_t)6 * ncols);
  if (!c->hip_ok(hipMemcpyAsync(r.data(), c->loop_res, sizeof(double) * r.size(), hipMemcpyDeviceToHost, c->stream),
                 "loops") ||
      !c->hip_ok(hipStreamSynchronize(c->stream), "sync"))
    return -3;
  // w21 = 2x1 + 1x2, w22 = 2x2, w42 = 4x2 + 2x4, w44 = 4x4
  const int from[4][2] = {{0, 1}, {2, -1}, {3, 4}, {5, -1}};
  for (int w = 0; w < 4; w++)
    for (int q = 0; q < ncols; q++)
      out[w * ncols + q] = r[(size_t)from[w][0] * ncols + q] + (from[w][1] >= 0 ? r[(size_t)from[w][1] * ncols + q] : 0.0);
  return 0;
}

int nspt_set_action(nspt_ctx* c, double c1) {
  if (!c) return -1;
  if (c->in_batch) return c->fail("batch replicas run the Wilson action") - 1;
  if (c1 != 0.0 && !c->R.data) {
    const int ncols = 2 * (c->nrt + 1);
    // R, and the 2-link field the rectangles are formed from
    if (!(c->alloc_field(c->R, 4L * c->V, true) && c->alloc(&c->obs2_part, (size_t)c->nblk_links * ncols) &&
          c->alloc(&c->obs2_mu, (size_t)4 * ncols) && (c->L2.data || c->alloc_field(c->L2, 4L * c->V, true)))) {
      c->R.data = nullptr;
      return -2;
    }
  }
  c->c1 = c1;
  // the observable row gains the rectangle plaquette (N+1 complex)
  c->obs_len = 2 * (c->nrt + 1) + c->nrt + 8 * c->nrt + (c1 != 0.0 ? 2 * (c->nrt + 1) : 0);
  c->obs_cap = 0;
  return ensure_obs(c, 1) ? 0 : -2;
}

int nspt_set_twist(nspt_ctx* c, int mask, const double* omega) {
  if (!c) return -1;
  if (c->in_batch) return c->fail("batch replicas run on periodic lattices") - 1;
  if (mask < 0 || mask > 15) return c->fail("twist mask must be in 0..15") - 1;
  if (mask && !omega) return c->fail("null twist matrices") - 1;
  const int nc = c->nc;
  TwistTab t{};
  for (int d = 0; d < 4; d++) {
    if (!((mask >> d) & 1)) {
      for (int i = 0; i < nc; i++) {
        t.ph[d][i] = make_double2(1.0, 0.0);
        t.perm[d] |= (unsigned)i << (2 * i) | (unsigned)i << (8 + 2 * i);
      }
      continue;
    }
    // one entry of modulus 1 per row and per column, zeros elsewhere
    int seen = 0;
    for (int i = 0; i < nc; i++) {
      int col = -1;
      for (int k = 0; k < nc; k++) {
        const double re = omega[2 * ((d * nc + i) * nc + k)], im = omega[2 * ((d * nc + i) * nc + k) + 1];
        const double m2 = re * re + im * im;
        if (m2 == 0.0) continue;
        if (col >= 0 || std::fabs(m2 - 1.0) > 1e-12) col = -2;
        if (col == -2) break;
        col = k;
        t.ph[d][i] = make_double2(re, im);
      }
      if (col < 0 || ((seen >> col) & 1))
        return c->fail("twist matrices must have one entry of modulus 1 per row and column (as twistmatrices.h)") - 1;
      seen |= 1 << col;
      t.perm[d] |= (unsigned)col << (2 * i) | (unsigned)i << (8 + 2 * col);
    }
  }
  if (!c->twist_tab && !c->alloc(&c->twist_tab, 1)) return -2;
  if (!c->hip_ok(hipMemcpyAsync(c->twist_tab, &t, sizeof(t), hipMemcpyHostToDevice, c->stream), "twist") ||
      !c->hip_ok(hipStreamSynchronize(c->stream), "sync"))
    return -3;
  c->g.twist = mask;
  c->g.tw = c->twist_tab;
  return 0;
}

int nspt_unitarize(nspt_ctx* c) {
  if (!c) return -1;
  // U <- exp(reH(log U)): the log kernel, then the zero-mode kernel with a zero mean
  c->ks.logf(c->nblk_links, c->stream, c->g, c->U, c->L);
  if (!launch_ok(c, "log") ||
      !c->hip_ok(hipMemsetAsync(c->MM, 0, sizeof(double2) * 4 * c->nrt * c->NE, c->stream), "memset"))
    return -3;
  c->ks.zero_mode(c->nblk_links, c->stream, c->g, c->L, c->MM, c->U, c->X);
  if (!launch_ok(c, "unitarize")) return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_gauge_wilson(nspt_ctx* c, const nspt_step_params* p, double* obs) {
  if (!c) return -1;
  if (!p) return c->fail("null step parameters") - 1;
  nspt_step_params q = *p;
  q.zero_modes = 0;
  q.gauge_fix = NSPT_GF_NONE;
  return nspt_step(c, &q, obs);
}

int nspt_zero_modes_subtraction(nspt_ctx* c) {
  if (!c) return -1;
  // MM_mu = (1/V) sum_x log U_mu(x), then U <- exp(reH(log U - MM_mu))
  c->ks.logf(c->nblk_links, c->stream, c->g, c->U, c->L);
  if (!launch_ok(c, "log")) return -3;
  hipLaunchKernelGGL(k_field_mean, dim3(4 * c->nrt * c->NE), dim3(MEAN_THREADS), 0, c->stream,
                     (const double2*)c->L.data, c->V, c->nrt, c->NE, reinterpret_cast<double2*>(c->MM));
  if (!launch_ok(c, "field mean")) return -3;
  c->ks.zero_mode(c->nblk_links, c->stream, c->g, c->L, c->MM, c->U, c->X);
  if (!launch_ok(c, "zero_mode")) return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_stochastic_gauge_fixing(nspt_ctx* c, double alpha) {
  if (!c) return -1;
  for (int d = 0; d < 4; d++)
    if (c->g.L[d] % 2) return c->fail("the checkerboard gauge fixing needs even lattice extents") - 1;
  Geo gs = c->g;
  for (int par = 0; par < 2; par++) {
    gs.sub = 1 + 4 * par;
    c->ks.sgf(c->g.nb_half, c->stream, gs, c->U, alpha);
  }
  if (!launch_ok(c, "checkerboard gauge fixing")) return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_fa_gauge_fixing(nspt_ctx* c, double alpha) {
  if (!c) return -1;
  if (c->g.twist) return c->fail("the FA gauge fixing needs periodic boundary conditions") - 1;
  c->ks.logf(c->nblk_links, c->stream, c->g, c->U, c->L);
  if (!launch_ok(c, "log") || !fa_apply(c, c->L, alpha, nullptr, false)) return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_momentum_norm(nspt_ctx* c, double* knorm) {
  if (!c || !knorm) return -1;
  if (c->U.P != 4L * c->V) return c->fail("nspt_momentum_norm: unexpected field layout") - 1;
  const long n = (long)c->nrt * c->V;
  if (!c->knorm_dev && !c->alloc(&c->knorm_dev, (size_t)n)) return -2;
  // per direction: its NE nrt planes through the FFT buffer (passes 3, 4),
  // then their norms added
  for (int mu = 0; mu < 4; mu++) {
    if (!fft_pass(c, 3, (const double2*)c->U.data, 0.0, mu) || !fft_pass(c, 4, (const double2*)c->U.data, 0.0))
      return -3;
    hipLaunchKernelGGL(k_knorm_acc, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, (const double2*)c->fb,
                       (long)c->V, c->NE, c->nrt, mu == 0 ? 1 : 0, c->knorm_dev);
    if (!launch_ok(c, "knorm")) return -3;
  }
  if (!c->hip_ok(hipMemcpyAsync(knorm, c->knorm_dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream),
                 "knorm"))
    return -3;
  return c->hip_ok(hipStreamSynchronize(c->stream), "sync") ? 0 : -3;
}

int nspt_drift(nspt_ctx* c, double eps, double* aos_out) {
  if (!c || !aos_out) return -1;
  SerField none{nullptr, nullptr, 0};
  if (c->c1 != 0.0) {
    c->ks.link_pair(c->nblk_links, c->stream, c->g, c->U, c->L2, 1);
    c->ks.rect(c->nblk_links, c->stream, c->g, c->U, c->L2, c->R, nullptr);
  }
  drift_p(c, c->U, c->D, none, -eps / (2.0 * c->nc), 0.0, 0, nullptr, c->c1 != 0.0 ? c->R : none);
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

namespace {
// FP64 FMA throughput: eight independent chains per lane, so that the
// FMA latency is covered by the chains and the waves of a SIMD.
__global__ void k_fma_probe(double* out, int iters, double a, double b) {
  const double t = 1e-9 * (double)(blockIdx.x * blockDim.x + threadIdx.x);
  double x[8];
#pragma unroll
  for (int c = 0; c < 8; c++) x[c] = t + 0.125 * c;
  for (int i = 0; i < iters; i++) {
#pragma unroll
    for (int c = 0; c < 8; c++) x[c] = fma(x[c], a, b);
  }
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < 8; c++) s += x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}
// Device-to-device stream copy, 16 bytes per lane per iteration.
__global__ void k_copy_probe(const double2* __restrict__ src, double2* __restrict__ dst, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) dst[i] = src[i];
}
}  // namespace

extern "C" int nspt_measure_peaks(double* fp64_tflops, double* hbm_gbs) {
  if (!fp64_tflops || !hbm_gbs) return -1;
  int dev = 0, ncu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
    return -3;
  const int threads = 256, blocks = ncu * 16, iters = 4096;
  const long n = 64L << 20;  // 64 Mi double2 = 1 GiB per buffer
  double* out = nullptr;
  double2 *src = nullptr, *dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = 0;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == 0) rc = -3;
    return rc == 0;
  };
  if (ok(hipMalloc(&out, sizeof(double) * threads * blocks)) && ok(hipMalloc(&src, sizeof(double2) * n)) &&
      ok(hipMalloc(&dst, sizeof(double2) * n)) && ok(hipEventCreate(&e0)) && ok(hipEventCreate(&e1)) &&
      ok(hipMemset(src, 0, sizeof(double2) * n))) {
    float best_f = 1e30f, best_c = 1e30f;
    for (int rep = 0; rep < 4 && rc == 0; rep++) {  // the first of each is a warm-up
      float ms = 0.f;
      hipLaunchKernelGGL(k_fma_probe, dim3(blocks), dim3(threads), 0, nullptr, out, 64, 1.0000001, 1e-7);
      (void)hipEventRecord(e0, nullptr);
      hipLaunchKernelGGL(k_fma_probe, dim3(blocks), dim3(threads), 0, nullptr, out, iters, 1.0000001, 1e-7);
      (void)hipEventRecord(e1, nullptr);
      if (ok(hipEventSynchronize(e1)) && ok(hipEventElapsedTime(&ms, e0, e1)) && rep > 0) best_f = std::min(best_f, ms);
      (void)hipEventRecord(e0, nullptr);
      hipLaunchKernelGGL(k_copy_probe, dim3(ncu * 32), dim3(256), 0, nullptr, (const double2*)src, dst, n);
      (void)hipEventRecord(e1, nullptr);
      if (ok(hipEventSynchronize(e1)) && ok(hipEventElapsedTime(&ms, e0, e1)) && rep > 0) best_c = std::min(best_c, ms);
    }
    ok(hipGetLastError());
    if (rc == 0) {
      *fp64_tflops = 2.0 * 8.0 * iters * (double)threads * blocks / (best_f * 1e-3) / 1e12;
      *hbm_gbs = 2.0 * sizeof(double2) * (double)n / (best_c * 1e-3) / 1e9;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void* p : {(void*)out, (void*)src, (void*)dst})
    if (p) (void)hipFree(p);
  return rc;
}

// ================================================================== batches
// Independent replicas stepped together.  A step of every replica is
// captured once into a HIP graph -- one branch per replica, on the
// replica's own stream, so the branches run concurrently -- and replayed.
// Two graphs are kept because a step swaps the roles of the U and F
// buffers: graph 0 is captured for the current buffer roles, graph 1 for the
// swapped ones; after each replay the host-side roles are swapped to match.
// One launch of a replica-table step: which kernel, its grid, and its
// table (an offset into the batch's packed tables until uploaded).
struct TabLaunch {
  enum Kind { NOISE, DRIFT, PLAQ, DRIFTP, REDUCE, OBSFIN, UPDATE, ZEROMODE, FFT, GEXP, TRANSFORM, COUNT } kind;
  dim3 grid;
  size_t shm = 0;
  size_t off = 0;
  const void* tab = nullptr;
  int pass = 0, maxn = 16;
  int nrep = 0;  // replicas of the table (COUNT)
};

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
  // Step form: one graph branch per replica (each replica's own launches),
  // or (table) one launch per stage for every replica of a group, each
  // replica's arguments in a device table, the groups (1 to 8) concurrent
  // graph branches.  The table form needs replicas that share the
  // integrator, gauge fixing (none or FA), zero-mode choice and kernel
  // build.
  bool table = false;
  bool table_ok = false;
  int groups = 1;
  std::vector<void*> tab_dev[2];                   // per group
  std::vector<std::vector<TabLaunch>> tab_seq[2];  // per group
  // the last autotuner run's candidates: form, tile edge, links per wave,
  // waves per SIMD
  std::vector<std::array<int, 4>> tune_cands;
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

// ---- replica-table step
int round8(long n) { return (int)((n + 7) / 8 * 8); }

struct TablePlan {
  std::vector<char> host;
  std::vector<TabLaunch> seq;
  template <class T>
  size_t add(const std::vector<T>& v) {
    const size_t off = (host.size() + 255) / 256 * 256;
    host.resize(off + sizeof(T) * v.size());
    std::memcpy(host.data() + off, v.data(), sizeof(T) * v.size());
    return off;
  }
  template <class T>
  void launch(TabLaunch::Kind k, dim3 grid, const std::vector<T>& v, size_t shm = 0, int pass = 0, int maxn = 16) {
    TabLaunch L;
    L.kind = k;
    L.grid = grid;
    L.shm = shm;
    L.off = add(v);
    L.pass = pass;
    L.maxn = maxn;
    seq.push_back(L);
  }
};

// Deterministic reduction of each replica's (4 rows x ncols) groups, as
// reduce_groups forms it alone: the same chunks per level, levels until
// one chunk; replicas with fewer levels sit out the later ones.
struct ReduceRep {
  const double* in;
  long ld, rows;
  double* out;
  double scale;
  double* tmp;
};
void plan_reduce(TablePlan& tp, const std::vector<ReduceRep>& rr, int ncols) {
  const size_t R = rr.size();
  std::vector<const double*> src(R);
  std::vector<long> src_ld(R), rows(R);
  std::vector<size_t> half(R);
  std::vector<bool> done(R, false);
  for (size_t r = 0; r < R; r++) {
    src[r] = rr[r].in;
    src_ld[r] = rr[r].ld;
    rows[r] = rr[r].rows;
    half[r] = reduce_tmp_doubles(rr[r].rows, 4, ncols) / 2;
  }
  for (int level = 0;; level++) {
    std::vector<ReduceArgs> lv(R);
    int maxnch = 0;
    for (size_t r = 0; r < R; r++) {
      ReduceArgs a{};
      a.ncols = ncols;
      a.per = REDUCE_PER;
      if (done[r]) {
        a.nch = 0;
        lv[r] = a;
        continue;
      }
      const long nch = (rows[r] + REDUCE_PER - 1) / REDUCE_PER;
      a.in = src[r];
      a.ld = src_ld[r];
      a.in_mu = rows[r];
      a.rows = rows[r];
      a.out_ld = ncols;
      if (nch == 1) {
        a.out = rr[r].out;
        a.out_mu = 1;
        a.scale = rr[r].scale;
        a.nch = 1;
        done[r] = true;
      } else {
        double* dst = rr[r].tmp + (level % 2) * half[r];
        a.out = dst;
        a.out_mu = nch;
        a.scale = 1.0;
        a.nch = (int)nch;
        src[r] = dst;
        src_ld[r] = ncols;
        rows[r] = nch;
      }
      maxnch = std::max(maxnch, a.nch);
      lv[r] = a;
    }
    if (maxnch == 0) break;
    tp.launch(TabLaunch::REDUCE, dim3((ncols + 63) / 64, maxnch, 4 * (unsigned)R), lv);
  }
}

// The launches of one step of every replica, parity q: q = 0 steps the
// fields the contexts call U (writing F), q = 1 the other way round (the
// roles after one step).  Mirrors do_step for the Jacobi Euler / RK2 step
// with zero modes optional and no or FA gauge fixing, Wilson action,
// periodic lattice, device noise.
void plan_table_step(nspt_batch* b, const std::vector<int>& idx, int q, bool with_obs, TablePlan& tp) {
  const size_t R = idx.size();
  nspt_ctx* c0 = b->reps[idx[0]];
  const nspt_step_params& p0 = b->params[idx[0]];
  // the group's replicas and step parameters, in table order
  std::vector<nspt_ctx*> reps(R);
  std::vector<nspt_step_params> params(R);
  for (size_t r = 0; r < R; r++) {
    reps[r] = b->reps[idx[r]];
    params[r] = b->params[idx[r]];
  }
  const bool rk2 = p0.integrator == NSPT_RK2;
  const bool fa = p0.gauge_fix == NSPT_GF_FA;
  const bool zm = p0.zero_modes != 0;
  const bool need_log = zm || fa;
  const bool pf = c0->plaq_form;
  const SerField none{nullptr, nullptr, 0};
  auto cur = [&](nspt_ctx* c) { return q == 0 ? c->U : c->F; };
  auto nxt = [&](nspt_ctx* c) { return q == 0 ? c->F : c->U; };
  long maxl = 0, maxs = 0, maxn4 = 0;
  for (auto* c : reps) {
    maxl = std::max(maxl, (long)c->nblk_links);
    maxs = std::max(maxs, (long)c->nblk_sites);
    maxn4 = std::max(maxn4, (long)grid_for(4L * c->V, 256));
  }
  const unsigned gl = round8(maxl), gs = round8(maxs), gp = round8(6 * maxs), nr = (unsigned)R;
  // noise
  {
    std::vector<NoiseArgs> v(R);
    for (size_t r = 0; r < R; r++) {
      nspt_ctx* c = reps[r];
      v[r] = NoiseArgs{c->nc, c->V, params[r].seed, params[r].step, c->ctr_dev, c->eta};
    }
    tp.launch(TabLaunch::NOISE, dim3((unsigned)maxn4, nr), v);
  }
  auto tau = [&](size_t r) { return -params[r].eps / (2.0 * reps[r]->nc); };
  auto drift = [&](bool second) {
    if (pf) {
      std::vector<PlaqArgs> pv(R);
      std::vector<DriftPArgs> dv(R);
      for (size_t r = 0; r < R; r++) {
        nspt_ctx* c = reps[r];
        const SerField Uin = second ? nxt(c) : cur(c);
        pv[r] = PlaqArgs{c->g, Uin, c->Pf, 6 * c->g.nb_mu};
        dv[r] = DriftPArgs{c->g, Uin, c->Pf, (rk2 && !second) ? c->D1 : c->D, second ? c->D1 : none, tau(r),
                           second ? c->nc * params[r].eps / 6.0 : 0.0, second ? 1 : 0,
                           (!second && with_obs) ? c->obs_part : nullptr, none, 0.0, c->nblk_links};
      }
      tp.launch(TabLaunch::PLAQ, dim3(gp, nr), pv);
      tp.launch(TabLaunch::DRIFTP, dim3(gl, nr), dv);
    } else {
      std::vector<DriftArgs> dv(R);
      for (size_t r = 0; r < R; r++) {
        nspt_ctx* c = reps[r];
        const SerField Uin = second ? nxt(c) : cur(c);
        dv[r] = DriftArgs{c->g, Uin, (rk2 && !second) ? c->D1 : c->D, second ? c->D1 : none, tau(r),
                          second ? c->nc * params[r].eps / 6.0 : 0.0, second ? 1 : 0,
                          (!second && with_obs) ? c->obs_part : nullptr, none, 0.0, c->nblk_links};
      }
      tp.launch(TabLaunch::DRIFT, dim3(gl, nr), dv);
    }
  };
  auto update = [&](bool final) {
    std::vector<UpdateArgs> v(R);
    for (size_t r = 0; r < R; r++) {
      nspt_ctx* c = reps[r];
      // the final update with zero modes writes only L (the links are
      // re-formed from exp(L - MM) afterwards)
      v[r] = UpdateArgs{c->g, (rk2 && !final) ? c->D1 : c->D, c->eta, std::sqrt(params[r].eps), cur(c),
                        (final && zm) ? none : nxt(c), (final && need_log) ? c->L : none,
                        (final && zm) ? c->mm_part : nullptr, c->nblk_links};
    }
    tp.launch(TabLaunch::UPDATE, dim3(gl, nr), v);
  };
  drift(false);
  if (with_obs) {
    std::vector<ReduceRep> rr(R);
    std::vector<ObsFinArgs> fv(R);
    for (size_t r = 0; r < R; r++) {
      nspt_ctx* c = reps[r];
      rr[r] = ReduceRep{c->obs_part, c->nob, c->g.nb_mu, c->obs_mu, 1.0, c->red_tmp};
      fv[r] = ObsFinArgs{c->obs_mu, c->nrt, c->nc, c->V, c->obs_dev, c->ctr_dev + 1, c->obs_len};
    }
    plan_reduce(tp, rr, c0->nob);
    tp.launch(TabLaunch::OBSFIN, dim3(nr), fv);
  }
  if (rk2) {
    update(false);
    drift(true);
  }
  update(true);
  if (zm) {
    const int ncol = 2 * c0->nrt * c0->NE;
    std::vector<ReduceRep> rr(R);
    for (size_t r = 0; r < R; r++) {
      nspt_ctx* c = reps[r];
      rr[r] = ReduceRep{reinterpret_cast<const double*>(c->mm_part), ncol, c->g.nb_mu,
                        reinterpret_cast<double*>(c->MM), 1.0 / (double)c->V, c->red_tmp};
    }
    plan_reduce(tp, rr, ncol);
    if (!fa) {
      std::vector<ZeroModeArgs> v(R);
      for (size_t r = 0; r < R; r++) {
        nspt_ctx* c = reps[r];
        v[r] = ZeroModeArgs{c->g, c->L, c->MM, nxt(c), c->X, c->nblk_links};
      }
      tp.launch(TabLaunch::ZEROMODE, dim3(gl, nr), v);
    }
  }
  if (fa) {
    const long ncomp = (long)c0->nrt * c0->NE;
    for (int pass = 0; pass < 3; pass++) {
      std::vector<FftArgs> v(R);
      int maxn = 16;
      for (auto* c : reps) {
        const FftTile t = fft_tile(c->g, pass);
        if (std::max(t.na, t.nb) > 16) maxn = FFT_MAXN;
      }
      long gx = 0;
      size_t shm = 0;
      for (size_t r = 0; r < R; r++) {
        nspt_ctx* c = reps[r];
        const FftTile t = fft_tile(c->g, pass);
        const bool direct = !fft_unrolled(t.na, maxn) || !fft_unrolled(t.nb, maxn);
        shm = std::max(shm, fft_shm(t, c->g, direct));
        v[r] = FftArgs{c->g, c->L.data, c->fb, params[r].alpha, (int)((long)t.tiles * ncomp), c->NE};
        gx = std::max(gx, (long)t.tiles * ncomp);
      }
      tp.launch(TabLaunch::FFT, dim3((unsigned)gx, nr), v, shm, pass, maxn);
    }
    std::vector<GexpArgs> gv(R);
    std::vector<TransformArgs> tv(R);
    for (size_t r = 0; r < R; r++) {
      nspt_ctx* c = reps[r];
      gv[r] = GexpArgs{c->g, c->fb, c->G, c->nblk_sites};
      tv[r] = TransformArgs{c->g, c->G, nxt(c), zm ? c->L : none,
                            zm ? reinterpret_cast<const double2*>(c->MM) : nullptr, c->nblk_links};
    }
    tp.launch(TabLaunch::GEXP, dim3(gs, nr), gv);
    tp.launch(TabLaunch::TRANSFORM, dim3(gl, nr), tv);
  }
  std::vector<unsigned long long*> ctrs(R);
  for (size_t r = 0; r < R; r++) ctrs[r] = reps[r]->ctr_dev;
  tp.launch(TabLaunch::COUNT, dim3(1), ctrs);
  tp.seq.back().nrep = (int)R;
}

void table_launch(nspt_batch* b, const TabLaunch& L, hipStream_t st) {
  const KernelSet& ks = b->reps[0]->ks;
  const int gx = (int)L.grid.x, gy = (int)L.grid.y;
  switch (L.kind) {
    case TabLaunch::NOISE:
      hipLaunchKernelGGL(k_noise_philox_tab, L.grid, dim3(256), 0, st, (const NoiseArgs*)L.tab);
      break;
    case TabLaunch::DRIFT: ks.drift_tab(gx, gy, st, (const nspt::DriftArgs*)L.tab); break;
    case TabLaunch::PLAQ: ks.plaq_field_tab(gx, gy, st, (const nspt::PlaqArgs*)L.tab); break;
    case TabLaunch::DRIFTP: ks.drift_p_tab(gx, gy, st, (const nspt::DriftPArgs*)L.tab); break;
    case TabLaunch::REDUCE:
      hipLaunchKernelGGL(k_reduce_rows_tab, L.grid, dim3(64), 0, st, (const ReduceArgs*)L.tab);
      break;
    case TabLaunch::OBSFIN:
      hipLaunchKernelGGL(k_obs_finalize_tab, L.grid, dim3(256), 0, st, (const ObsFinArgs*)L.tab);
      break;
    case TabLaunch::UPDATE: ks.update_tab(gx, gy, st, (const nspt::UpdateArgs*)L.tab); break;
    case TabLaunch::ZEROMODE: ks.zero_mode_tab(gx, gy, st, (const nspt::ZeroModeArgs*)L.tab); break;
    case TabLaunch::FFT:
      if (L.maxn == 16)
        hipLaunchKernelGGL(k_fa_fft_tab<16>, L.grid, dim3(FFT_THREADS), L.shm, st, L.pass, (const FftArgs*)L.tab);
      else
        hipLaunchKernelGGL(k_fa_fft_tab<FFT_MAXN>, L.grid, dim3(FFT_THREADS), L.shm, st, L.pass,
                           (const FftArgs*)L.tab);
      break;
    case TabLaunch::GEXP: ks.gexp_tab(gx, gy, st, (const nspt::GexpArgs*)L.tab); break;
    case TabLaunch::TRANSFORM: ks.transform_tab(gx, gy, st, (const nspt::TransformArgs*)L.tab); break;
    case TabLaunch::COUNT:
      hipLaunchKernelGGL(k_count_tab, dim3(1), dim3(64), 0, st, (unsigned long long* const*)L.tab, L.nrep);
      break;
  }
}

// Whether the replicas can share one launch per stage.
bool table_possible(const nspt_batch* b) {
  const nspt_ctx* c0 = b->reps[0];
  const nspt_step_params& p0 = b->params[0];
  if (p0.gauge_fix != NSPT_GF_NONE && p0.gauge_fix != NSPT_GF_FA) return false;
  for (size_t r = 0; r < b->reps.size(); r++) {
    const nspt_ctx* c = b->reps[r];
    const nspt_step_params& p = b->params[r];
    if (p.integrator != p0.integrator || p.gauge_fix != p0.gauge_fix || (p.zero_modes != 0) != (p0.zero_modes != 0))
      return false;
    if (c->big != c0->big || c->plaq_form != c0->plaq_form || c->c1 != 0.0 || c->g.twist) return false;
  }
  return true;
}

void batch_free_tables(nspt_batch* b) {
  for (int q = 0; q < 2; q++) {
    for (void* p : b->tab_dev[q])
      if (p) (void)hipFree(p);
    b->tab_dev[q].clear();
    b->tab_seq[q].clear();
  }
}

// The replicas of each table group: largest first, each to the group with
// the fewest links so far (ties to the lower group), so that the groups'
// launches carry about the same work; ascending replica order within a group.
std::vector<std::vector<int>> table_groups(const nspt_batch* b) {
  const int R = (int)b->reps.size();
  const int G = std::max(1, std::min(b->groups, R));
  std::vector<int> order(R);
  for (int r = 0; r < R; r++) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return b->reps[x]->V > b->reps[y]->V; });
  std::vector<std::vector<int>> grp(G);
  std::vector<long> load(G, 0);
  for (int r : order) {
    const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    grp[g].push_back(r);
    load[g] += b->reps[r]->V;
  }
  for (auto& v : grp) std::sort(v.begin(), v.end());
  return grp;
}

// Plan and upload the tables of both parities, one per group.
bool batch_plan_tables(nspt_batch* b, bool with_obs) {
  batch_free_tables(b);
  const auto grp = table_groups(b);
  for (int q = 0; q < 2; q++) {
    for (const auto& idx : grp) {
      TablePlan tp;
      plan_table_step(b, idx, q, with_obs, tp);
      void* dev = nullptr;
      if (hipMalloc(&dev, tp.host.size()) != hipSuccess ||
          hipMemcpy(dev, tp.host.data(), tp.host.size(), hipMemcpyHostToDevice) != hipSuccess) {
        if (dev) (void)hipFree(dev);
        b->err = "replica tables: allocation or upload failed";
        return false;
      }
      b->tab_dev[q].push_back(dev);
      for (auto& L : tp.seq) L.tab = static_cast<const char*>(dev) + L.off;
      b->tab_seq[q].push_back(tp.seq);
    }
  }
  return true;
}

// Capture the two step graphs (with or without observables).
bool batch_build(nspt_batch* b, bool with_obs) {
  batch_drop_graphs(b);
  if (b->table) {
    if (!batch_plan_tables(b, with_obs)) return false;
    bool ok = true;
    for (int g = 0; g < 2 && ok; g++) {
      if (hipStreamBeginCapture(b->master, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        b->err = "hipStreamBeginCapture failed";
        return false;
      }
      const size_t G = b->tab_seq[g].size();
      if (G == 1) {
        for (const auto& L : b->tab_seq[g][0]) table_launch(b, L, b->master);
      } else {
        // each group a branch forked from the master stream, on the stream
        // of a replica of its own (group j: replica j's stream)
        (void)hipEventRecord(b->fork, b->master);
        for (size_t j = 0; j < G; j++) {
          hipStream_t st = b->reps[j]->stream;
          (void)hipStreamWaitEvent(st, b->fork, 0);
          for (const auto& L : b->tab_seq[g][j]) table_launch(b, L, st);
          (void)hipEventRecord(b->join[j], st);
          (void)hipStreamWaitEvent(b->master, b->join[j], 0);
        }
      }
      hipGraph_t graph = nullptr;
      if (hipStreamEndCapture(b->master, &graph) != hipSuccess || !graph) {
        b->err = "hipStreamEndCapture failed";
        ok = false;
      }
      if (ok && hipGraphInstantiate(&b->exec[g], graph, nullptr, nullptr, 0) != hipSuccess) {
        b->err = "hipGraphInstantiate failed";
        ok = false;
      }
      if (graph) (void)hipGraphDestroy(graph);
    }
    if (!ok) batch_drop_graphs(b);
    b->parity = 0;
    b->with_obs = with_obs;
    return ok;
  }
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
  // the host source stays alive until the copies have completed
  const unsigned long long v[2] = {step, row};
  bool ok = true;
  for (auto* c : b->reps) {
    if (!c->hip_ok(hipMemcpyAsync(c->ctr_dev, v, sizeof(v), hipMemcpyHostToDevice, b->master), "counters")) {
      b->err = c->err;
      ok = false;
      break;
    }
  }
  return batch_sync(b) && ok;
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
    c->in_batch = true;
    b->reps.push_back(c);
    b->params.push_back(params[r]);
  }
  if (hipStreamCreate(&b->master) != hipSuccess || hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
    return b->fail("nspt_batch_create: stream/event creation failed");
  b->join.resize(nrep);
  for (auto& e : b->join)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return b->fail("event creation failed");
  b->table_ok = table_possible(b);
  return batch_set_counters(b, 0, 0) ? 0 : -3;
}

int nspt_batch_set_form(nspt_batch* b, int table) {
  if (!b) return -1;
  if (table < 0 || table > 8) return b->fail("nspt_batch_set_form: form is 0 (branches) or 1-8 table groups");
  if (table && !b->table_ok)
    return b->fail("nspt_batch_set_form: the replicas differ in integrator, gauge fixing (none or FA only), "
                   "zero modes or kernel build");
  const int groups = table ? std::min(table, (int)b->reps.size()) : b->groups;
  if ((table != 0) != b->table || groups != b->groups) {
    b->table = table != 0;
    b->groups = groups;
    batch_drop_graphs(b);
  }
  return 0;
}

int nspt_batch_form(const nspt_batch* b) { return b ? (b->table ? b->groups : 0) : -1; }

int nspt_batch_destroy(nspt_batch* b) {
  if (!b) return -1;
  batch_drop_graphs(b);
  batch_free_tables(b);
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
      static const unsigned long long zero = 0;  // outlives the queued copies
      for (auto* c : b->reps) {
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
  // the compiled shapes (tile edge, links per wave, waves per SIMD) of the
  // replicas' order
  std::vector<std::array<int, 3>> shapes;
  auto see = [&](KernelSet s) {
    if (s.nc == c0->nc && s.nmax == c0->ks.nmax && s.big == c0->big) shapes.push_back({s.qt, s.lpw, s.wpe});
  };
#define NSPT_SEE(a, bb, q, l, w)                         \
  see(nspt::kernel_set_##a##_##bb##_q##q##_l##l##_w##w()); \
  see(nspt::kernel_set_##a##_##bb##_q##q##_l##l##_w##w##_big());
  NSPT_INSTANCES(NSPT_SEE)
#undef NSPT_SEE
  // stage 1: step form (graph branches, then when the replicas allow it the
  // replica table in 1, 2 and 4 concurrent groups, as far as there are
  // replicas) x compiled tile edge at the default occupancy; stage 2: the
  // occupancy variants compiled for the chosen edge, in the chosen form
  std::vector<int> edges;
  for (int q : {4, 3, 5, 2})
    for (const auto& sh : shapes)
      if (sh[0] == q && sh[1] == 0 && sh[2] == 1) edges.push_back(q);
  std::vector<int> forms = {0};
  if (b->table_ok)
    for (int g : {1, 2, 4})
      if (g == 1 || g <= (int)b->reps.size() / 2) forms.push_back(g);
  b->tune_cands.clear();
  for (int f : forms)
    for (int q : edges) b->tune_cands.push_back({f, q, 0, 1});
  const size_t stage1 = b->tune_cands.size();
  const int form0 = b->table ? b->groups : 0;
  const int qt0 = c0->ks.qt, lpw0 = c0->ks.lpw, wpe0 = c0->ks.wpe;
  // snapshot fields and counters: tuning steps are not part of the run
  std::vector<double2*> snap(b->reps.size(), nullptr);
  std::vector<unsigned long long> ctr(2 * b->reps.size());
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool snapped = false;
  double best = 1e30;
  std::array<int, 4> best_c = {form0, qt0, lpw0, wpe0};
  int rc = 0;
  auto set_form = [&](const std::array<int, 4>& k) {
    b->table = k[0] != 0;
    if (k[0]) b->groups = k[0];
    for (auto* c : b->reps) set_kernels(c, find_set(c->nc, c->nrt, k[1], c->big, k[2], k[3]));
  };
  for (size_t r = 0; r < b->reps.size() && rc == 0; r++) {
    nspt_ctx* c = b->reps[r];
    const size_t n = (size_t)nspt::ser_alloc(4L * c->V, c->nrt, c->NE) + 4 * (size_t)c->V;
    if (hipMalloc(&snap[r], n * sizeof(double2)) != hipSuccess) {
      rc = b->fail("tune: snapshot allocation failed");
      break;
    }
    (void)hipMemcpyAsync(snap[r], c->U.data, (n - 4 * (size_t)c->V) * sizeof(double2), hipMemcpyDeviceToDevice,
                         b->master);
    (void)hipMemcpyAsync(snap[r] + (n - 4 * (size_t)c->V), c->U.flag, 4 * (size_t)c->V * sizeof(double2),
                         hipMemcpyDeviceToDevice, b->master);
    (void)hipMemcpyAsync(&ctr[2 * r], c->ctr_dev, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, b->master);
  }
  if (rc == 0 && !batch_sync(b)) rc = -3;
  snapped = rc == 0;
  if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = b->fail("tune: events");
  for (size_t i = 0; i < b->tune_cands.size() && rc == 0; i++) {
    set_form(b->tune_cands[i]);
    if (!batch_build(b, false) || !batch_replay(b, 1)) {
      rc = -3;
      break;
    }
    (void)hipEventRecord(e0, b->master);
    if (!batch_replay(b, steps)) {
      rc = -3;
      break;
    }
    (void)hipEventRecord(e1, b->master);
    if (!batch_sync(b)) {
      rc = -3;
      break;
    }
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double per = ms / steps;
    if (ms_per_step && i < 16) ms_per_step[i] = per;
    if (per < best) {
      best = per;
      best_c = b->tune_cands[i];
    }
    // stage 2 joins the list once stage 1 is timed, in the best form and
    // edge so far
    if (i + 1 == stage1) {
      for (const auto& sh : shapes)
        if (sh[0] == best_c[1] && !(sh[1] == 0 && sh[2] == 1) && b->tune_cands.size() < 16)
          b->tune_cands.push_back({best_c[0], sh[0], sh[1], sh[2]});
    }
  }
  // every exit restores the run: the form and kernel build (the best on
  // success, the original otherwise), the fields and the counters
  if (rc == 0)
    set_form(best_c);
  else
    set_form({form0, qt0, lpw0, wpe0});
  batch_drop_graphs(b);
  if (snapped) {
    for (size_t r = 0; r < b->reps.size(); r++) {
      nspt_ctx* c = b->reps[r];
      const size_t n = (size_t)nspt::ser_alloc(4L * c->V, c->nrt, c->NE);
      (void)hipMemcpyAsync(c->U.data, snap[r], n * sizeof(double2), hipMemcpyDeviceToDevice, b->master);
      (void)hipMemcpyAsync(c->U.flag, snap[r] + n, 4 * (size_t)c->V * sizeof(double2), hipMemcpyDeviceToDevice,
                           b->master);
      (void)hipMemcpyAsync(c->ctr_dev, &ctr[2 * r], 2 * sizeof(unsigned long long), hipMemcpyHostToDevice,
                           b->master);
    }
    if (!batch_sync(b) && rc == 0) rc = -3;
  }
  for (auto* p : snap)
    if (p) (void)hipFree(p);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (rc == 0 && qt_chosen) *qt_chosen = best_c[1];
  return rc;
}

int nspt_batch_tune_candidates(const nspt_batch* b, int* desc, int max) {
  if (!b || (max > 0 && !desc)) return -1;
  const int n = (int)b->tune_cands.size();
  for (int i = 0; i < n && i < max; i++)
    for (int k = 0; k < 4; k++) desc[4 * i + k] = b->tune_cands[i][k];
  return n;
}

int nspt_batch_kernel_shape(const nspt_batch* b, int* qt, int* links_per_wave, int* waves_per_simd) {
  if (!b || b->reps.empty()) return -1;
  return nspt_kernel_shape(b->reps[0], qt, links_per_wave, waves_per_simd);
}

}  // extern "C"
