// nspt_rng.cpp -- host noise sources of the engine:
//  * Luescher's RANLUX double-precision generator "ranlxd" (portable
//    formulation, ranlxd/ranlxd.cc:332-531; published in M. Luescher,
//    "Algorithms used in ranlux v3.0", 2001) with the su(N) noise mapping of
//    MyMath.cc:295-322 (the reference built without USE_RAND55);
//  * RAND55, the additive lagged-Fibonacci generator of MyRand (MyRand.h:81-108,
//    MyRand.cc:30-130), the generator of the reference's shipped build
//    (Makefile:2 -DUSE_RAND55), with its Box-Muller pairs and the su(N)
//    noise mapping of MyMath.cc:256-292, 326-346.
//
// ranlxd.
// State: 12 vectors, each a 48-bit word per lane held as (low, high) 24-bit
// halves, for 4 interleaved lanes; one update performs p subtract-with-borrow
// steps x_i <- x_j - x_i - carry with lag (i, j = i + 7 mod 12).  A draw
// returns the next lane's 48-bit word as a double in [0, 1).
#include <cmath>
#include <cstring>
#include <new>

#include "../../include/nspt.h"

// RAND55: 55 words x_n = x_{n-55} + x_{n-24} (mod 2^64), read at positions
// n1 and n2 = n1 + 31 (mod 55) that both advance by one per draw; a draw
// returns x / (2^64 - 1).  The ring is seeded from 161803398 + seed by the
// reference's recurrence (a wrapping 64-bit sum; the reference's signed
// long overflows there) and 220 draws are discarded.  Normal deviates come
// in Box-Muller pairs with the reference's truncated 2 pi (MyRand.h:40):
// t = 2 pi u1, r = sqrt(-ln(1 - u2)), first r cos t, then r sin t.
struct Rand55 {
  unsigned long long x[55];
  int n1 = 0, n2 = 31;
  bool have_sin = false;
  double r = 0.0, t = 0.0;

  explicit Rand55(long seed) {
    unsigned long long prev = 161803398ULL + (unsigned long long)seed, acc = 1;
    x[54] = prev;
    for (int i = 0, pos = 20; i < 54; i++, pos = (pos + 21) % 55) {
      x[pos] = acc;
      acc += prev;
      prev = x[pos];
    }
    for (int i = 0; i < 220; i++) uniform();
  }
  double uniform() {
    n1 = (n1 == 54) ? 0 : n1 + 1;
    n2 = (n2 == 54) ? 0 : n2 + 1;
    x[n1] += x[n2];
    return (double)x[n1] * (1.0 / (double)~0ULL);
  }
  double normal() {
    if (have_sin) {
      have_sin = false;
      return r * std::sin(t);
    }
    t = 6.28318530717959 * uniform();
    r = std::sqrt(-std::log(1.0 - uniform()));
    have_sin = true;
    return r * std::cos(t);
  }
  // i sum_a g_a lambda_a in the reference's draw order: SU(3) entries
  // (0,1)/(1,0), (0,2)/(2,0), (1,2)/(2,1), then the diagonal from two
  // deviates (the second times sqrt(1/3)); SU(2) (0,1)/(1,0), then the
  // diagonal.  Interleaved complex, row-major.
  void noise(int nc, double* o) {
    for (int q = 0; q < 2 * nc * nc; q++) o[q] = 0.0;
    auto pair = [&](int a, int b) {  // element (a) = g1 i + g2, (b) = g1 i - g2
      const double g1 = normal(), g2 = normal();
      o[2 * a + 1] = g1;
      o[2 * b + 1] = g1;
      o[2 * a] = g2;
      o[2 * b] = -g2;
    };
    if (nc == 3) {
      pair(1, 3);
      pair(2, 6);
      pair(5, 7);
      const double g = normal(), gg = normal() * std::sqrt(1.0 / 3.0);
      o[1] = gg + g;
      o[9] = gg - g;
      o[17] = -2.0 * gg;
    } else {
      pair(1, 2);
      const double g = normal();
      o[1] = g;
      o[7] = -g;
    }
  }
};

struct nspt_rng {
  Rand55* r55 = nullptr;  // RAND55 source when set, ranlxd otherwise
  int lo[12][4], hi[12][4];
  int carry[4];
  int p, pmod, ir, jr;
  int pos;     // lanes read since the last update (48 per update)
  bool first;  // no update done yet

  ~nspt_rng() { delete r55; }
  nspt_rng(int seed, int level) {
    p = (level == 2) ? 397 : 202;
    pmod = p % 12;
    // bit generator of the seed (31-bit shift register, taps 0 and 18)
    int bits[31];
    unsigned s = (unsigned)seed;
    for (int k = 0; k < 31; k++, s >>= 1) bits[k] = (int)(s & 1u);
    int b0 = 0, b1 = 18;
    int word[96];
    for (int lane = 0; lane < 4; lane++)
      for (int k = 0; k < 24; k++) {
        int v = 0;
        for (int l = 0; l < 24; l++) {
          v = (v << 1) | bits[b0];
          bits[b0] ^= bits[b1];
          b0 = (b0 + 1) % 31;
          b1 = (b1 + 1) % 31;
        }
        if (k % 4 != lane) v = 0xffffff - v;
        word[4 * k + lane] = v;
      }
    // word index 4k+lane: vector k/2, half k%2 (0 = low, 1 = high)
    for (int w = 0; w < 96; w++) {
      const int lane = w % 4, k = w / 4;
      (k % 2 ? hi : lo)[k / 2][lane] = word[w];
    }
    for (int l = 0; l < 4; l++) carry[l] = 0;
    ir = 0;
    jr = 7;
    pos = 0;
    first = true;
  }

  void update() {
    for (int k = 0, i = ir, j = jr; k < p; k++) {
      for (int l = 0; l < 4; l++) {
        int d = lo[j][l] - lo[i][l] - carry[l];
        hi[i][l] += (d < 0);
        lo[i][l] = (d + 0x1000000) & 0xffffff;
      }
      for (int l = 0; l < 4; l++) {
        int d = hi[j][l] - hi[i][l];
        carry[l] = (d < 0);
        hi[i][l] = (d + 0x1000000) & 0xffffff;
      }
      i = (i + 1) % 12;
      j = (j + 1) % 12;
    }
    ir = (ir + pmod) % 12;
    jr = (jr + pmod) % 12;
  }

  // After each update the 48 words are read vector by vector, lanes 0..3,
  // starting at vector ir; the first draw triggers the first update.
  double next() {
    if (first) {
      first = false;
      update();
      pos = 0;
    } else if (++pos == 48) {
      update();
      pos = 0;
    }
    const int v = ((ir + 12) + pos / 4) % 12;
    const int l = pos % 4;
    const double one_bit = std::ldexp(1.0, -24);
    return one_bit * ((double)hi[v][l] + one_bit * (double)lo[v][l]);
  }
};

extern "C" {

int nspt_rng_create_ranlxd(int seed, int level, nspt_rng** out) {
  if (!out) return -1;
  *out = nullptr;
  if (seed <= 0 || (level != 1 && level != 2)) return -2;
  *out = new (std::nothrow) nspt_rng(seed, level);
  return *out ? 0 : -3;
}

int nspt_rng_create_rand55(long seed, nspt_rng** out) {
  if (!out) return -1;
  *out = new (std::nothrow) nspt_rng(1, 1);
  if (!*out) return -3;
  (*out)->r55 = new (std::nothrow) Rand55(seed);
  if (!(*out)->r55) {
    delete *out;
    *out = nullptr;
    return -3;
  }
  return 0;
}

int nspt_rng_destroy(nspt_rng* r) {
  delete r;
  return 0;
}

int nspt_rng_uniform(nspt_rng* r, int n, double* out) {
  if (!r || !out || n < 0) return -1;
  for (int i = 0; i < n; i++) out[i] = r->r55 ? r->r55->uniform() : r->next();
  return 0;
}

int nspt_rng_noise(nspt_rng* r, int nc, int nlinks, double* eta) {
  if (!r || !eta || nlinks < 0 || (nc != 2 && nc != 3)) return -1;
  if (r->r55) {
    for (int link = 0; link < nlinks; link++) r->r55->noise(nc, eta + (size_t)link * 2 * nc * nc);
    return 0;
  }
  const double twopi = std::atan(1.0) * 8.0;
  const int ng = (nc == 3) ? 8 : 4;
  for (int link = 0; link < nlinks; link++) {
    double g[8];
    for (int i = 0; i < ng; i++) g[i] = r->next();
    for (int i = 0; i < ng; i += 2) {
      const double t = twopi * g[i];
      const double rr = std::sqrt(-std::log(1.0 - g[i + 1]));
      g[i] = rr * std::cos(t);
      g[i + 1] = rr * std::sin(t);
    }
    double* o = eta + (size_t)link * 2 * nc * nc;
    if (nc == 3) {
      g[7] *= std::sqrt(1.0 / 3.0);
      const double m[18] = {0, g[7] + g[6], g[1], g[0], g[3], g[2], -g[1], g[0], 0, g[7] - g[6],
                            g[5], g[4], -g[3], g[2], -g[5], g[4], 0, -g[7] * 2};
      std::memcpy(o, m, sizeof(m));
    } else {
      const double m[8] = {0, g[2], g[1], g[0], -g[1], g[0], 0, -g[2]};
      std::memcpy(o, m, sizeof(m));
    }
  }
  return 0;
}

}  // extern "C"
