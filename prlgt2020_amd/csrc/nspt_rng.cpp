// nspt_rng.cpp -- host noise source of the engine: Luescher's RANLUX
// double-precision generator "ranlxd" (portable formulation,
// ranlxd/ranlxd.cc:332-531; published in M. Luescher, "Algorithms used in
// ranlux v3.0", 2001) and the su(N) noise mapping of MyMath.cc:295-322.
//
// State: 12 vectors, each a 48-bit word per lane held as (low, high) 24-bit
// halves, for 4 interleaved lanes; one update performs p subtract-with-borrow
// steps x_i <- x_j - x_i - carry with lag (i, j = i + 7 mod 12).  A draw
// returns the next lane's 48-bit word as a double in [0, 1).
#include <cmath>
#include <cstring>
#include <new>

#include "../../include/nspt.h"

struct nspt_rng {
  int lo[12][4], hi[12][4];
  int carry[4];
  int p, pmod, ir, jr;
  int pos;     // lanes read since the last update (48 per update)
  bool first;  // no update done yet

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

int nspt_rng_destroy(nspt_rng* r) {
  delete r;
  return 0;
}

int nspt_rng_uniform(nspt_rng* r, int n, double* out) {
  if (!r || !out || n < 0) return -1;
  for (int i = 0; i < n; i++) out[i] = r->next();
  return 0;
}

int nspt_rng_noise(nspt_rng* r, int nc, int nlinks, double* eta) {
  if (!r || !eta || nlinks < 0 || (nc != 2 && nc != 3)) return -1;
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
