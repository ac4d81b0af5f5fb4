// CPU check of the GPU formulations of Goldilocks arithmetic (whir_amd/csrc/gl.h
// with GL_HOST_CARRY_TEST: the carry-chain add/sub, the 4-product 128-bit multiply,
// the carry-chain reduction, mul_pow2, the 128-bit-plus-carries sums of products and
// the unreduced extension dot products)
// against unsigned __int128 arithmetic,
// over every pair of edge values and a stream of pseudo-random pairs.
// Prints "ok <count>" or the first mismatch; exit status 0 iff all agree.
#define GL_HOST_CARRY_TEST 1
#include <stdio.h>
#include <stdlib.h>

#include "../whir_amd/csrc/gl.h"

typedef unsigned __int128 u128;
static const uint64_t P = 0xFFFFFFFF00000001ull;

static uint64_t r_add(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a + b) % P); }
static uint64_t r_sub(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a + P - b) % P); }
static uint64_t r_mul(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % P); }
static uint64_t r_pow2(uint64_t x, unsigned k) {
  uint64_t r = x % P;
  for (unsigned i = 0; i < k; i++) r = (uint64_t)(((u128)r * 2) % P);
  return r;
}

static uint64_t state = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {  // splitmix64
  uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static uint64_t canon(uint64_t x) { return x >= P ? x - P : x; }

static long checked = 0;
static int check(const char* op, uint64_t a, uint64_t b, uint64_t got, uint64_t want) {
  checked++;
  if (got != want) {
    printf("MISMATCH %s a=%016llx b=%016llx got=%016llx want=%016llx\n", op, (unsigned long long)a,
           (unsigned long long)b, (unsigned long long)got, (unsigned long long)want);
    return 1;
  }
  return 0;
}

// Extension products of (x0, x1[, x2]) and (y0, y1[, y2]) from per-product reductions.
static void r_emul2(const uint64_t* x, const uint64_t* y, uint64_t* z) {
  z[0] = r_add(r_mul(x[0], y[0]), r_mul(7, r_mul(x[1], y[1])));
  z[1] = r_add(r_mul(x[0], y[1]), r_mul(x[1], y[0]));
}
static void r_emul3(const uint64_t* x, const uint64_t* y, uint64_t* z) {
  z[0] = r_add(r_mul(x[0], y[0]), r_mul(2, r_add(r_mul(x[1], y[2]), r_mul(x[2], y[1]))));
  z[1] = r_add(r_add(r_mul(x[0], y[1]), r_mul(x[1], y[0])), r_mul(2, r_mul(x[2], y[2])));
  z[2] = r_add(r_add(r_mul(x[0], y[2]), r_mul(x[1], y[1])), r_mul(x[2], y[0]));
}

static int ext(const uint64_t* x, const uint64_t* y) {
  int bad = 0;
  gl::Ext<2> a2{{x[0], x[1]}}, b2{{y[0], y[1]}};
  gl::Ext<2> c2 = gl::emul<2>(a2, b2);
  uint64_t w2[2];
  r_emul2(x, y, w2);
  bad |= check("emul2.c0", x[0], y[0], c2.c[0], w2[0]);
  bad |= check("emul2.c1", x[1], y[1], c2.c[1], w2[1]);
  gl::Ext<3> a3{{x[0], x[1], x[2]}}, b3{{y[0], y[1], y[2]}};
  gl::Ext<3> c3 = gl::emul<3>(a3, b3);
  uint64_t w3[3];
  r_emul3(x, y, w3);
  bad |= check("emul3.c0", x[0], y[0], c3.c[0], w3[0]);
  bad |= check("emul3.c1", x[1], y[1], c3.c[1], w3[1]);
  bad |= check("emul3.c2", x[2], y[2], c3.c[2], w3[2]);
  return bad;
}

static int pair(uint64_t a, uint64_t b) {
  int bad = 0;
  bad |= check("add", a, b, gl::add(a, b), r_add(a, b));
  bad |= check("sub", a, b, gl::sub(a, b), r_sub(a, b));
  bad |= check("mul", a, b, gl::mul(a, b), r_mul(a, b));
  uint64_t lo, hi;
  gl::mul128(a, b, lo, hi);
  u128 prod = (u128)a * b;
  bad |= check("mul128.lo", a, b, lo, (uint64_t)prod);
  bad |= check("mul128.hi", a, b, hi, (uint64_t)(prod >> 64));
  // reduce128 of arbitrary 128-bit values (not only products of canonical inputs)
  bad |= check("reduce128", a, b, gl::reduce128(a, b), (uint64_t)((((u128)b << 64) | a) % P));
  return bad;
}

// Sums of products in the 128-bit-plus-carries accumulator (wide_mac, wide_add,
// wide_sub, wide_reduce) against an exact 192-bit sum (u128 + carry count).
struct Exact {
  u128 low;
  uint64_t top;
};
static void ex_add(Exact& e, u128 v) {
  const u128 s = e.low + v;
  e.top += s < e.low;
  e.low = s;
}
static int wide_check(const char* op, const gl::Wide& w, const Exact& e) {
  int bad = 0;
  bad |= check(op, (uint64_t)e.low, 0, w.lo, (uint64_t)e.low);
  bad |= check(op, (uint64_t)(e.low >> 64), 1, w.hi, (uint64_t)(e.low >> 64));
  bad |= check(op, e.top, 2, w.top, e.top);
  // value mod p = low + top 2^128, 2^128 = -2^32 (mod p)
  const uint64_t want = r_sub((uint64_t)(e.low % P), r_mul(e.top % P, 1ull << 32));
  bad |= check(op, e.top, 3, gl::wide_reduce(w), want);
  return bad;
}
static int wide_sums(const uint64_t* vals, int nv, long rounds) {
  int bad = 0;
  for (long r = 0; r < rounds && !bad; r++) {
    gl::Wide w{0, 0, 0};
    Exact e{0, 0};
    const int len = 1 + (int)(rnd() % 300);
    for (int i = 0; i < len && !bad; i++) {
      const uint64_t a = (rnd() & 3) ? rnd() : vals[rnd() % nv], b = (rnd() & 3) ? rnd() : vals[rnd() % nv];
      gl::wide_mac(w, a, b);
      ex_add(e, (u128)a * b);
      if ((i & 31) == 31) {  // a whole product added and taken away again
        const gl::Wide x = gl::wide_prod(a, b);
        gl::wide_add(w, x);
        gl::wide_sub(w, x);
      }
    }
    bad |= wide_check("wide_mac", w, e);
  }
  // every product of two all-ones words: each one carries out of 128 bits in turn
  gl::Wide w{0, 0, 0};
  Exact e{0, 0};
  for (int i = 0; i < 1000 && !bad; i++) {
    gl::wide_mac(w, ~0ull, ~0ull);
    ex_add(e, (u128)~0ull * ~0ull);
    if (i % 97 == 0) bad |= wide_check("wide_mac.ones", w, e);
  }
  return bad | wide_check("wide_mac.ones", w, e);
}

// Dot products of extension vectors through emul_acc / ext_reduce against sums of the
// per-product reductions, with terms of edge values (p - 1, ...) to push the carries.
static int ext_sums(const uint64_t* vals, int nv, long rounds) {
  int bad = 0;
  for (long r = 0; r < rounds && !bad; r++) {
    const int len = 1 + (int)(rnd() % 200);
    gl::Wide w2[2], w3[3];
    gl::wide_zero<2>(w2);
    gl::wide_zero<3>(w3);
    uint64_t e2[2] = {0, 0}, e3[3] = {0, 0, 0};
    for (int i = 0; i < len; i++) {
      uint64_t x[3], y[3];
      for (int d = 0; d < 3; d++) {
        x[d] = (rnd() & 1) ? canon(rnd()) : canon(vals[rnd() % nv]);
        y[d] = (rnd() & 1) ? canon(rnd()) : canon(vals[rnd() % nv]);
      }
      gl::emul_acc<2>(w2, gl::Ext<2>{{x[0], x[1]}}, gl::Ext<2>{{y[0], y[1]}});
      gl::emul_acc<3>(w3, gl::Ext<3>{{x[0], x[1], x[2]}}, gl::Ext<3>{{y[0], y[1], y[2]}});
      uint64_t z2[2], z3[3];
      r_emul2(x, y, z2);
      r_emul3(x, y, z3);
      for (int d = 0; d < 2; d++) e2[d] = r_add(e2[d], z2[d]);
      for (int d = 0; d < 3; d++) e3[d] = r_add(e3[d], z3[d]);
    }
    const gl::Ext<2> g2 = gl::ext_reduce<2>(w2);
    const gl::Ext<3> g3 = gl::ext_reduce<3>(w3);
    for (int d = 0; d < 2; d++) bad |= check("emul_acc2", len, d, g2.c[d], e2[d]);
    for (int d = 0; d < 3; d++) bad |= check("emul_acc3", len, d, g3.c[d], e3[d]);
  }
  // long runs of the largest coefficients: every term near 9 * 2^128 (Fp2, c0)
  gl::Wide w2[2];
  gl::wide_zero<2>(w2);
  uint64_t e2[2] = {0, 0};
  const uint64_t x[3] = {P - 1, P - 1, P - 1};
  for (int i = 0; i < 5000 && !bad; i++) {
    gl::emul_acc<2>(w2, gl::Ext<2>{{x[0], x[1]}}, gl::Ext<2>{{x[0], x[1]}});
    uint64_t z2[2];
    r_emul2(x, x, z2);
    for (int d = 0; d < 2; d++) e2[d] = r_add(e2[d], z2[d]);
  }
  const gl::Ext<2> g2 = gl::ext_reduce<2>(w2);
  for (int d = 0; d < 2; d++) bad |= check("emul_acc2.max", 5000, d, g2.c[d], e2[d]);
  return bad;
}

int main(int argc, char** argv) {
  long n_random = argc > 1 ? atol(argv[1]) : 2000000;
  const uint64_t E = 0xFFFFFFFFull;
  uint64_t edge[] = {0, 1, 2, 3, E - 1, E, E + 1, E + 2, 1ull << 32, (1ull << 32) + 1, 1ull << 63, (1ull << 63) - 1,
                     P - 1, P - 2, P - E, P - E - 1, P - (1ull << 32), P / 2, P / 2 + 1, 0x00000000FFFFFFFEull,
                     0xFFFFFFFE00000000ull, 0xFFFFFFFE00000001ull, 0xFFFFFFFEFFFFFFFFull,
                     0x8000000000000000ull, 0x7FFFFFFF80000000ull, 0x00000001FFFFFFFFull};
  const int ne = sizeof(edge) / sizeof(edge[0]);
  int bad = 0;
  for (int i = 0; i < ne && !bad; i++)
    for (int j = 0; j < ne && !bad; j++) bad |= pair(edge[i], edge[j]);
  // reduce128 also at the top of its input range
  uint64_t top[] = {~0ull, ~0ull - 1, 0xFFFFFFFF00000000ull, 0x00000000FFFFFFFFull, P, P + 1};
  for (int i = 0; i < 6 && !bad; i++)
    for (int j = 0; j < 6 && !bad; j++)
      bad |= check("reduce128", top[i], top[j], gl::reduce128(top[i], top[j]),
                   (uint64_t)((((u128)top[j] << 64) | top[i]) % P));
  for (int i = 0; i < ne && !bad; i++)
    for (unsigned k = 0; k < 192 && !bad; k++)
      bad |= check("mul_pow2", edge[i], k, gl::mul_pow2(canon(edge[i]), k), r_pow2(edge[i], k));
  // extension products over all triples of edge values (canonical), and random ones
  for (int i = 0; i < ne && !bad; i++)
    for (int j = 0; j < ne && !bad; j++)
      for (int k = 0; k < ne && !bad; k += 5) {
        uint64_t x[3] = {canon(edge[i]), canon(edge[j]), canon(edge[k])};
        uint64_t y[3] = {canon(edge[j]), canon(edge[k]), canon(edge[i])};
        bad |= ext(x, y);
      }
  for (long t = 0; t < n_random / 8 && !bad; t++) {
    uint64_t x[3] = {canon(rnd()), canon(rnd()), canon(rnd())}, y[3] = {canon(rnd()), canon(rnd()), canon(rnd())};
    if (t & 1) x[1] = P - 1 - (rnd() & 0xFF), y[1] = P - 1 - (rnd() & 0xFF);
    bad |= ext(x, y);
  }
  for (long t = 0; t < n_random && !bad; t++) {
    uint64_t a = canon(rnd()), b = canon(rnd());
    // mix in values near p and near the 32-bit boundaries
    if ((t & 7) == 1) a = P - 1 - (rnd() & 0xFFFF);
    if ((t & 7) == 2) b = (rnd() & 0xFFFFFFFFull);
    if ((t & 7) == 3) a = (rnd() << 32) | 0xFFFFFFFFull, a = canon(a);
    bad |= pair(a, b);
    if ((t & 63) == 0) bad |= check("mul_pow2", a, t % 192, gl::mul_pow2(a, (unsigned)(t % 192)), r_pow2(a, (unsigned)(t % 192)));
  }
  if (!bad) bad |= wide_sums(edge, ne, n_random / 400);
  if (!bad) bad |= ext_sums(edge, ne, n_random / 2000);
  if (!bad) printf("ok %ld\n", checked);
  return bad;
}
