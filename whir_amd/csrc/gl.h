// Goldilocks field p = 2^64 - 2^32 + 1 and its degree-2/3 extensions, for
// CDNA4 device code and the host.
//
// Arithmetic spec: src/algebra/fields.rs:85-143 (Field64, Field64_2 = Fp[u]/(u^2-7),
// Field64_3 = Fp[u]/(u^3-2)).  Values are always kept canonical (< p), so two
// equal field elements are equal bit patterns; the reference's Montgomery form is
// a storage detail that its encoder removes (matrix_commit.rs:143-151).
//
// Reduction of a 128-bit product uses 2^64 = 2^32 - 1 and 2^96 = -1 (mod p):
//   lo + 2^64 (hh 2^32 + hl) = lo - hh + hl (2^32 - 1).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GL_FN __host__ __device__ __forceinline__
#else
#define GL_FN static inline
#endif

// GL_CARRY selects the formulations written for the GPU (branch-free selects,
// explicit carry chains, a 128-bit product from 32x32->64 partial products).
// Device code always uses them; a host build can select them with
// GL_HOST_CARRY_TEST to check them against 128-bit integer arithmetic on the CPU.
#if defined(__HIP_DEVICE_COMPILE__) || defined(GL_HOST_CARRY_TEST)
#define GL_CARRY 1
#endif

namespace gl {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

GL_FN uint64_t umulhi(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

GL_FN uint64_t reduce128(uint64_t lo, uint64_t hi) {
#if defined(GL_CARRY)
  // Same steps with explicit carry chains, so every correction is a masked
  // subtract or add instead of a 64-bit compare and two selects.
  unsigned long c;
  const uint64_t hh = hi >> 32;
  const uint32_t hl = (uint32_t)hi;
  uint64_t t0 = __builtin_subcl(lo, hh, 0, &c);                // lo - hh, c = borrow
  t0 -= (uint64_t)(0u - (uint32_t)c);                          // borrow: subtract 2^32 - 1 (add p)
  // hl * (2^32 - 1) as one v_mad_u64_u32: the shift-and-subtract form is two
  // full-rate ops plus a carry hazard and measured no faster on MI355X
  const uint64_t t1 = (uint64_t)hl * 0xFFFFFFFFu;
  uint64_t t2 = __builtin_addcl(t0, t1, 0, &c);
  t2 += (uint64_t)(0u - (uint32_t)c);                          // carry: 2^64 = 2^32 - 1
  uint64_t t3 = __builtin_addcl(t2, EPS, 0, &c);               // t2 + 2^32 - 1 wraps iff t2 >= p
  return c ? t3 : t2;
#else
  uint64_t hh = hi >> 32;
  uint64_t hl = hi & EPS;
  uint64_t t0 = lo - hh;
  if (lo < hh) t0 -= EPS;               // borrow: add p
  uint64_t t1 = (hl << 32) - hl;        // hl * (2^32 - 1)
  uint64_t t2 = t0 + t1;
  if (t2 < t1) t2 += EPS;               // carry: 2^64 = EPS
  if (t2 >= P) t2 -= P;
  return t2;
#endif
}

GL_FN uint64_t add(uint64_t a, uint64_t b) {
#if defined(GL_CARRY)
  // Branch-free: one correction covers both cases.  If a + b wrapped, the sum is
  // s + 2^64 = s + EPS (mod p) and s + EPS < p; if s >= p, s - p = s + EPS mod 2^64.
  // (A 32-bit carry-chain form of the same tests measured no faster on MI355X.)
  const uint64_t s = a + b, t = s + EPS;
  return ((s < a) | (t < s)) ? t : s;
#else
  uint64_t s = a + b;
  if (s < a) return s + EPS;            // wrapped past 2^64
  return s >= P ? s - P : s;
#endif
}
GL_FN uint64_t sub(uint64_t a, uint64_t b) {
#if defined(GL_CARRY)
  // The borrow of the 64-bit subtraction selects d + p; taking it from the
  // carry chain saves the separate compare.
  uint32_t c0, c1;
  const uint32_t lo = __builtin_subc((uint32_t)a, (uint32_t)b, 0u, &c0);
  const uint32_t hi = __builtin_subc((uint32_t)(a >> 32), (uint32_t)(b >> 32), c0, &c1);
  const uint64_t d = ((uint64_t)hi << 32) | lo;
  return c1 ? d + P : d;
#else
  uint64_t d = a - b;
  return a < b ? d - EPS : d;           // d + p mod 2^64
#endif
}
GL_FN uint64_t neg(uint64_t a) { return a ? P - a : 0; }

// 64x64 -> 128-bit product from four 32x32 -> 64 partial products, each carrying
// the previous one's spill as its addend (v_mad_u64_u32 on CDNA): no partial sum
// overflows 64 bits, and the low product is not recomputed.
GL_FN void mul128(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
#if defined(GL_CARRY)
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t p00 = (uint64_t)a0 * b0;
  const uint64_t mid = (uint64_t)a0 * b1 + (p00 >> 32);           // <= 2^64 - 2^32
  const uint64_t mid2 = (uint64_t)a1 * b0 + (uint32_t)mid;         // <  2^64
  hi = (uint64_t)a1 * b1 + (mid >> 32) + (mid2 >> 32);              // <= 2^64 - 1
  lo = (mid2 << 32) | (uint32_t)p00;
#else
  unsigned __int128 p = (unsigned __int128)a * b;
  lo = (uint64_t)p;
  hi = (uint64_t)(p >> 64);
#endif
}

GL_FN uint64_t mul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul128(a, b, lo, hi);
  return reduce128(lo, hi);
}
GL_FN uint64_t sqr(uint64_t a) { return mul(a, a); }

// x * 2^k for 0 <= k < 192 (2 has order 192: 2^96 = -1).
GL_FN uint64_t mul_pow2(uint64_t x, unsigned k) {
  bool negate = false;
  if (k >= 96) { k -= 96; negate = true; }
  uint64_t r;
  if (k == 0) {
    r = x;
  } else if (k < 64) {
    r = reduce128(x << k, x >> (64 - k));
  } else {
    // x * 2^64 * 2^(k-64) = (x * 2^(k-64)) * (2^32 - 1)
    uint64_t y = (k == 64) ? x : reduce128(x << (k - 64), x >> (128 - k));
    uint64_t lo = (y << 32) - y;                 // low word of y*(2^32-1)
    uint64_t hi = (y >> 32) - ((y << 32) < y);   // high word (borrow from low)
    r = reduce128(lo, hi);
  }
  return negate ? neg(r) : r;
}

GL_FN uint64_t pow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = mul(r, a);
    a = mul(a, a);
    e >>= 1;
  }
  return r;
}
GL_FN uint64_t inv(uint64_t a) { return pow(a, P - 2); }

// Extension element of degree D (1, 2, 3); lowest coefficient first.
template <int D>
struct Ext {
  uint64_t c[D];
};

template <int D>
GL_FN Ext<D> ext_zero() {
  Ext<D> r;
#pragma unroll
  for (int i = 0; i < D; i++) r.c[i] = 0;
  return r;
}
template <int D>
GL_FN Ext<D> ext_base(uint64_t x) {
  Ext<D> r = ext_zero<D>();
  r.c[0] = x;
  return r;
}
template <int D>
GL_FN Ext<D> eadd(const Ext<D>& a, const Ext<D>& b) {
  Ext<D> r;
#pragma unroll
  for (int i = 0; i < D; i++) r.c[i] = add(a.c[i], b.c[i]);
  return r;
}
template <int D>
GL_FN Ext<D> esub(const Ext<D>& a, const Ext<D>& b) {
  Ext<D> r;
#pragma unroll
  for (int i = 0; i < D; i++) r.c[i] = sub(a.c[i], b.c[i]);
  return r;
}
template <int D>
GL_FN Ext<D> escale(const Ext<D>& a, uint64_t s) {
  Ext<D> r;
#pragma unroll
  for (int i = 0; i < D; i++) r.c[i] = mul(a.c[i], s);
  return r;
}
template <int D>
GL_FN bool eeq(const Ext<D>& a, const Ext<D>& b) {
  bool r = true;
#pragma unroll
  for (int i = 0; i < D; i++) r = r && a.c[i] == b.c[i];
  return r;
}

GL_FN uint64_t mul7(uint64_t x) { return sub(mul_pow2(x, 3), x); }
GL_FN uint64_t dbl(uint64_t x) { return add(x, x); }

template <int D>
GL_FN Ext<D> emul(const Ext<D>& a, const Ext<D>& b);

template <>
GL_FN Ext<1> emul<1>(const Ext<1>& a, const Ext<1>& b) {
  Ext<1> r;
  r.c[0] = mul(a.c[0], b.c[0]);
  return r;
}
// Sums of 128-bit products, reduced once: value = lo + hi 2^64 + top 2^128, where top
// counts the carries out of hi (a handful of products, so top stays small).  With
// 2^64 = 2^32 - 1 and 2^128 = -2^32 (mod p), the value is reduce128(lo, hi) - top 2^32,
// and top 2^32 < p.  An extension product then costs one reduction per output
// coefficient instead of one per partial product.
struct Wide {
  uint64_t lo, hi;
  uint32_t top;
};

GL_FN Wide wide_prod(uint64_t a, uint64_t b) {
  Wide w;
  mul128(a, b, w.lo, w.hi);
  w.top = 0;
  return w;
}
// The GPU forms below add and subtract 32 bits at a time through the carry flag
// (v_add_co_u32 / v_addc_co_u32, full rate): a 64-bit __builtin_addcl compiles to a
// half-rate 64-bit add and a half-rate 64-bit compare per word instead.
GL_FN void wide_add(Wide& w, const Wide& x) {
#if defined(GL_CARRY)
  uint32_t c;
  const uint32_t w0 = __builtin_addc((uint32_t)w.lo, (uint32_t)x.lo, 0u, &c);
  const uint32_t w1 = __builtin_addc((uint32_t)(w.lo >> 32), (uint32_t)(x.lo >> 32), c, &c);
  const uint32_t w2 = __builtin_addc((uint32_t)w.hi, (uint32_t)x.hi, c, &c);
  const uint32_t w3 = __builtin_addc((uint32_t)(w.hi >> 32), (uint32_t)(x.hi >> 32), c, &c);
  w.top = __builtin_addc(w.top, x.top, c, &c);
  w.lo = ((uint64_t)w1 << 32) | w0;
  w.hi = ((uint64_t)w3 << 32) | w2;
#else
  const unsigned __int128 s = (((unsigned __int128)w.hi << 64) | w.lo) + (((unsigned __int128)x.hi << 64) | x.lo);
  const bool c = (uint64_t)(s >> 64) < w.hi || ((uint64_t)(s >> 64) == w.hi && (uint64_t)s < w.lo);
  w.lo = (uint64_t)s;
  w.hi = (uint64_t)(s >> 64);
  w.top += x.top + (c ? 1u : 0u);
#endif
}
// w -= x, for w >= x as integers.
GL_FN void wide_sub(Wide& w, const Wide& x) {
#if defined(GL_CARRY)
  uint32_t b;
  const uint32_t w0 = __builtin_subc((uint32_t)w.lo, (uint32_t)x.lo, 0u, &b);
  const uint32_t w1 = __builtin_subc((uint32_t)(w.lo >> 32), (uint32_t)(x.lo >> 32), b, &b);
  const uint32_t w2 = __builtin_subc((uint32_t)w.hi, (uint32_t)x.hi, b, &b);
  const uint32_t w3 = __builtin_subc((uint32_t)(w.hi >> 32), (uint32_t)(x.hi >> 32), b, &b);
  w.top = __builtin_subc(w.top, x.top, b, &b);
  w.lo = ((uint64_t)w1 << 32) | w0;
  w.hi = ((uint64_t)w3 << 32) | w2;
#else
  const unsigned __int128 a = ((unsigned __int128)w.hi << 64) | w.lo, b = ((unsigned __int128)x.hi << 64) | x.lo;
  w.top -= x.top + (a < b ? 1u : 0u);
  const unsigned __int128 d = a - b;
  w.lo = (uint64_t)d;
  w.hi = (uint64_t)(d >> 64);
#endif
}
// w += a b.  (Forming the product with four independent v_mad_u64_u32 and adding its
// halves in two carry chains, instead of the addend-chained product, measured slower
// on MI355X: more registers and condition-register hazards.)
GL_FN void wide_mac(Wide& w, uint64_t a, uint64_t b) { wide_add(w, wide_prod(a, b)); }
// w <<= k for 1 <= k <= 3.
GL_FN void wide_shl(Wide& w, unsigned k) {
  w.top = (w.top << k) | (uint32_t)(w.hi >> (64 - k));
  w.hi = (w.hi << k) | (w.lo >> (64 - k));
  w.lo <<= k;
}
GL_FN uint64_t wide_reduce(const Wide& w) { return sub(reduce128(w.lo, w.hi), (uint64_t)w.top << 32); }

// Fp2: (a0 + a1 u)(b0 + b1 u), u^2 = 7: c0 = a0 b0 + 8 a1 b1 - a1 b1, c1 = a0 b1 + a1 b0,
// each a sum of 128-bit products reduced once.
template <>
GL_FN Ext<2> emul<2>(const Ext<2>& a, const Ext<2>& b) {
  Wide c0 = wide_prod(a.c[0], b.c[0]);
  const Wide t = wide_prod(a.c[1], b.c[1]);
  Wide t8 = t;
  wide_shl(t8, 3);
  wide_add(c0, t8);
  wide_sub(c0, t);
  Wide c1 = wide_prod(a.c[0], b.c[1]);
  wide_mac(c1, a.c[1], b.c[0]);
  Ext<2> r;
  r.c[0] = wide_reduce(c0);
  r.c[1] = wide_reduce(c1);
  return r;
}
// Fp3: u^3 = 2: c0 = a0 b0 + 2 (a1 b2 + a2 b1), c1 = a0 b1 + a1 b0 + 2 a2 b2,
// c2 = a0 b2 + a1 b1 + a2 b0, each reduced once.
template <>
GL_FN Ext<3> emul<3>(const Ext<3>& a, const Ext<3>& b) {
  Wide c0 = wide_prod(a.c[1], b.c[2]);
  wide_mac(c0, a.c[2], b.c[1]);
  wide_shl(c0, 1);
  wide_mac(c0, a.c[0], b.c[0]);
  Wide c1 = wide_prod(a.c[2], b.c[2]);
  wide_shl(c1, 1);
  wide_mac(c1, a.c[0], b.c[1]);
  wide_mac(c1, a.c[1], b.c[0]);
  Wide c2 = wide_prod(a.c[0], b.c[2]);
  wide_mac(c2, a.c[1], b.c[1]);
  wide_mac(c2, a.c[2], b.c[0]);
  Ext<3> r;
  r.c[0] = wide_reduce(c0);
  r.c[1] = wide_reduce(c1);
  r.c[2] = wide_reduce(c2);
  return r;
}

// Sums of extension products without a reduction per term: w[d] += (a b)_d as an
// integer (the coefficient before reduction mod p), so a dot product reduces once per
// coefficient at the end (ext_reduce).  Each term adds less than 9 * 2^128 per
// coefficient, so the 32-bit carry word holds 2^28 terms.
template <int D>
GL_FN void emul_acc(Wide (&w)[D], const Ext<D>& a, const Ext<D>& b);
template <>
GL_FN void emul_acc<1>(Wide (&w)[1], const Ext<1>& a, const Ext<1>& b) {
  wide_mac(w[0], a.c[0], b.c[0]);
}
// c0 += a0 b0 + 7 a1 b1 (as + 8 t - t, w >= 8t >= t), c1 += a0 b1 + a1 b0
template <>
GL_FN void emul_acc<2>(Wide (&w)[2], const Ext<2>& a, const Ext<2>& b) {
  const Wide t = wide_prod(a.c[1], b.c[1]);
  Wide t8 = t;
  wide_shl(t8, 3);
  wide_mac(w[0], a.c[0], b.c[0]);
  wide_add(w[0], t8);
  wide_sub(w[0], t);
  wide_mac(w[1], a.c[0], b.c[1]);
  wide_mac(w[1], a.c[1], b.c[0]);
}
// c0 += a0 b0 + 2 (a1 b2 + a2 b1), c1 += a0 b1 + a1 b0 + 2 a2 b2, c2 += a0 b2 + a1 b1 + a2 b0
template <>
GL_FN void emul_acc<3>(Wide (&w)[3], const Ext<3>& a, const Ext<3>& b) {
  Wide t = wide_prod(a.c[1], b.c[2]);
  wide_mac(t, a.c[2], b.c[1]);
  wide_shl(t, 1);
  wide_add(w[0], t);
  wide_mac(w[0], a.c[0], b.c[0]);
  Wide u = wide_prod(a.c[2], b.c[2]);
  wide_shl(u, 1);
  wide_add(w[1], u);
  wide_mac(w[1], a.c[0], b.c[1]);
  wide_mac(w[1], a.c[1], b.c[0]);
  wide_mac(w[2], a.c[0], b.c[2]);
  wide_mac(w[2], a.c[1], b.c[1]);
  wide_mac(w[2], a.c[2], b.c[0]);
}
template <int D>
GL_FN void wide_zero(Wide (&w)[D]) {
  for (int d = 0; d < D; d++) w[d] = Wide{0, 0, 0};
}
template <int D>
GL_FN Ext<D> ext_reduce(const Wide (&w)[D]) {
  Ext<D> r;
  for (int d = 0; d < D; d++) r.c[d] = wide_reduce(w[d]);
  return r;
}

template <int D>
GL_FN Ext<D> epow(Ext<D> a, uint64_t e) {
  Ext<D> r = ext_base<D>(1);
  while (e) {
    if (e & 1) r = emul<D>(r, a);
    a = emul<D>(a, a);
    e >>= 1;
  }
  return r;
}

// Roots of unity of the reference's NTT engine (cooley_tukey.rs:53-79):
// omega_order = 7^((p-1)/order), order = 3 * 2^32; root(n) = omega_order^(order/n).
constexpr uint64_t NTT_ORDER = 3ull << 32;
GL_FN uint64_t root_of_unity(uint64_t n) { return pow(pow(7, (P - 1) / NTT_ORDER), NTT_ORDER / n); }

}  // namespace gl
