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
  const uint64_t t1 = (uint64_t)hl * 0xFFFFFFFFu;              // hl * (2^32 - 1)
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
// Fp2: (a0 + a1 u)(b0 + b1 u), u^2 = 7 (Karatsuba).
template <>
GL_FN Ext<2> emul<2>(const Ext<2>& a, const Ext<2>& b) {
  uint64_t v0 = mul(a.c[0], b.c[0]);
  uint64_t v1 = mul(a.c[1], b.c[1]);
  uint64_t m = mul(add(a.c[0], a.c[1]), add(b.c[0], b.c[1]));
  Ext<2> r;
  r.c[0] = add(v0, mul7(v1));
  r.c[1] = sub(sub(m, v0), v1);
  return r;
}
// Fp3: u^3 = 2.
template <>
GL_FN Ext<3> emul<3>(const Ext<3>& a, const Ext<3>& b) {
  uint64_t a0b0 = mul(a.c[0], b.c[0]), a1b1 = mul(a.c[1], b.c[1]), a2b2 = mul(a.c[2], b.c[2]);
  uint64_t a0b1 = mul(a.c[0], b.c[1]), a1b0 = mul(a.c[1], b.c[0]);
  uint64_t a0b2 = mul(a.c[0], b.c[2]), a2b0 = mul(a.c[2], b.c[0]);
  uint64_t a1b2 = mul(a.c[1], b.c[2]), a2b1 = mul(a.c[2], b.c[1]);
  Ext<3> r;
  r.c[0] = add(a0b0, dbl(add(a1b2, a2b1)));
  r.c[1] = add(add(a0b1, a1b0), dbl(a2b2));
  r.c[2] = add(add(a0b2, a2b0), a1b1);
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
