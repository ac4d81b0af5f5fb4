// Sumcheck prover round and covector builders on CDNA4.
//
// All of these are streaming passes over HBM with exact Goldilocks arithmetic,
// so their order of summation does not change any result.  Reductions run as
// grid-stride partial sums per block (LDS tree) followed by one small block.
#include "algebra.h"
#include "common.h"
#include "gl.h"
#include "ntt.h"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

namespace whir {

using gl::Ext;

static constexpr int RT = 256;            // threads per block for streaming kernels
static constexpr unsigned MAX_BLOCKS = 2048;

template <int D>
__device__ __forceinline__ Ext<D> ld(const uint64_t* p, size_t i) {
  Ext<D> r;
#pragma unroll
  for (int d = 0; d < D; d++) r.c[d] = p[i * D + d];
  return r;
}
// f(std::integral_constant<int, 0>{}), ..., f(std::integral_constant<int, N - 1>{}):
// a loop whose index is a compile-time constant in the body, so arrays indexed by it
// stay in registers however large the body is (an unroll pragma gives up on large
// bodies, and the arrays then go to scratch memory).
template <class F, int... J>
__device__ __forceinline__ void static_for_seq(F&& f, std::integer_sequence<int, J...>) {
  (f(std::integral_constant<int, J>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
  static_for_seq(f, std::make_integer_sequence<int, N>{});
}

template <int D>
__device__ __forceinline__ void st(uint64_t* p, size_t i, const Ext<D>& v) {
#pragma unroll
  for (int d = 0; d < D; d++) p[i * D + d] = v.c[d];
}
template <int D>
__device__ __forceinline__ Ext<D> ld_host(const uint64_t* const* unused, const Ext<D>& v) { return v; }

// Sum NV extension values over the block into thread 0's v (other threads' v are left
// partial).  Each wave sums its lanes by cross-lane exchanges (no barrier), its first
// lane puts the wave's sum in LDS, and after one barrier thread 0 adds the waves' sums.
template <int D, int NV>
__device__ void block_sum(Ext<D> v[NV]) {
  __shared__ uint64_t red[RT / 64][NV * D];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
#pragma unroll
    for (int k = 0; k < NV; k++) {
      Ext<D> o;
#pragma unroll
      for (int d = 0; d < D; d++) o.c[d] = __shfl_xor(v[k].c[d], off);
      v[k] = gl::eadd<D>(v[k], o);
    }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int k = 0; k < NV; k++)
#pragma unroll
      for (int d = 0; d < D; d++) red[wave][k * D + d] = v[k].c[d];
  __syncthreads();
  if (threadIdx.x == 0) {
    const int waves = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
#pragma unroll
      for (int d = 0; d < D; d++) v[k].c[d] = red[0][k * D + d];
      for (int w = 1; w < waves; w++) {
        Ext<D> x;
#pragma unroll
        for (int d = 0; d < D; d++) x.c[d] = red[w][k * D + d];
        v[k] = gl::eadd<D>(v[k], x);
      }
    }
  }
}

// Sum NV extension values over the block; thread 0 writes them to out (and then raises
// the mailbox flag when sig asks it to).
template <int D, int NV>
__device__ void block_reduce_store(Ext<D> v[NV], uint64_t* out, MailboxSig sig = {}) {
  block_sum<D, NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++)
#pragma unroll
      for (int d = 0; d < D; d++) out[k * D + d] = v[k].c[d];
    mailbox_raise(sig);
  }
}

// Sum NV extension values over a block of RT threads through LDS: each thread stores
// its values, 16 threads per value add 16 of the stored rows each, and one thread per
// value adds those 16 partials.  The totals are in tot[0..NV) after the call (all
// threads).  For single-block reductions of many values, where the cross-lane sum of
// block_sum is a long chain of exchanges per value.
template <int D, int NV>
__device__ void block_sum_lds(const Ext<D> (&v)[NV], Ext<D> (&tot)[NV]) {
  static_assert(NV * 16 <= RT, "block_sum_lds: at most RT / 16 values");
  __shared__ uint64_t rows[RT][NV * D];
  __shared__ uint64_t part[16][NV * D];
  __shared__ uint64_t sum[NV * D];
#pragma unroll
  for (int k = 0; k < NV; k++)
#pragma unroll
    for (int d = 0; d < D; d++) rows[threadIdx.x][k * D + d] = v[k].c[d];
  __syncthreads();
  if (threadIdx.x < NV * 16) {
    const int k = threadIdx.x >> 4, p = threadIdx.x & 15;
    Ext<D> acc = gl::ext_zero<D>();
    for (int r = p; r < (int)blockDim.x; r += 16) {
      Ext<D> x;
#pragma unroll
      for (int d = 0; d < D; d++) x.c[d] = rows[r][k * D + d];
      acc = gl::eadd<D>(acc, x);
    }
#pragma unroll
    for (int d = 0; d < D; d++) part[p][k * D + d] = acc.c[d];
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    Ext<D> acc = gl::ext_zero<D>();
#pragma unroll
    for (int p = 0; p < 16; p++) {
      Ext<D> x;
#pragma unroll
      for (int d = 0; d < D; d++) x.c[d] = part[p][threadIdx.x * D + d];
      acc = gl::eadd<D>(acc, x);
    }
#pragma unroll
    for (int d = 0; d < D; d++) sum[threadIdx.x * D + d] = acc.c[d];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; k++)
#pragma unroll
    for (int d = 0; d < D; d++) tot[k].c[d] = sum[k * D + d];
}

// Final reduction of `nparts` block partials of NV*D words each into `out`
// (device memory, or the stream's pinned mailbox).  A second launch rather than a
// last-block-done reduction in the first: on CDNA4 the per-block device-scope
// release that pattern needs writes back the XCD's L2, and the producing kernels
// have just filled it with their outputs (measured 1.5x slower on MI355X).
template <int D, int NV>
__global__ __launch_bounds__(RT) void k_reduce_parts(const uint64_t* parts, size_t nparts, uint64_t* out,
                                                     MailboxSig sig) {
  Ext<D> v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = gl::ext_zero<D>();
  for (size_t p = threadIdx.x; p < nparts; p += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = gl::eadd<D>(v[k], ld<D>(parts + p * NV * D, k));
  Ext<D> tot[NV];
  block_sum_lds<D, NV>(v, tot);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; k++) st<D>(out, k, tot[k]);
    mailbox_raise(sig);
  }
}

// last: this launch ends the work a mailbox wait may be armed for.
template <int D, int NV>
static void finish_reduce(const uint64_t* parts, unsigned nparts, uint64_t* out, hipStream_t s, bool last = true) {
  k_reduce_parts<D, NV><<<1, RT, 0, s>>>(parts, nparts, out, last ? mailbox_take(out) : MailboxSig{});
  WHIR_HIP(hipGetLastError());
}

static unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>(blocks_for(n, RT), MAX_BLOCKS); }

// Up to this many element pairs a reducing pass runs as one block (four pairs per
// lane), which stores its sums directly: cheaper than a second launch (256 and 64
// measured no faster at d = 24 on MI355X).
static constexpr size_t SMALL_PAIRS = 1024;

// Pass a small host constant (D words) to kernels by value.
template <int D>
static Ext<D> host_ext(const uint64_t* w) {
  Ext<D> r;
  for (int d = 0; d < D; d++) r.c[d] = w[d];
  return r;
}

static size_t next_pow2(size_t n) {
  size_t p = 1;
  while (p < n) p <<= 1;
  return p;
}

// ------------------------------------------------------------------ sumcheck polynomial
// Element i of a vector of Ds-word elements, embedded in the degree-D extension
// (Ds = 1: a base-field vector, as the initial round sees the committed vector).
template <int Ds, int D>
__device__ __forceinline__ Ext<D> ldx(const uint64_t* p, size_t i) {
  if constexpr (Ds == D) return ld<D>(p, i);
  else return gl::ext_base<D>(p[i]);
}

// a * b with a embedded from Ds words.
template <int Ds, int D>
__device__ __forceinline__ Ext<D> mulx(const uint64_t* a, size_t i, const Ext<D>& b) {
  if constexpr (Ds == D) return gl::emul<D>(ld<D>(a, i), b);
  else return gl::escale<D>(b, a[i]);
}

template <int Ds, int D>
__global__ __launch_bounds__(RT) void k_sc_poly(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                                size_t half, size_t nhi, uint64_t* parts, MailboxSig sig) {
  Ext<D> acc[2] = {gl::ext_zero<D>(), gl::ext_zero<D>()};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < half; i += (size_t)gridDim.x * blockDim.x) {
    const Ext<D> b0 = ld<D>(b, i);
    const Ext<D> p0 = mulx<Ds, D>(a, i, b0);
    acc[0] = gl::eadd<D>(acc[0], p0);
    if (i < nhi) {
      const Ext<D> db = gl::esub<D>(ld<D>(b, half + i), b0);
      if constexpr (Ds == D) {
        acc[1] = gl::eadd<D>(acc[1], gl::emul<D>(gl::esub<D>(ld<D>(a, half + i), ld<D>(a, i)), db));
      } else {
        acc[1] = gl::eadd<D>(acc[1], gl::escale<D>(db, gl::sub(a[half + i], a[i])));
      }
    } else {
      acc[1] = gl::eadd<D>(acc[1], p0);  // implicit zero padding: (0 - a)(0 - b) = a b
    }
  }
  block_reduce_store<D, 2>(acc, parts + (size_t)blockIdx.x * 2 * D, sig);
}

template <int Ds, int D>
__global__ void k_sc_single(const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out, MailboxSig sig) {
  if (threadIdx.x != 0) return;
  Ext<D> c0 = n ? mulx<Ds, D>(a, 0, ld<D>(b, 0)) : gl::ext_zero<D>();
  st<D>(out, 0, c0);
  st<D>(out, 1, gl::ext_zero<D>());
  mailbox_raise(sig);
}

template <int Ds, int D>
static void sc_poly(const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out, hipStream_t s) {
  if (n <= 1) {
    k_sc_single<Ds, D><<<1, 64, 0, s>>>(a, b, n, out, mailbox_take(out));
    WHIR_HIP(hipGetLastError());
    return;
  }
  size_t half = next_pow2(n) >> 1, nhi = n - half;
  // Small rounds: one block, which then writes the sums itself (no second launch).
  const unsigned g = half <= SMALL_PAIRS ? 1u : grid_for(half);
  DevBuf parts(g == 1 ? 0 : (size_t)g * 2 * D * 8, s);
  {
    ProfScope prof(kernel_name("k_sc_poly", Ds, D), s, 8.0 * (Ds + D) * (double)n);
    k_sc_poly<Ds, D><<<g, RT, 0, s>>>(a, b, half, nhi, g == 1 ? out : parts.as<uint64_t>(),
                                      g == 1 ? mailbox_take(out) : MailboxSig{});
    WHIR_HIP(hipGetLastError());
  }
  if (g > 1) finish_reduce<D, 2>(parts.as<uint64_t>(), g, out, s);
}

void sumcheck_poly_mixed_dev(int Ds, int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out,
                             hipStream_t s) {
  if (Ds == D) {
    switch (D) {
      case 1: sc_poly<1, 1>(a, b, n, out, s); return;
      case 2: sc_poly<2, 2>(a, b, n, out, s); return;
      case 3: sc_poly<3, 3>(a, b, n, out, s); return;
    }
  } else if (Ds == 1) {
    switch (D) {
      case 2: sc_poly<1, 2>(a, b, n, out, s); return;
      case 3: sc_poly<1, 3>(a, b, n, out, s); return;
    }
  }
  throw Error("sumcheck: extension degree must be 1, 2 or 3");
}

void sumcheck_poly_dev(int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out, hipStream_t s) {
  sumcheck_poly_mixed_dev(D, D, a, b, n, out, s);
}

// ------------------------------------------------------------------ fold
// Folded element i of v (pairs (i, half + i), zero padded past n):
//   v[i] + w (v[half + i] - v[i])  or  v[i] (1 - w) when the partner is padding.
template <int Ds, int D>
__device__ __forceinline__ Ext<D> fold_at(const uint64_t* v, size_t i, size_t half, size_t nhi, const Ext<D>& w,
                                          const Ext<D>& om) {
  if constexpr (Ds == D) {
    const Ext<D> lo = ld<D>(v, i);
    if (i < nhi) return gl::eadd<D>(lo, gl::emul<D>(gl::esub<D>(ld<D>(v, half + i), lo), w));
    return gl::emul<D>(lo, om);
  } else {
    const uint64_t lo = v[i];
    if (i < nhi) {
      Ext<D> r = gl::escale<D>(w, gl::sub(v[half + i], lo));
      r.c[0] = gl::add(r.c[0], lo);
      return r;
    }
    return gl::escale<D>(om, lo);
  }
}

// out may alias in when Ds == D: lane i reads positions i and half + i and writes i only.
template <int Ds, int D>
__global__ __launch_bounds__(RT) void k_fold(const uint64_t* in, uint64_t* out, size_t half, size_t nhi, Ext<D> w,
                                             Ext<D> one_minus_w) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < half; i += (size_t)gridDim.x * blockDim.x)
    st<D>(out, i, fold_at<Ds, D>(in, i, half, nhi, w, one_minus_w));
}

template <int Ds, int D>
static size_t fold_t(const uint64_t* in, uint64_t* out, size_t n, const uint64_t* w, hipStream_t s) {
  if (n <= 1) {
    if (n == 1 && Ds != D) lift_dev(D, in, 1, out, s);  // nothing to fold: embed
    if (n == 1 && Ds == D && in != out) WHIR_HIP(hipMemcpyAsync(out, in, D * 8, hipMemcpyDeviceToDevice, s));
    return n;
  }
  size_t half = next_pow2(n) >> 1, nhi = n - half;
  Ext<D> ww = host_ext<D>(w);
  Ext<D> om = gl::esub<D>(gl::ext_base<D>(1), ww);
  ProfScope prof(kernel_name("k_fold", Ds, D), s, 8.0 * ((double)Ds * n + (double)D * half));
  k_fold<Ds, D><<<grid_for(half), RT, 0, s>>>(in, out, half, nhi, ww, om);
  WHIR_HIP(hipGetLastError());
  return half;
}

size_t fold_into_dev(int Ds, int D, const uint64_t* in, uint64_t* out, size_t n, const uint64_t* w, hipStream_t s) {
  if (Ds == D) {
    switch (D) {
      case 1: return fold_t<1, 1>(in, out, n, w, s);
      case 2: return fold_t<2, 2>(in, out, n, w, s);
      case 3: return fold_t<3, 3>(in, out, n, w, s);
    }
  } else if (Ds == 1) {
    switch (D) {
      case 2: return fold_t<1, 2>(in, out, n, w, s);
      case 3: return fold_t<1, 3>(in, out, n, w, s);
    }
  }
  throw Error("fold: extension degree must be 1, 2 or 3");
}

size_t fold_dev(int D, uint64_t* v, size_t n, const uint64_t* w, hipStream_t s) {
  return fold_into_dev(D, D, v, v, n, w, s);
}

// Both sumcheck vectors folded by the same challenge in one launch (a from Ds words
// into a_out, b in place).
template <int Ds, int D>
__global__ __launch_bounds__(RT) void k_fold2(const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t half,
                                              size_t nhi, Ext<D> w, Ext<D> one_minus_w) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < half; i += (size_t)gridDim.x * blockDim.x) {
    st<D>(a_out, i, fold_at<Ds, D>(a_in, i, half, nhi, w, one_minus_w));
    st<D>(b, i, fold_at<D, D>(b, i, half, nhi, w, one_minus_w));
  }
}

template <int Ds, int D>
static size_t fold2_t(const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t n, const uint64_t* w, hipStream_t s) {
  if (n <= 1) {
    fold_t<Ds, D>(a_in, a_out, n, w, s);
    return n;
  }
  size_t half = next_pow2(n) >> 1, nhi = n - half;
  Ext<D> ww = host_ext<D>(w);
  Ext<D> om = gl::esub<D>(gl::ext_base<D>(1), ww);
  ProfScope prof(kernel_name("k_fold2", Ds, D), s, 8.0 * ((double)(Ds + D) * n + 2.0 * D * half));
  k_fold2<Ds, D><<<grid_for(half), RT, 0, s>>>(a_in, a_out, b, half, nhi, ww, om);
  WHIR_HIP(hipGetLastError());
  return half;
}

size_t fold_both_dev(int Ds, int D, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t n, const uint64_t* w,
                     hipStream_t s) {
  if (Ds == D) {
    switch (D) {
      case 1: return fold2_t<1, 1>(a_in, a_out, b, n, w, s);
      case 2: return fold2_t<2, 2>(a_in, a_out, b, n, w, s);
      case 3: return fold2_t<3, 3>(a_in, a_out, b, n, w, s);
    }
  } else if (Ds == 1) {
    switch (D) {
      case 2: return fold2_t<1, 2>(a_in, a_out, b, n, w, s);
      case 3: return fold2_t<1, 3>(a_in, a_out, b, n, w, s);
    }
  }
  throw Error("fold: extension degree must be 1, 2 or 3");
}

// ------------------------------------------------------------------ fused fold + polynomial
// Fold a (Ds words, into a_out) and b (in place) by w, then the round polynomial of
// the folded pair.  Folded length h = next_pow2(n)/2 (a power of two >= 2); the new
// pairs are (i, h/2 + i).  Lane i writes positions i and h/2 + i only and reads
// i, h/2 + i, h + i, h + h/2 + i, so a_out may alias a_in when Ds == D.
template <int Ds, int D>
__global__ __launch_bounds__(RT) void k_fold_poly(const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t h,
                                                  size_t nhi, Ext<D> w, Ext<D> om, uint64_t* parts,
                                                  MailboxSig sig) {
  Ext<D> acc[2] = {gl::ext_zero<D>(), gl::ext_zero<D>()};
  const size_t q = h >> 1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < q; i += (size_t)gridDim.x * blockDim.x) {
    Ext<D> a0 = fold_at<Ds, D>(a_in, i, h, nhi, w, om), a1 = fold_at<Ds, D>(a_in, q + i, h, nhi, w, om);
    Ext<D> b0 = fold_at<D, D>(b, i, h, nhi, w, om), b1 = fold_at<D, D>(b, q + i, h, nhi, w, om);
    st<D>(a_out, i, a0); st<D>(a_out, q + i, a1);
    st<D>(b, i, b0); st<D>(b, q + i, b1);
    acc[0] = gl::eadd<D>(acc[0], gl::emul<D>(a0, b0));
    acc[1] = gl::eadd<D>(acc[1], gl::emul<D>(gl::esub<D>(a1, a0), gl::esub<D>(b1, b0)));
  }
  block_reduce_store<D, 2>(acc, parts + (size_t)blockIdx.x * 2 * D, sig);
}

// Block partials of one fold + polynomial pass: none when one block does it all.
static unsigned fold_poly_grid(size_t n) {
  if (n <= 2) return 1;
  const size_t q = (next_pow2(n) >> 1) >> 1;
  return q <= SMALL_PAIRS ? 1u : grid_for(q);
}

// parts_buf: caller-provided room for the block partials (fold_poly_parts_words), or
// null to take it from the stream's pool.
template <int Ds, int D>
static size_t fold_poly_t(const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t n, const uint64_t* w,
                          uint64_t* out, hipStream_t s, uint64_t* parts_buf = nullptr) {
  if (n <= 2) {
    // folded length <= 1: fold then the one-element polynomial
    size_t m = fold_t<Ds, D>(a_in, a_out, n, w, s);
    fold_t<D, D>(b, b, n, w, s);
    sc_poly<D, D>(a_out, b, m, out, s);
    return m;
  }
  size_t h = next_pow2(n) >> 1, nhi = n - h;
  Ext<D> ww = host_ext<D>(w);
  Ext<D> om = gl::esub<D>(gl::ext_base<D>(1), ww);
  const unsigned g = fold_poly_grid(n);
  DevBuf parts(g == 1 || parts_buf ? 0 : (size_t)g * 2 * D * 8, s);
  uint64_t* pp = parts_buf ? parts_buf : parts.as<uint64_t>();
  {
    ProfScope prof(kernel_name("k_fold_poly", Ds, D), s, 8.0 * ((Ds + D) * (double)n + 2.0 * D * h));
    k_fold_poly<Ds, D><<<g, RT, 0, s>>>(a_in, a_out, b, h, nhi, ww, om, g == 1 ? out : pp,
                                        g == 1 ? mailbox_take(out) : MailboxSig{});
    WHIR_HIP(hipGetLastError());
  }
  if (g > 1) finish_reduce<D, 2>(pp, g, out, s);
  return h;
}

size_t fold_poly_parts_words(int D, size_t n) {
  const unsigned g = fold_poly_grid(n);
  return g == 1 ? 0 : (size_t)g * 2 * D;
}

size_t fold_and_poly_ws_dev(int D, uint64_t* a, uint64_t* b, size_t n, const uint64_t* w, uint64_t* out,
                            uint64_t* parts, hipStream_t s) {
  WHIR_REQUIRE(fold_poly_parts_words(D, n) == 0 || parts, "fold_and_poly: no room for the block partials");
  switch (D) {
    case 1: return fold_poly_t<1, 1>(a, a, b, n, w, out, s, parts);
    case 2: return fold_poly_t<2, 2>(a, a, b, n, w, out, s, parts);
    case 3: return fold_poly_t<3, 3>(a, a, b, n, w, out, s, parts);
  }
  throw Error("fold: extension degree must be 1, 2 or 3");
}

size_t fold_and_poly_mixed_dev(int Ds, int D, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t n,
                               const uint64_t* w, uint64_t* out, hipStream_t s) {
  if (Ds == D) {
    switch (D) {
      case 1: return fold_poly_t<1, 1>(a_in, a_out, b, n, w, out, s);
      case 2: return fold_poly_t<2, 2>(a_in, a_out, b, n, w, out, s);
      case 3: return fold_poly_t<3, 3>(a_in, a_out, b, n, w, out, s);
    }
  } else if (Ds == 1) {
    switch (D) {
      case 2: return fold_poly_t<1, 2>(a_in, a_out, b, n, w, out, s);
      case 3: return fold_poly_t<1, 3>(a_in, a_out, b, n, w, out, s);
    }
  }
  throw Error("fold: extension degree must be 1, 2 or 3");
}

size_t fold_and_poly_dev(int D, uint64_t* a, uint64_t* b, size_t n, const uint64_t* w, uint64_t* out,
                         hipStream_t s) {
  return fold_and_poly_mixed_dev(D, D, a, a, b, n, w, out, s);
}

// ------------------------------------------------------------------ two sumcheck rounds per pass
// Folding is linear in the challenge, so round k+1's polynomial is a polynomial in
// round k's challenge r whose coefficients the pass over round k's vectors can sum
// too.  For vectors x, y of length L = 4q (a power of two), h = L/2, dx_i = x_(h+i) -
// x_i and, after the fold by r, x'_i = x_i + r dx_i:
//   round k:    c0 = sum_(i<h) x_i y_i,  c2 = sum_(i<h) dx_i dy_i
//   round k+1:  c0' = S0 + r S1 + r^2 S2,  c2' = T0 + r T1 + r^2 T2, with (i < q)
//     S0 = sum x_i y_i,  S1 = sum (x_i dy_i + dx_i y_i),  S2 = sum dx_i dy_i,
//     U = x_(q+i) - x_i,  V = dx_(q+i) - dx_i:  T0 = sum Ux Uy,  T1 = sum (Ux Vy + Vx Uy),
//     T2 = sum Vx Vy
// (the same field values as folding first and summing after, the order of exact
// additions aside).  Lane i takes positions i, q+i, 2q+i, 3q+i.  With FOLD the vectors
// are first folded by two pending challenges (r1, then r2) from length 4L, position t
// from t, L+t, 2L+t, 3L+t of the source, and written (a to a_out, b in place: lane i
// writes only its own four positions, the only ones below L that it reads).  NS: sums
// per block, 8 (c0, c2, S0, S1, S2, T0, T1, T2), 2 (c0, c2) or 0.
template <int Ds, int D>
__device__ __forceinline__ Ext<D> fold_twice(const uint64_t* v, size_t t, size_t L, const Ext<D>& r1,
                                             const Ext<D>& r2) {
  Ext<D> p, q;
  if constexpr (Ds == 1 && D > 1) {
    // base-field source: the first fold scales r1 by a base-field difference
    const uint64_t x0 = v[t], x1 = v[L + t], x2 = v[2 * L + t], x3 = v[3 * L + t];
    p = gl::escale<D>(r1, gl::sub(x2, x0));
    p.c[0] = gl::add(p.c[0], x0);
    q = gl::escale<D>(r1, gl::sub(x3, x1));
    q.c[0] = gl::add(q.c[0], x1);
  } else {
    const Ext<D> x0 = ldx<Ds, D>(v, t), x1 = ldx<Ds, D>(v, L + t), x2 = ldx<Ds, D>(v, 2 * L + t),
                 x3 = ldx<Ds, D>(v, 3 * L + t);
    p = gl::eadd<D>(x0, gl::emul<D>(gl::esub<D>(x2, x0), r1));
    q = gl::eadd<D>(x1, gl::emul<D>(gl::esub<D>(x3, x1), r1));
  }
  return gl::eadd<D>(p, gl::emul<D>(gl::esub<D>(q, p), r2));
}

// Element i of a covector view: sum_k hi_k[i >> sh] lo_k[i & (2^sh - 1)].
template <int D>
__device__ __forceinline__ Ext<D> view_at(const CovView& v, size_t i) {
  const size_t h = i >> v.sh, l = i & (((size_t)1 << v.sh) - 1);
  Ext<D> r = gl::emul<D>(ld<D>(v.hi, h), ld<D>(v.lo, l));
  for (uint32_t k = 1; k < v.nt; k++)
    r = gl::eadd<D>(r, gl::emul<D>(ld<D>(v.hi, (size_t)k * v.nhi + h), ld<D>(v.lo, ((size_t)k << v.sh) + l)));
  return r;
}

// VIEW: b's values come from the view v (mode 0: y itself; folding modes: y already
// folded by (r1, r2)) instead of from b, and the folding modes write them into b.
template <int Ds, int D, bool FOLD, int NS, bool VIEW>
__global__ __launch_bounds__(RT) void k_sc_pair(const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t L,
                                                Ext<D> r1, Ext<D> r2, uint64_t* parts, MailboxSig sig, CovView v) {
  constexpr int NA = NS > 0 ? NS : 1;
  Ext<D> acc[NA];
#pragma unroll
  for (int k = 0; k < NA; k++) acc[k] = gl::ext_zero<D>();
  const size_t q = L >> 2;
  if constexpr (FOLD && NS == 0) {
    // the double fold alone: one output position per lane (four times the lanes of the
    // summing passes, which the short vectors at the end of a sumcheck need)
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < L; t += (size_t)gridDim.x * blockDim.x) {
      const Ext<D> x = fold_twice<Ds, D>(a_in, t, L, r1, r2);
      const Ext<D> y = VIEW ? view_at<D>(v, t) : fold_twice<D, D>(b, t, L, r1, r2);
      st<D>(a_out, t, x);
      st<D>(b, t, y);
    }
    return;
  }
  if constexpr (!FOLD && Ds == 1 && D > 1 && NS == 8) {
    // x in the base field: every product is base x extension (D multiplies, not an
    // extension product), and the differences of x are base-field differences.
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < q; i += (size_t)gridDim.x * blockDim.x) {
      uint64_t x[4];
      Ext<D> y[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        x[k] = a_in[k * q + i];
        y[k] = VIEW ? view_at<D>(v, k * q + i) : ld<D>(b, k * q + i);
      }
      const uint64_t dx0 = gl::sub(x[2], x[0]), dx1 = gl::sub(x[3], x[1]);
      const Ext<D> dy0 = gl::esub<D>(y[2], y[0]), dy1 = gl::esub<D>(y[3], y[1]);
      const Ext<D> m0 = gl::escale<D>(y[0], x[0]), d0 = gl::escale<D>(dy0, dx0);
      acc[0] = gl::eadd<D>(acc[0], gl::eadd<D>(m0, gl::escale<D>(y[1], x[1])));
      acc[1] = gl::eadd<D>(acc[1], gl::eadd<D>(d0, gl::escale<D>(dy1, dx1)));
      acc[2] = gl::eadd<D>(acc[2], m0);
      acc[3] = gl::eadd<D>(acc[3], gl::eadd<D>(gl::escale<D>(dy0, x[0]), gl::escale<D>(y[0], dx0)));
      acc[4] = gl::eadd<D>(acc[4], d0);
      const uint64_t ux = gl::sub(x[1], x[0]), vx = gl::sub(dx1, dx0);
      const Ext<D> uy = gl::esub<D>(y[1], y[0]), vy = gl::esub<D>(dy1, dy0);
      acc[5] = gl::eadd<D>(acc[5], gl::escale<D>(uy, ux));
      acc[6] = gl::eadd<D>(acc[6], gl::eadd<D>(gl::escale<D>(vy, ux), gl::escale<D>(uy, vx)));
      acc[7] = gl::eadd<D>(acc[7], gl::escale<D>(vy, vx));
    }
    block_reduce_store<D, NS>(acc, parts + (size_t)blockIdx.x * NS * D, sig);
    return;
  }
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < q; i += (size_t)gridDim.x * blockDim.x) {
    Ext<D> x[4], y[4];
    static_for<4>([&](auto kc) {  // a constant k keeps x and y in registers for Fp3 too
      constexpr int k = decltype(kc)::value;
      const size_t t = k * q + i;
      if constexpr (FOLD) {
        x[k] = fold_twice<Ds, D>(a_in, t, L, r1, r2);
        y[k] = VIEW ? view_at<D>(v, t) : fold_twice<D, D>(b, t, L, r1, r2);
      } else {
        x[k] = ldx<Ds, D>(a_in, t);
        y[k] = VIEW ? view_at<D>(v, t) : ld<D>(b, t);
      }
    });
    if constexpr (FOLD) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        st<D>(a_out, k * q + i, x[k]);
        st<D>(b, k * q + i, y[k]);
      }
    }
    if constexpr (NS > 0) {
      const Ext<D> dx0 = gl::esub<D>(x[2], x[0]), dx1 = gl::esub<D>(x[3], x[1]);
      const Ext<D> dy0 = gl::esub<D>(y[2], y[0]), dy1 = gl::esub<D>(y[3], y[1]);
      const Ext<D> m0 = gl::emul<D>(x[0], y[0]), d0 = gl::emul<D>(dx0, dy0);
      acc[0] = gl::eadd<D>(acc[0], gl::eadd<D>(m0, gl::emul<D>(x[1], y[1])));
      acc[1] = gl::eadd<D>(acc[1], gl::eadd<D>(d0, gl::emul<D>(dx1, dy1)));
      if constexpr (NS == 8) {
        acc[2] = gl::eadd<D>(acc[2], m0);
        acc[3] = gl::eadd<D>(acc[3], gl::eadd<D>(gl::emul<D>(x[0], dy0), gl::emul<D>(dx0, y[0])));
        acc[4] = gl::eadd<D>(acc[4], d0);
        const Ext<D> ux = gl::esub<D>(x[1], x[0]), vx = gl::esub<D>(dx1, dx0);
        const Ext<D> uy = gl::esub<D>(y[1], y[0]), vy = gl::esub<D>(dy1, dy0);
        acc[5] = gl::eadd<D>(acc[5], gl::emul<D>(ux, uy));
        acc[6] = gl::eadd<D>(acc[6], gl::eadd<D>(gl::emul<D>(ux, vy), gl::emul<D>(vx, uy)));
        acc[7] = gl::eadd<D>(acc[7], gl::emul<D>(vx, vy));
      }
    }
  }
  if constexpr (NS > 0) block_reduce_store<D, NS>(acc, parts + (size_t)blockIdx.x * NS * D, sig);
}

template <int Ds, int D, bool FOLD, int NS, bool VIEW>
static void sc_pair_t(const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t L, const uint64_t* r1,
                      const uint64_t* r2, uint64_t* out, const CovView& v, hipStream_t s) {
  const size_t q = L >> 2;
  const unsigned g = NS == 0 ? grid_for(L) : (q <= SMALL_PAIRS ? 1u : grid_for(q));
  DevBuf parts(NS > 0 && g > 1 ? (size_t)g * NS * D * 8 : 0, s);
  const Ext<D> w1 = r1 ? host_ext<D>(r1) : gl::ext_zero<D>(), w2 = r2 ? host_ext<D>(r2) : gl::ext_zero<D>();
  {
    ProfScope prof(kernel_name(FOLD ? (NS ? "k_sc_pair_fold" : "k_fold4") : "k_sc_pair", Ds, D), s,
                   FOLD ? 8.0 * ((Ds + D) * 4.0 * L + 2.0 * D * L) : 8.0 * (Ds + D) * (double)L);
    const bool direct = NS > 0 && g == 1;  // one block stores the sums itself
    k_sc_pair<Ds, D, FOLD, NS, VIEW><<<g, RT, 0, s>>>(a_in, a_out, b, L, w1, w2,
                                                      direct ? out : parts.as<uint64_t>(),
                                                      direct ? mailbox_take(out) : MailboxSig{}, v);
    WHIR_HIP(hipGetLastError());
  }
  if constexpr (NS > 0) {
    if (g > 1) finish_reduce<D, NS>(parts.as<uint64_t>(), g, out, s);
  }
}

template <int Ds, int D, bool VIEW>
static void sc_pair_d(int mode, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t L, const uint64_t* r1,
                      const uint64_t* r2, uint64_t* out, const CovView& v, hipStream_t s) {
  switch (mode) {
    case 0: sc_pair_t<Ds, D, false, 8, VIEW>(a_in, a_out, b, L, r1, r2, out, v, s); return;
    case 1: sc_pair_t<Ds, D, true, 8, VIEW>(a_in, a_out, b, L, r1, r2, out, v, s); return;
    case 2: sc_pair_t<Ds, D, true, 2, VIEW>(a_in, a_out, b, L, r1, r2, out, v, s); return;
    default: sc_pair_t<Ds, D, true, 0, VIEW>(a_in, a_out, b, L, r1, r2, out, v, s); return;
  }
}

template <bool VIEW>
static void sc_pair_any(int mode, int Ds, int D, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t L,
                        const uint64_t* r1, const uint64_t* r2, uint64_t* out, const CovView& v, hipStream_t s) {
  WHIR_REQUIRE(L >= 4 && (L & (L - 1)) == 0, "sumcheck pair: length must be a power of two >= 4");
  if (Ds == D) {
    switch (D) {
      case 1: sc_pair_d<1, 1, VIEW>(mode, a_in, a_out, b, L, r1, r2, out, v, s); return;
      case 2: sc_pair_d<2, 2, VIEW>(mode, a_in, a_out, b, L, r1, r2, out, v, s); return;
      case 3: sc_pair_d<3, 3, VIEW>(mode, a_in, a_out, b, L, r1, r2, out, v, s); return;
    }
  } else if (Ds == 1) {
    switch (D) {
      case 2: sc_pair_d<1, 2, VIEW>(mode, a_in, a_out, b, L, r1, r2, out, v, s); return;
      case 3: sc_pair_d<1, 3, VIEW>(mode, a_in, a_out, b, L, r1, r2, out, v, s); return;
    }
  }
  throw Error("sumcheck pair: extension degree must be 1, 2 or 3");
}

void sumcheck_pair_dev(int mode, int Ds, int D, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t L,
                       const uint64_t* r1, const uint64_t* r2, uint64_t* out, hipStream_t s) {
  sc_pair_any<false>(mode, Ds, D, a_in, a_out, b, L, r1, r2, out, CovView{}, s);
}

// ------------------------------------------------------------------ view passes
// The sums of a pairs pass whose covector y is a view, factored by the view's blocks.
// With quarters x_a(i) = x(a L/4 + i), y_b likewise (i < L/4), every sum of the pass is
// a fixed bilinear combination of the 16 values G[a][b] = sum_i x_a(i) y_b(i), and
//   G[a][b] = sum_t sum_h hi_t[b Qb + h] P_(a,t)(h),  P_(a,t)(h) = sum_l x_a(h 2^sh + l) lo_t[l]
// (Qb = L/4 / 2^sh blocks per quarter): the inner products run against the low tables
// (a base-field x costs D products per term, summed as 128-bit values and reduced once
// per block), and the high tables enter once per block.  One GPU block per (a, h).
constexpr int VIEW_MAX_TERMS = 4;

template <int Ds, int D, bool FOLD, int NT>
__global__ __launch_bounds__(RT) void k_view_pass(const uint64_t* a_in, uint64_t* a_out, size_t L, Ext<D> r1,
                                                  Ext<D> r2, CovView v, uint64_t* parts) {
  const size_t q = L >> 2, B = (size_t)1 << v.sh, Qb = q >> v.sh;
  const size_t a = blockIdx.x / Qb, h = blockIdx.x % Qb;
  const size_t base = a * q + h * B;
  Ext<D> P[NT];
  if constexpr (!FOLD && Ds == 1) {
    gl::Wide W[NT][D];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int d = 0; d < D; d++) W[t][d] = gl::Wide{0, 0, 0};
#pragma unroll 1
    for (size_t l = threadIdx.x; l < B; l += blockDim.x) {
      const uint64_t x = a_in[base + l];
      Ext<D> lo[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) lo[t] = ld<D>(v.lo, ((size_t)t << v.sh) + l);
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int d = 0; d < D; d++) gl::wide_mac(W[t][d], x, lo[t].c[d]);
    }
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int d = 0; d < D; d++) P[t].c[d] = gl::wide_reduce(W[t][d]);
  } else {
    // extension x: the products' coefficients are summed unreduced, one reduction per
    // coefficient per block
    gl::Wide W[NT][D];
#pragma unroll
    for (int t = 0; t < NT; t++) gl::wide_zero<D>(W[t]);
#pragma unroll 1
    for (size_t l = threadIdx.x; l < B; l += blockDim.x) {
      const size_t p = base + l;
      Ext<D> lo[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) lo[t] = ld<D>(v.lo, ((size_t)t << v.sh) + l);
      Ext<D> x;
      if constexpr (FOLD) {
        x = fold_twice<Ds, D>(a_in, p, L, r1, r2);
        st<D>(a_out, p, x);
      } else {
        x = ldx<Ds, D>(a_in, p);
      }
#pragma unroll
      for (int t = 0; t < NT; t++) gl::emul_acc<D>(W[t], x, lo[t]);
    }
#pragma unroll
    for (int t = 0; t < NT; t++) P[t] = gl::ext_reduce<D>(W[t]);
  }
  block_sum<D, NT>(P);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int b = 0; b < 4; b++) {
      Ext<D> g = gl::ext_zero<D>();
#pragma unroll
      for (int t = 0; t < NT; t++) g = gl::eadd<D>(g, gl::emul<D>(ld<D>(v.hi, (size_t)t * v.nhi + b * Qb + h), P[t]));
      st<D>(parts, (size_t)blockIdx.x * 4 + b, g);
    }
  }
}

// Sums the blocks' G contributions and writes the pass's NS sums, the same values the
// pairs kernel's accumulators hold (see k_sc_pair: quarters 0..3 of x and y, dx0 = x2 - x0,
// dx1 = x3 - x1, ux = x1 - x0, vx = dx1 - dx0, likewise for y).
template <int D, int NS>
__global__ __launch_bounds__(RT) void k_view_reduce(const uint64_t* parts, size_t Qb, uint64_t* out, MailboxSig sig) {
  Ext<D> G[16];
#pragma unroll
  for (int k = 0; k < 16; k++) G[k] = gl::ext_zero<D>();
#pragma unroll
  for (int a = 0; a < 4; a++)
    for (size_t h = threadIdx.x; h < Qb; h += blockDim.x)
#pragma unroll
      for (int b = 0; b < 4; b++) G[a * 4 + b] = gl::eadd<D>(G[a * 4 + b], ld<D>(parts, (a * Qb + h) * 4 + b));
  Ext<D> T[16];
  block_sum_lds<D, 16>(G, T);
  // Thread k < NS forms sum k = B(al1, be1) + B(al2, be2) from the totals, all threads on
  // one code path: the coefficient of G[a][b] is al1_a be1_b + al2_a be2_b (in -2..2).
  __shared__ uint64_t res[8 * D];
  if (threadIdx.x < NS) {
    // al1, be1, al2, be2 per sum (coefficient vectors over the quarters 0..3)
    constexpr int V[8][4][4] = {
        {{1, 0, 0, 0}, {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 1, 0, 0}},          // x0 y0 + x1 y1
        {{-1, 0, 1, 0}, {-1, 0, 1, 0}, {0, -1, 0, 1}, {0, -1, 0, 1}},      // dx0 dy0 + dx1 dy1
        {{1, 0, 0, 0}, {1, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},          // x0 y0
        {{1, 0, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 1, 0}, {1, 0, 0, 0}},        // x0 dy0 + dx0 y0
        {{-1, 0, 1, 0}, {-1, 0, 1, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},        // dx0 dy0
        {{-1, 1, 0, 0}, {-1, 1, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 0}},        // ux uy
        {{-1, 1, 0, 0}, {1, -1, -1, 1}, {1, -1, -1, 1}, {-1, 1, 0, 0}},    // ux vy + vx uy
        {{1, -1, -1, 1}, {1, -1, -1, 1}, {0, 0, 0, 0}, {0, 0, 0, 0}}};     // vx vy
    const int k = threadIdx.x;
    Ext<D> r = gl::ext_zero<D>();
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) {
        const int c = V[k][0][a] * V[k][1][b] + V[k][2][a] * V[k][3][b];
        const Ext<D>& t = T[a * 4 + b];
        const Ext<D> z = gl::ext_zero<D>();
        r = gl::eadd<D>(r, c > 0 ? t : z);
        r = gl::esub<D>(r, c < 0 ? t : z);
        r = gl::eadd<D>(r, c > 1 ? t : z);
        r = gl::esub<D>(r, c < -1 ? t : z);
      }
#pragma unroll
    for (int d = 0; d < D; d++) res[k * D + d] = r.c[d];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < NS * D; k++) out[k] = res[k];
    mailbox_raise(sig);
  }
}

template <int Ds, int D>
static void view_pass_t(int mode, const uint64_t* a_in, uint64_t* a_out, size_t L, const uint64_t* r1,
                        const uint64_t* r2, const CovView& v, uint64_t* out, hipStream_t s) {
  const size_t Qb = (L >> 2) >> v.sh;
  const unsigned blocks = (unsigned)(4 * Qb);
  DevBuf parts((size_t)blocks * 4 * D * 8, s);
  const Ext<D> w1 = r1 ? host_ext<D>(r1) : gl::ext_zero<D>(), w2 = r2 ? host_ext<D>(r2) : gl::ext_zero<D>();
  {
    ProfScope prof(kernel_name(mode ? "k_view_pass_fold" : "k_view_pass", Ds, D), s,
                   mode ? 8.0 * (Ds * 4.0 * L + D * (double)L) : 8.0 * Ds * (double)L);
    uint64_t* pp = parts.as<uint64_t>();
    auto go = [&](auto nt) {
      constexpr int NT = decltype(nt)::value;
      if (mode == 0) k_view_pass<Ds, D, false, NT><<<blocks, RT, 0, s>>>(a_in, a_out, L, w1, w2, v, pp);
      else k_view_pass<Ds, D, true, NT><<<blocks, RT, 0, s>>>(a_in, a_out, L, w1, w2, v, pp);
    };
    switch (v.nt) {
      case 1: go(std::integral_constant<int, 1>{}); break;
      case 2: go(std::integral_constant<int, 2>{}); break;
      case 3: go(std::integral_constant<int, 3>{}); break;
      default: go(std::integral_constant<int, 4>{}); break;
    }
    WHIR_HIP(hipGetLastError());
  }
  if (mode == 2) k_view_reduce<D, 2><<<1, RT, 0, s>>>(parts.as<uint64_t>(), Qb, out, mailbox_take(out));
  else k_view_reduce<D, 8><<<1, RT, 0, s>>>(parts.as<uint64_t>(), Qb, out, mailbox_take(out));
  WHIR_HIP(hipGetLastError());
}

void sumcheck_pair_view_dev(int mode, int Ds, int D, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t L,
                            const uint64_t* r1, const uint64_t* r2, const CovView& v, uint64_t* out, hipStream_t s) {
  WHIR_REQUIRE(v.nt > 0 && v.nt <= VIEW_MAX_TERMS && v.hi && v.lo, "sumcheck pair: bad covector view");
  WHIR_REQUIRE(((size_t)v.nhi << v.sh) == L, "sumcheck pair: covector view length mismatch");
  if (mode == 3) {  // the double fold alone; it writes the view's values into b
    sc_pair_any<true>(mode, Ds, D, a_in, a_out, b, L, r1, r2, out, v, s);
    return;
  }
  WHIR_REQUIRE(v.nhi >= 4 && mode >= 0 && mode <= 2, "sumcheck pair: covector view too short for a view pass");
  if (Ds == D) {
    switch (D) {
      case 1: return view_pass_t<1, 1>(mode, a_in, a_out, L, r1, r2, v, out, s);
      case 2: return view_pass_t<2, 2>(mode, a_in, a_out, L, r1, r2, v, out, s);
      case 3: return view_pass_t<3, 3>(mode, a_in, a_out, L, r1, r2, v, out, s);
    }
  } else if (Ds == 1) {
    switch (D) {
      case 2: return view_pass_t<1, 2>(mode, a_in, a_out, L, r1, r2, v, out, s);
      case 3: return view_pass_t<1, 3>(mode, a_in, a_out, L, r1, r2, v, out, s);
    }
  }
  throw Error("sumcheck pair: extension degree must be 1, 2 or 3");
}

// ------------------------------------------------------------------ covector terms
// covector_dev adds (or writes) any mix of
//   eq terms     s_p eq(z_p, i)      eval_eq (src/algebra/multilinear.rs:170-190)
//   power terms  t_j x_j^i           geometric_accumulate (src/algebra/mod.rs:154-177)
// for every i < n in one pass over HBM.  Element layout: block b covers
// [b S, (b+1) S) with S = RT K; lane t owns i = b S + k RT + t (k < K), so loads
// and stores are coalesced.  k_cov_tables builds every table on the device:
//   eq:    hi_p[h] = s_p prod_{v < k1} (h_v ? z_v : 1 - z_v) over the first k1
//          variables (h_0 = MSB), lo_p[l] the same over the last k2; eq = hi lo.
//   power: T0[j nb + b] = t_j x_j^(b S), T1[j RT + t] = x_j^t, X[j] = x_j^RT.
//          Base-field points (upper coefficients zero; the STIR in-domain points)
//          keep T1 and X in the base field, so a step costs D base products
//          instead of an extension product.
struct CovPlan {
  size_t np, nv, k1, k2, nhi, nlo;        // eq terms
  size_t m_e, m_b, nb, S;                 // power terms: extension points, then base points
  size_t in_z, in_s, in_x, in_t;          // uploaded inputs (word offsets)
  size_t hi, lo, T0, T1e, Xe, T1b, Xb;    // tables (word offsets)
  size_t words, entries;
  size_t in_sq;                           // per extension point: x^(S 2^k) for k < kb, then x^(2^k)
  int kb;                                 // for k <= 8 (squares made on the host)
};

// Inputs of the covector tables (eq points and scalars, power points and scalars) up
// to this many words travel with the launch itself, which saves a staged copy and
// its launch; larger inputs are copied into the table buffer first.
static constexpr size_t COV_ARG_WORDS = 320;
struct CovIn {
  uint64_t w[COV_ARG_WORDS];
};

// Element i (D words) of the input region starting at word `off`: from device memory
// when dev is given, else from the launch's arguments.
template <int D>
__device__ __forceinline__ Ext<D> in_ld(const uint64_t* dev, const CovIn& arg, size_t off, size_t i) {
  Ext<D> r;
#pragma unroll
  for (int d = 0; d < D; d++) r.c[d] = dev ? dev[off + i * D + d] : arg.w[off + i * D + d];
  return r;
}

template <int D>
__global__ void k_cov_tables(uint64_t* buf, CovPlan L, const uint64_t* in_dev, CovIn in_arg) {
  size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= L.entries) return;
  const size_t m = L.m_e + L.m_b;
  const size_t sec[7] = {L.np * L.nhi, L.np * L.nlo, m * L.nb, L.m_e * RT, L.m_e, L.m_b * RT, L.m_b};
  if (e < sec[0]) {
    const size_t p = e / L.nhi, h = e % L.nhi;
    Ext<D> v = in_ld<D>(in_dev, in_arg, L.in_s, p);
    for (size_t q = 0; q < L.k1; q++) {
      Ext<D> zq = in_ld<D>(in_dev, in_arg, L.in_z, p * L.nv + q);
      v = gl::emul<D>(v, ((h >> (L.k1 - 1 - q)) & 1) ? zq : gl::esub<D>(gl::ext_base<D>(1), zq));
    }
    st<D>(buf + L.hi, e, v);
    return;
  }
  e -= sec[0];
  if (e < sec[1]) {
    const size_t p = e / L.nlo, l = e % L.nlo;
    Ext<D> v = gl::ext_base<D>(1);
    for (size_t q = 0; q < L.k2; q++) {
      Ext<D> zq = in_ld<D>(in_dev, in_arg, L.in_z, p * L.nv + L.k1 + q);
      v = gl::emul<D>(v, ((l >> (L.k2 - 1 - q)) & 1) ? zq : gl::esub<D>(gl::ext_base<D>(1), zq));
    }
    st<D>(buf + L.lo, e, v);
    return;
  }
  e -= sec[1];
  // extension points take their powers from squares made on the host (a product per set
  // bit of the exponent instead of a square-and-multiply chain per lane)
  const size_t nsq = (size_t)L.kb + 9;
  if (e < sec[2]) {
    const size_t j = e / L.nb, b = e % L.nb;
    const Ext<D> t = in_ld<D>(in_dev, in_arg, L.in_t, j);
    Ext<D> v;
    if (j < L.m_e) {
      v = t;
      for (int k = 0; k < L.kb; k++)
        if ((b >> k) & 1) v = gl::emul<D>(v, in_ld<D>(in_dev, in_arg, L.in_sq, j * nsq + k));
    } else {
      v = gl::escale<D>(t, gl::pow(in_ld<D>(in_dev, in_arg, L.in_x, j).c[0], (uint64_t)b * L.S));
    }
    st<D>(buf + L.T0, e, v);
    return;
  }
  e -= sec[2];
  if (e < sec[3]) {
    const size_t j = e / RT, r = e % RT;
    Ext<D> v = gl::ext_base<D>(1);
    for (int k = 0; k < 8; k++)
      if ((r >> k) & 1) v = gl::emul<D>(v, in_ld<D>(in_dev, in_arg, L.in_sq, j * nsq + L.kb + k));
    st<D>(buf + L.T1e, e, v);
    return;
  }
  e -= sec[3];
  if (e < sec[4]) {
    static_assert(RT == 256, "x^RT is the square x^(2^8)");
    st<D>(buf + L.Xe, e, in_ld<D>(in_dev, in_arg, L.in_sq, e * nsq + L.kb + 8));
    return;
  }
  e -= sec[4];
  if (e < sec[5]) {
    buf[L.T1b + e] = gl::pow(in_ld<D>(in_dev, in_arg, L.in_x, L.m_e + e / RT).c[0], e % RT);
    return;
  }
  e -= sec[5];
  buf[L.Xb + e] = gl::pow(in_ld<D>(in_dev, in_arg, L.in_x, L.m_e + e).c[0], RT);
}

template <int D, int K>
__global__ __launch_bounds__(RT) void k_cov(uint64_t* acc, size_t n, int accumulate, const uint64_t* buf,
                                            CovPlan L) {
  const size_t b = blockIdx.x;
  const size_t i0 = b * (size_t)RT * K + threadIdx.x;
  Ext<D> sum[K];
#pragma unroll
  for (int k = 0; k < K; k++) sum[k] = gl::ext_zero<D>();
  if (L.np) {
    const size_t mask = L.nlo - 1;
    // term and query loops unrolled so several iterations' loads share one wait
#pragma unroll 2
    for (size_t p = 0; p < L.np; p++) {
      const uint64_t* hi = buf + L.hi + p * L.nhi * D;
      const uint64_t* lo = buf + L.lo + p * L.nlo * D;
#pragma unroll
      for (int k = 0; k < K; k++) {
        const size_t i = i0 + (size_t)k * RT;
        if (i < n) sum[k] = gl::eadd<D>(sum[k], gl::emul<D>(ld<D>(hi, i >> L.k2), ld<D>(lo, i & mask)));
      }
    }
  }
#pragma unroll 2
  for (size_t j = 0; j < L.m_e; j++) {
    Ext<D> p = gl::emul<D>(ld<D>(buf + L.T0, j * L.nb + b), ld<D>(buf + L.T1e, j * RT + threadIdx.x));
    const Ext<D> step = ld<D>(buf + L.Xe, j);
#pragma unroll
    for (int k = 0; k < K; k++) {
      sum[k] = gl::eadd<D>(sum[k], p);
      if (k + 1 < K) p = gl::emul<D>(p, step);
    }
  }
#pragma unroll 4
  for (size_t j = 0; j < L.m_b; j++) {
    Ext<D> p = gl::escale<D>(ld<D>(buf + L.T0, (L.m_e + j) * L.nb + b), buf[L.T1b + j * RT + threadIdx.x]);
    const uint64_t step = buf[L.Xb + j];
#pragma unroll
    for (int k = 0; k < K; k++) {
      sum[k] = gl::eadd<D>(sum[k], p);
      if (k + 1 < K) p = gl::escale<D>(p, step);
    }
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const size_t i = i0 + (size_t)k * RT;
    if (i < n) st<D>(acc, i, accumulate ? gl::eadd<D>(ld<D>(acc, i), sum[k]) : sum[k]);
  }
}

// Elements per lane: the largest K (up to a per-degree maximum) that still gives
// >= 1024 blocks.
// The largest K is 16 for Fp, 8 for Fp2 and 4 for Fp3 (fewer accumulators per lane,
// more waves: the d = 24 Fp2 covector 0.198 -> 0.177 ms, the d = 26 Fp3 ones 3.53 ->
// 2.82 ms on MI355X); WHIR_COV_KMAX2 / WHIR_COV_KMAX3 override them.
static int cov_k(size_t n, int D) {
  static const int kmax2 = [] {
    const char* e = getenv("WHIR_COV_KMAX2");
    return e ? atoi(e) : 8;
  }();
  static const int kmax3 = [] {
    const char* e = getenv("WHIR_COV_KMAX3");
    return e ? atoi(e) : 4;
  }();
  int K = D == 3 ? kmax3 : (D == 2 ? kmax2 : 16);
  while (K > 1 && (n + (size_t)RT * K - 1) / ((size_t)RT * K) < 1024) K >>= 1;
  return K;
}

// ------------------------------------------------------------------ many power terms
// sum_j t_j x_j^i for many points (the STIR terms of a round: the OOD samples and the
// in-domain queries), factored by i = 256 h + l:
//   t_j x_j^i = HI_j[h] * LO_j[l],  HI_j[h] = t_j x_j^(256 h),  LO_j[l] = x_j^l.
// Lane l of a block holds LO_j[l] for one point at a time and adds it times the block's
// K high values (the same for every lane: scalar loads) into 128-bit sums, reduced once
// per element at the end; in-domain points lie in the base field, so their terms cost D
// 64x64 products per element instead of an extension scale and an addition.  In-domain
// points come as exponents of the domain's root of unity (x = w_N^e), so each of their
// table entries is one twiddle-table product (exponents multiply mod N); the few OOD
// points' tables are powers computed in short runs.
constexpr int GEO_K = 4;        // rows h per block
constexpr int GEO_RUN = 8;      // consecutive OOD-point table entries per thread
struct GeoPlan {
  size_t m_e, m_b, H;           // OOD points first, then in-domain points; H = n / 256
  size_t in_x, in_t, in_e;      // inputs: OOD points (m_e D words), scalars (all points),
                                // in-domain exponents (m_b words)
  size_t hi, lo_e, lo_b;        // tables: HI (m H elements), LO of OOD points (m_e 256
                                // elements), LO of in-domain points (m_b 256 words)
  uint64_t N;                   // domain size of the in-domain points (a power of two)
  size_t in_sq;                 // inputs: per OOD point, x^(256 2^k) for k < kh, then x^(2^k)
  int kh;                       // for k < 8 (squares made on the host)
};

template <int D>
__global__ void k_geo_tables(uint64_t* buf, GeoPlan L, const uint64_t* in_dev, CovIn in_arg, TwTable tw) {
  const size_t runs_hi = (L.H + GEO_RUN - 1) / GEO_RUN;
  const size_t sec[4] = {L.m_b * L.H, L.m_b * 256, L.m_e * runs_hi, L.m_e * (256 / GEO_RUN)};
  size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e < sec[0]) {  // in-domain HI: t_j w^(256 h e_j)
    const size_t j = e / L.H, h = e % L.H;
    const uint64_t ej = in_dev ? in_dev[L.in_e + j] : in_arg.w[L.in_e + j];
    const uint64_t x = tw_at(tw, ((uint64_t)h * 256 * ej) & (L.N - 1));
    st<D>(buf + L.hi, (L.m_e + j) * L.H + h, gl::escale<D>(in_ld<D>(in_dev, in_arg, L.in_t, L.m_e + j), x));
    return;
  }
  e -= sec[0];
  if (e < sec[1]) {  // in-domain LO: w^(l e_j)
    const size_t j = e / 256, l = e % 256;
    const uint64_t ej = in_dev ? in_dev[L.in_e + j] : in_arg.w[L.in_e + j];
    buf[L.lo_b + e] = tw_at(tw, ((uint64_t)l * ej) & (L.N - 1));
    return;
  }
  e -= sec[1];
  const size_t nsq = (size_t)L.kh + 8;  // squares per OOD point
  if (e < sec[2]) {  // OOD HI in runs: t_j x_j^(256 h), x^(256 h0) from the squares of x^256
    const size_t j = e / runs_hi, h0 = (e % runs_hi) * GEO_RUN;
    const Ext<D> step = in_ld<D>(in_dev, in_arg, L.in_sq, j * nsq);
    Ext<D> v = in_ld<D>(in_dev, in_arg, L.in_t, j);
    for (int k = 0; k < L.kh; k++)
      if ((h0 >> k) & 1) v = gl::emul<D>(v, in_ld<D>(in_dev, in_arg, L.in_sq, j * nsq + k));
    for (size_t h = h0; h < h0 + GEO_RUN && h < L.H; h++) {
      st<D>(buf + L.hi, j * L.H + h, v);
      v = gl::emul<D>(v, step);
    }
    return;
  }
  e -= sec[2];
  if (e >= sec[3]) return;
  const size_t j = e / (256 / GEO_RUN), l0 = (e % (256 / GEO_RUN)) * GEO_RUN;  // OOD LO in runs: x_j^l
  const Ext<D> x = in_ld<D>(in_dev, in_arg, L.in_x, j);
  Ext<D> v = gl::ext_base<D>(1);
  for (int k = 0; k < 8; k++)
    if ((l0 >> k) & 1) v = gl::emul<D>(v, in_ld<D>(in_dev, in_arg, L.in_sq, j * nsq + L.kh + k));
  for (size_t l = l0; l < l0 + GEO_RUN; l++) {
    st<D>(buf + L.lo_e, j * 256 + l, v);
    v = gl::emul<D>(v, x);
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_geo(uint64_t* acc, size_t n, int accumulate, const uint64_t* buf, GeoPlan L) {
  const size_t l = threadIdx.x, h0 = (size_t)blockIdx.x * GEO_K;
  gl::Wide W[GEO_K][D];
#pragma unroll
  for (int k = 0; k < GEO_K; k++)
#pragma unroll
    for (int d = 0; d < D; d++) W[k][d] = gl::Wide{0, 0, 0};
  const uint64_t* hi = buf + L.hi;
  // in-domain points: HI (extension) x LO (base), D products per row.  The query loops
  // are unrolled so that the loads of several queries share one wait (short vectors run
  // about one wave per SIMD, with nothing else to hide the latency).
#pragma unroll 4
  for (size_t j = 0; j < L.m_b; j++) {
    const uint64_t lo = buf[L.lo_b + j * 256 + l];
    const uint64_t* hj = hi + ((L.m_e + j) * L.H + h0) * D;
#pragma unroll
    for (int k = 0; k < GEO_K; k++)
#pragma unroll
      for (int d = 0; d < D; d++) gl::wide_mac(W[k][d], hj[k * D + d], lo);
  }
  Ext<D> sum[GEO_K];
#pragma unroll
  for (int k = 0; k < GEO_K; k++)
#pragma unroll
    for (int d = 0; d < D; d++) sum[k].c[d] = gl::wide_reduce(W[k][d]);
  // OOD points: full extension products
#pragma unroll 2
  for (size_t j = 0; j < L.m_e; j++) {
    const Ext<D> lo = ld<D>(buf + L.lo_e, j * 256 + l);
#pragma unroll
    for (int k = 0; k < GEO_K; k++) sum[k] = gl::eadd<D>(sum[k], gl::emul<D>(ld<D>(hi, j * L.H + h0 + k), lo));
  }
#pragma unroll
  for (int k = 0; k < GEO_K; k++) {
    const size_t i = (h0 + k) * 256 + l;
    if (i < n) st<D>(acc, i, accumulate ? gl::eadd<D>(ld<D>(acc, i), sum[k]) : sum[k]);
  }
}

template <int D>
static void geo_terms_t(uint64_t* acc, size_t n, bool accumulate, const uint64_t* scalars, const uint64_t* ext_points,
                        size_t m_e, const uint64_t* exps, size_t m_b, int logN, hipStream_t s) {
  GeoPlan L{};
  L.H = n / 256;
  L.m_e = m_e;
  L.m_b = m_b;
  L.N = (uint64_t)1 << logN;
  const size_t m = m_e + m_b;
  std::vector<uint64_t> in;
  L.in_x = 0;
  in.insert(in.end(), ext_points, ext_points + m_e * D);
  L.in_t = in.size();
  in.insert(in.end(), scalars, scalars + m * D);
  L.in_e = in.size();
  in.insert(in.end(), exps, exps + m_b);
  // Squares for the OOD tables: x^(256 2^k) for k < kh (2^kh >= H) and x^(2^k) for k < 8,
  // so a lane forms its run's first power from the bits of its index.
  L.kh = 0;
  while (((size_t)1 << L.kh) < L.H) L.kh++;
  L.in_sq = in.size();
  for (size_t j = 0; j < m_e; j++) {
    Ext<D> x;
    for (int d = 0; d < D; d++) x.c[d] = ext_points[j * D + d];
    Ext<D> lo[8];
    for (int k = 0; k < 8; k++) {
      lo[k] = x;
      x = gl::emul<D>(x, x);
    }
    for (int k = 0; k < L.kh; k++) {  // x is x_j^256 here
      in.insert(in.end(), x.c, x.c + D);
      x = gl::emul<D>(x, x);
    }
    for (int k = 0; k < 8; k++) in.insert(in.end(), lo[k].c, lo[k].c + D);
  }
  const bool by_arg = in.size() <= COV_ARG_WORDS;
  size_t w = by_arg ? 0 : in.size();
  L.hi = w;   w += m * L.H * D;
  L.lo_e = w; w += m_e * 256 * D;
  L.lo_b = w; w += m_b * 256;
  DevBuf buf(w * 8, s);
  CovIn args{};
  if (by_arg) std::copy(in.begin(), in.end(), args.w);
  else upload_async(buf.ptr, in.data(), in.size() * 8, s);
  const TwTable tw = m_b ? twiddles(logN) : TwTable{};
  const size_t entries = m_b * L.H + m_b * 256 + m_e * ((L.H + GEO_RUN - 1) / GEO_RUN) + m_e * (256 / GEO_RUN);
  k_geo_tables<D><<<blocks_for(entries, RT), RT, 0, s>>>(buf.as<uint64_t>(), L, by_arg ? nullptr : buf.as<uint64_t>(),
                                                          args, tw);
  WHIR_HIP(hipGetLastError());
  ProfScope prof(kernel_name("k_geo", D), s, 8.0 * D * (accumulate ? 2.0 : 1.0) * (double)n);
  k_geo<D><<<(unsigned)(L.H / GEO_K), 256, 0, s>>>(acc, n, accumulate ? 1 : 0, buf.as<uint64_t>(), L);
  WHIR_HIP(hipGetLastError());
}

// WHIR_COV_GEO=0 keeps power terms on the one-pass kernel (for comparison).
static bool geo_path_enabled() {
  static const bool on = [] {
    const char* e = getenv("WHIR_COV_GEO");
    return !e || atoi(e) != 0;
  }();
  return on;
}

template <int D>
static void covector_t(uint64_t* acc, size_t n, bool accumulate, const uint64_t* eq_points,
                       const uint64_t* eq_scalars, size_t np, size_t nv, const uint64_t* scalars,
                       const uint64_t* points, size_t m, hipStream_t s) {
  if (!n) return;
  WHIR_REQUIRE(!np || n == ((size_t)1 << nv), "eval_eq: the covector length must be 2^(number of variables)");
  if (!np && !m) {
    if (!accumulate) fill_zero(acc, n * D * 8, s);
    return;
  }
  const int K = cov_k(n, D);
  CovPlan L{};
  L.np = np;
  L.nv = np ? nv : 0;
  L.k1 = L.nv / 2;
  L.k2 = L.nv - L.k1;
  L.nhi = (size_t)1 << L.k1;
  L.nlo = (size_t)1 << L.k2;
  L.S = (size_t)RT * K;
  L.nb = (n + L.S - 1) / L.S;
  // Extension points first, then points that lie in the base field.
  std::vector<size_t> order_e, order_b;
  for (size_t j = 0; j < m; j++) {
    bool base = D > 1;
    for (int d = 1; d < D; d++) base = base && points[j * D + d] == 0;
    (base ? order_b : order_e).push_back(j);
  }
  L.m_e = order_e.size();
  L.m_b = order_b.size();
  std::vector<uint64_t> in;
  L.in_z = 0;
  if (np) in.insert(in.end(), eq_points, eq_points + np * L.nv * D);
  L.in_s = in.size();
  if (np) in.insert(in.end(), eq_scalars, eq_scalars + np * D);
  L.in_x = in.size();
  for (size_t j : order_e) in.insert(in.end(), points + j * D, points + (j + 1) * D);
  for (size_t j : order_b) in.insert(in.end(), points + j * D, points + (j + 1) * D);
  L.in_t = in.size();
  for (size_t j : order_e) in.insert(in.end(), scalars + j * D, scalars + (j + 1) * D);
  for (size_t j : order_b) in.insert(in.end(), scalars + j * D, scalars + (j + 1) * D);
  // Squares of the extension points: x^(S 2^k) for k < kb (2^kb >= nb), then x^(2^k)
  // for k <= 8.
  L.kb = 0;
  while (((size_t)1 << L.kb) < L.nb) L.kb++;
  L.in_sq = in.size();
  for (size_t j : order_e) {
    Ext<D> x;
    for (int d = 0; d < D; d++) x.c[d] = points[j * D + d];
    Ext<D> xs = gl::epow<D>(x, L.S);
    for (int k = 0; k < L.kb; k++) {
      in.insert(in.end(), xs.c, xs.c + D);
      xs = gl::emul<D>(xs, xs);
    }
    for (int k = 0; k <= 8; k++) {
      in.insert(in.end(), x.c, x.c + D);
      x = gl::emul<D>(x, x);
    }
  }
  // Small inputs go with the launch; larger ones are copied to the front of the buffer.
  const bool by_arg = in.size() <= COV_ARG_WORDS;
  size_t w = by_arg ? 0 : in.size();
  L.hi = w;  w += np * L.nhi * D;
  L.lo = w;  w += np * L.nlo * D;
  L.T0 = w;  w += m * L.nb * D;
  L.T1e = w; w += L.m_e * RT * D;
  L.Xe = w;  w += L.m_e * D;
  L.T1b = w; w += L.m_b * RT;
  L.Xb = w;  w += L.m_b;
  L.words = w;
  L.entries = np * (L.nhi + L.nlo) + m * L.nb + L.m_e * (RT + 1) + L.m_b * (RT + 1);
  DevBuf buf(L.words * 8, s);
  CovIn args{};
  if (by_arg) std::copy(in.begin(), in.end(), args.w);
  else upload_async(buf.ptr, in.data(), in.size() * 8, s);
  k_cov_tables<D><<<blocks_for(L.entries, RT), RT, 0, s>>>(buf.as<uint64_t>(), L, by_arg ? nullptr : buf.as<uint64_t>(),
                                                            args);
  WHIR_HIP(hipGetLastError());
  ProfScope prof(kernel_name("k_cov", D, K), s, 8.0 * D * (accumulate ? 2.0 : 1.0) * (double)n);
  const unsigned g = (unsigned)L.nb;
  const int ac = accumulate ? 1 : 0;
  uint64_t* tb = buf.as<uint64_t>();
  switch (K) {
    case 16: k_cov<D, 16><<<g, RT, 0, s>>>(acc, n, ac, tb, L); break;
    case 8: k_cov<D, 8><<<g, RT, 0, s>>>(acc, n, ac, tb, L); break;
    case 4: k_cov<D, 4><<<g, RT, 0, s>>>(acc, n, ac, tb, L); break;
    case 2: k_cov<D, 2><<<g, RT, 0, s>>>(acc, n, ac, tb, L); break;
    default: k_cov<D, 1><<<g, RT, 0, s>>>(acc, n, ac, tb, L); break;
  }
  WHIR_HIP(hipGetLastError());
}

void covector_dev(int D, uint64_t* acc, size_t n, bool accumulate, const uint64_t* eq_points,
                  const uint64_t* eq_scalars, size_t np, size_t nv, const uint64_t* scalars, const uint64_t* points,
                  size_t m, hipStream_t s) {
  switch (D) {
    case 1: covector_t<1>(acc, n, accumulate, eq_points, eq_scalars, np, nv, scalars, points, m, s); break;
    case 2: covector_t<2>(acc, n, accumulate, eq_points, eq_scalars, np, nv, scalars, points, m, s); break;
    case 3: covector_t<3>(acc, n, accumulate, eq_points, eq_scalars, np, nv, scalars, points, m, s); break;
    default: throw Error("covector: extension degree must be 1, 2 or 3");
  }
}

// ------------------------------------------------------------------ covector views
// Table layout in the buffer: the uploaded inputs (when they do not travel with the
// launch), then the nt hi tables, then the nt lo tables.  Terms: np eq terms (points
// z_p of nv elements, scalars s_p), then m geometric terms (points x_j, scalars t_j):
//   eq:  hi[h] = s_p prod_{q < nv-sh} (h_q ? z_q : 1 - z_q)   (h_0 = MSB of h)
//        lo[l] = prod_{q < sh} (l_q ? z_(nv-sh+q) : 1 - z_(nv-sh+q))
//   geo: hi[h] = t_j x_j^(h 2^sh), lo[l] = x_j^l.
struct CovViewPlan {
  size_t np, m, nv, sh, nhi, nlo;
  size_t in_z, in_s, in_x, in_t;  // input offsets (words)
  size_t hi, lo;                  // table offsets (words)
};

template <int D>
__global__ void k_covview_tables(uint64_t* buf, CovViewPlan L, const uint64_t* in_dev, CovIn in_arg) {
  size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t nt = L.np + L.m, k1 = L.nv - L.sh;
  if (e < nt * L.nhi) {
    const size_t k = e / L.nhi, h = e % L.nhi;
    Ext<D> v;
    if (k < L.np) {
      v = in_ld<D>(in_dev, in_arg, L.in_s, k);
      for (size_t q = 0; q < k1; q++) {
        const Ext<D> zq = in_ld<D>(in_dev, in_arg, L.in_z, k * L.nv + q);
        v = gl::emul<D>(v, ((h >> (k1 - 1 - q)) & 1) ? zq : gl::esub<D>(gl::ext_base<D>(1), zq));
      }
    } else {
      const size_t j = k - L.np;
      v = gl::emul<D>(in_ld<D>(in_dev, in_arg, L.in_t, j),
                      gl::epow<D>(in_ld<D>(in_dev, in_arg, L.in_x, j), (uint64_t)h << L.sh));
    }
    st<D>(buf + L.hi, e, v);
    return;
  }
  e -= nt * L.nhi;
  if (e >= nt * L.nlo) return;
  const size_t k = e / L.nlo, l = e % L.nlo;
  Ext<D> v;
  if (k < L.np) {
    v = gl::ext_base<D>(1);
    for (size_t q = 0; q < L.sh; q++) {
      const Ext<D> zq = in_ld<D>(in_dev, in_arg, L.in_z, k * L.nv + k1 + q);
      v = gl::emul<D>(v, ((l >> (L.sh - 1 - q)) & 1) ? zq : gl::esub<D>(gl::ext_base<D>(1), zq));
    }
  } else {
    v = gl::epow<D>(in_ld<D>(in_dev, in_arg, L.in_x, k - L.np), l);
  }
  st<D>(buf + L.lo, e, v);
}

size_t covview_words(int D, size_t np, size_t m, size_t nv) {
  const size_t sh = nv - nv / 2, nt = np + m;
  return (np * nv + np + 2 * m) * D + nt * (((size_t)1 << (nv - sh)) + ((size_t)1 << sh)) * D;
}

template <int D>
static CovView covview_build_t(uint64_t* buf, size_t nv, const uint64_t* eq_points, const uint64_t* eq_scalars,
                               size_t np, const uint64_t* scalars, const uint64_t* points, size_t m, hipStream_t s) {
  CovViewPlan L{};
  L.np = np;
  L.m = m;
  L.nv = nv;
  L.sh = nv - nv / 2;
  L.nhi = (size_t)1 << (nv - L.sh);
  L.nlo = (size_t)1 << L.sh;
  std::vector<uint64_t> in;
  L.in_z = 0;
  in.insert(in.end(), eq_points, eq_points + np * nv * D);
  L.in_s = in.size();
  in.insert(in.end(), eq_scalars, eq_scalars + np * D);
  L.in_x = in.size();
  in.insert(in.end(), points, points + m * D);
  L.in_t = in.size();
  in.insert(in.end(), scalars, scalars + m * D);
  const bool by_arg = in.size() <= COV_ARG_WORDS;
  const size_t nt = np + m;
  L.hi = in.size();  // the input region stays reserved either way (covview_words)
  L.lo = L.hi + nt * L.nhi * D;
  CovIn args{};
  if (by_arg) std::copy(in.begin(), in.end(), args.w);
  else upload_async(buf, in.data(), in.size() * 8, s);
  const size_t entries = nt * (L.nhi + L.nlo);
  k_covview_tables<D><<<blocks_for(entries, RT), RT, 0, s>>>(buf, L, by_arg ? nullptr : buf, args);
  WHIR_HIP(hipGetLastError());
  CovView v;
  v.hi = buf + L.hi;
  v.lo = buf + L.lo;
  v.nt = (uint32_t)nt;
  v.sh = (uint32_t)L.sh;
  v.nhi = L.nhi;
  return v;
}

CovView covview_build_dev(int D, uint64_t* buf, size_t nv, const uint64_t* eq_points, const uint64_t* eq_scalars,
                          size_t np, const uint64_t* scalars, const uint64_t* points, size_t m, hipStream_t s) {
  WHIR_REQUIRE(np + m > 0 && nv >= 2 && nv < 63, "covector view: needs terms and 2 <= nv < 63");
  switch (D) {
    case 1: return covview_build_t<1>(buf, nv, eq_points, eq_scalars, np, scalars, points, m, s);
    case 2: return covview_build_t<2>(buf, nv, eq_points, eq_scalars, np, scalars, points, m, s);
    case 3: return covview_build_t<3>(buf, nv, eq_points, eq_scalars, np, scalars, points, m, s);
    default: throw Error("covector view: extension degree must be 1, 2 or 3");
  }
}

template <int D>
__global__ void k_covview_fold2(const uint64_t* hi, uint64_t* out, size_t nt, size_t nhi, Ext<D> r1, Ext<D> r2) {
  const size_t q = nhi >> 2;
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= nt * q) return;
  const size_t k = e / q, t = e % q;
  st<D>(out, e, fold_twice<D, D>(hi + k * nhi * D, t, q, r1, r2));
}

template <int D>
__global__ void k_covview_write(CovView v, uint64_t* out, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    st<D>(out, i, view_at<D>(v, i));
}

CovView covview_fold2_dev(int D, const CovView& v, const uint64_t* r1, const uint64_t* r2, uint64_t* out_hi,
                          hipStream_t s) {
  WHIR_REQUIRE(v.nt > 0 && v.nhi >= 4, "covector view: nothing left to fold in the high tables");
  const size_t n = (size_t)v.nt * (v.nhi >> 2);
  switch (D) {
    case 1: k_covview_fold2<1><<<blocks_for(n, RT), RT, 0, s>>>(v.hi, out_hi, v.nt, v.nhi, host_ext<1>(r1), host_ext<1>(r2)); break;
    case 2: k_covview_fold2<2><<<blocks_for(n, RT), RT, 0, s>>>(v.hi, out_hi, v.nt, v.nhi, host_ext<2>(r1), host_ext<2>(r2)); break;
    case 3: k_covview_fold2<3><<<blocks_for(n, RT), RT, 0, s>>>(v.hi, out_hi, v.nt, v.nhi, host_ext<3>(r1), host_ext<3>(r2)); break;
    default: throw Error("covector view: extension degree must be 1, 2 or 3");
  }
  WHIR_HIP(hipGetLastError());
  CovView f = v;
  f.hi = out_hi;
  f.nhi = v.nhi >> 2;
  return f;
}

void covview_write_dev(int D, const CovView& v, uint64_t* out, size_t n, hipStream_t s) {
  WHIR_REQUIRE(n <= ((size_t)v.nhi << v.sh), "covector view: output longer than the view");
  if (!n) return;
  const unsigned g = grid_for(n);
  switch (D) {
    case 1: k_covview_write<1><<<g, RT, 0, s>>>(v, out, n); break;
    case 2: k_covview_write<2><<<g, RT, 0, s>>>(v, out, n); break;
    case 3: k_covview_write<3><<<g, RT, 0, s>>>(v, out, n); break;
    default: throw Error("covector view: extension degree must be 1, 2 or 3");
  }
  WHIR_HIP(hipGetLastError());
}

void covector_stir_dev(int D, uint64_t* acc, size_t n, bool accumulate, const uint64_t* scalars,
                       const uint64_t* ext_points, size_t m_e, const uint64_t* exps, const uint64_t* points, size_t m_b,
                       int logN, hipStream_t s) {
  const size_t m = m_e + m_b;
  if (!geo_path_enabled() || m < 8 || n < ((size_t)1 << 14) || n % (256 * GEO_K) || logN > 62) {
    // the one-pass kernel, with every point as a field element (in-domain points lifted)
    std::vector<uint64_t> pts(ext_points, ext_points + m_e * D);
    for (size_t j = 0; j < m_b; j++) {
      pts.push_back(points[j]);
      for (int d = 1; d < D; d++) pts.push_back(0);
    }
    covector_dev(D, acc, n, accumulate, nullptr, nullptr, 0, 0, scalars, pts.data(), m, s);
    return;
  }
  switch (D) {
    case 1: geo_terms_t<1>(acc, n, accumulate, scalars, ext_points, m_e, exps, m_b, logN, s); break;
    case 2: geo_terms_t<2>(acc, n, accumulate, scalars, ext_points, m_e, exps, m_b, logN, s); break;
    case 3: geo_terms_t<3>(acc, n, accumulate, scalars, ext_points, m_e, exps, m_b, logN, s); break;
    default: throw Error("covector: extension degree must be 1, 2 or 3");
  }
}

void eval_eq_dev(int D, uint64_t* acc, const uint64_t* point, size_t k, const uint64_t* scalar, hipStream_t s) {
  covector_dev(D, acc, (size_t)1 << k, true, point, scalar, 1, k, nullptr, nullptr, 0, s);
}

void geometric_accumulate_dev(int D, uint64_t* acc, size_t n, const uint64_t* scalars, const uint64_t* points,
                              size_t m, hipStream_t s) {
  covector_dev(D, acc, n, true, nullptr, nullptr, 0, 0, scalars, points, m, s);
}

// ------------------------------------------------------------------ univariate evaluation
// mixed_univariate_evaluate (src/algebra/mod.rs:105-129) at up to UNI_MP points per
// pass: f(x_j) = sum_b x_j^(b S) sum_{q < S} c_(b S + q) x_j^q, S = RT UNI_K.  Block
// b sums its segment against the power table P[j][q] = x_j^q (one base-by-extension
// or extension product per element and point, no chained step); the final kernel
// scales block partials by B[j][b] = x_j^(b S) and adds them.
static constexpr int UNI_K = 16;
static constexpr int UNI_MP = 4;

struct UniPlan {
  size_t m, nb, S;
  size_t P, B, words, entries;
  uint64_t x[UNI_MP * 3];  // the points (m x D words), passed with the launch
};

// x_j^(2^b) for b < UNI_PW, computed on the host and passed with the launch: a table
// entry x_j^e is then the product over the set bits of e, one multiply per set bit
// instead of a squaring and a multiply per bit (the table launch runs one short wave
// per SIMD, so its time is its instruction count).
static constexpr int UNI_PW = 32;
struct UniPow {
  uint64_t w[UNI_MP * UNI_PW * 3];
};

template <int D>
__global__ void k_uni_tables(uint64_t* buf, UniPlan L, UniPow pw) {
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= L.entries) return;
  const bool in_p = e < L.m * L.S;
  const size_t j = in_p ? e / L.S : (e - L.m * L.S) / L.nb;
  uint64_t ex = in_p ? e % L.S : (uint64_t)((e - L.m * L.S) % L.nb) * L.S;
  Ext<D> r = gl::ext_base<D>(1);
  for (int b = 0; ex; b++, ex >>= 1) {
    if (ex & 1) {
      Ext<D> p;
#pragma unroll
      for (int d = 0; d < D; d++) p.c[d] = pw.w[(j * UNI_PW + b) * D + d];
      r = gl::emul<D>(r, p);
    }
  }
  if (in_p) st<D>(buf + L.P, e, r);
  else st<D>(buf + L.B, e - L.m * L.S, r);
}

template <int Ds, int D>
__global__ __launch_bounds__(RT) void k_uni(const uint64_t* c, size_t n, const uint64_t* buf, UniPlan L, int m,
                                            uint64_t* parts) {
  const size_t b = blockIdx.x;
  const size_t i0 = b * L.S + threadIdx.x;
  Ext<D> acc[UNI_MP];
#pragma unroll
  for (int j = 0; j < UNI_MP; j++) acc[j] = gl::ext_zero<D>();
#pragma unroll
  for (int k = 0; k < UNI_K; k++) {
    const size_t i = i0 + (size_t)k * RT;
    if (i >= n) break;
    const size_t q = (size_t)k * RT + threadIdx.x;
#pragma unroll
    for (int j = 0; j < UNI_MP; j++) {
      if (j < m) {
        const Ext<D> pw = ld<D>(buf + L.P, j * L.S + q);
        if constexpr (Ds == 1) acc[j] = gl::eadd<D>(acc[j], gl::escale<D>(pw, c[i]));
        else acc[j] = gl::eadd<D>(acc[j], gl::emul<D>(pw, ld<D>(c, i)));
      }
    }
  }
#pragma unroll
  for (int j = 0; j < UNI_MP; j++)
    if (j < m) acc[j] = gl::emul<D>(acc[j], ld<D>(buf + L.B, j * L.nb + b));  // segment offset x_j^(b S)
  block_reduce_store<D, UNI_MP>(acc, parts + (size_t)blockIdx.x * UNI_MP * D);
}

// Vectors of at most UNI_SMALL_BLOCKS blocks skip the table launch: each thread forms
// x_j^(i0) (its first index) by square-and-multiply from the squares x_j^(2^k) passed
// with the launch, then steps by x_j^RT; one block stores its sums itself.
static constexpr size_t UNI_SMALL_BLOCKS = 256;

template <int Ds, int D, int MP>
__global__ __launch_bounds__(RT) void k_uni_small(const uint64_t* c, size_t n, UniPow pw, uint64_t* parts,
                                                  MailboxSig sig) {
  static_assert(RT == 256, "the step x^RT is the square x^(2^8)");
  const size_t i0 = blockIdx.x * (size_t)RT * UNI_K + threadIdx.x;
  // Per point: x_j^i0 by square-and-multiply, then the lane's elements in order.  The
  // elements are loaded a chunk at a time before the chunk's arithmetic (zero past the
  // end), so a chunk's loads are in flight together rather than one per step of the
  // power chain; Fp3 takes chunks of four (sixteen unrolled Fp3 steps do not fit the
  // registers).
  constexpr int UC = D == 3 ? 4 : UNI_K;
  Ext<D> acc[UNI_MP];
#pragma unroll
  for (int j = 0; j < UNI_MP; j++) acc[j] = gl::ext_zero<D>();
  if (i0 < n) {
    static_for<MP>([&](auto jc) {
      constexpr int j = decltype(jc)::value;
      Ext<D> p = gl::ext_base<D>(1);
      for (int bit = 0; (i0 >> bit) != 0; bit++)
        if ((i0 >> bit) & 1) {
          Ext<D> sq;
#pragma unroll
          for (int d = 0; d < D; d++) sq.c[d] = pw.w[(j * UNI_PW + bit) * D + d];
          p = gl::emul<D>(p, sq);
        }
      Ext<D> step;
#pragma unroll
      for (int d = 0; d < D; d++) step.c[d] = pw.w[(j * UNI_PW + 8) * D + d];
      for (int k0 = 0; k0 < UNI_K; k0 += UC) {
        Ext<Ds> cv[UC];
#pragma unroll
        for (int k = 0; k < UC; k++) {
          const size_t i = i0 + (size_t)(k0 + k) * RT;
          cv[k] = i < n ? ldx<Ds, Ds>(c, i) : gl::ext_zero<Ds>();
        }
#pragma unroll
        for (int k = 0; k < UC; k++) {
          if constexpr (Ds == 1) acc[j] = gl::eadd<D>(acc[j], gl::escale<D>(p, cv[k].c[0]));
          else acc[j] = gl::eadd<D>(acc[j], gl::emul<D>(p, cv[k]));
          p = gl::emul<D>(p, step);
        }
      }
    });
  }
  block_reduce_store<D, UNI_MP>(acc, parts + (size_t)blockIdx.x * UNI_MP * D, sig);
}

template <int Ds, int D>
static void uni_t(const uint64_t* c, size_t n, const uint64_t* xs, size_t m, uint64_t* out, hipStream_t s,
                  bool padded) {
  for (size_t j0 = 0; j0 < m; j0 += UNI_MP) {
    const size_t mj = std::min<size_t>(UNI_MP, m - j0);
    uint64_t* o = out + j0 * D;
    if (!n) {
      fill_zero(o, mj * D * 8, s);
      continue;
    }
    UniPlan L{};
    L.m = mj;
    L.S = (size_t)RT * UNI_K;
    L.nb = (n + L.S - 1) / L.S;
    std::copy(xs + j0 * D, xs + (j0 + mj) * D, L.x);
    L.P = 0;
    L.B = L.P + mj * L.S * D;
    L.words = L.B + mj * L.nb * D;
    L.entries = mj * (L.S + L.nb);
    const bool small = L.nb <= UNI_SMALL_BLOCKS;
    DevBuf buf(small ? 0 : L.words * 8, s), parts(L.nb > 1 ? L.nb * UNI_MP * D * 8 : 0, s),
        res((mj == UNI_MP || padded) ? 0 : UNI_MP * D * 8, s);
    WHIR_REQUIRE((uint64_t)(L.nb - 1) * L.S < ((uint64_t)1 << UNI_PW), "univariate: vector too long");
    UniPow pw;
    for (size_t j = 0; j < mj; j++) {
      Ext<D> p;
      for (int d = 0; d < D; d++) p.c[d] = xs[(j0 + j) * D + d];
      for (int b = 0; b < UNI_PW; b++) {
        for (int d = 0; d < D; d++) pw.w[(j * UNI_PW + b) * D + d] = p.c[d];
        p = gl::emul<D>(p, p);
      }
    }
    // The reduction writes all UNI_MP slots; only the first mj go to the caller
    // (straight into `out` when it has room for whole groups).
    uint64_t* dst = (mj == UNI_MP || padded) ? o : res.as<uint64_t>();
    const bool last = j0 + UNI_MP >= m;
    if (small) {
      {
        ProfScope prof(kernel_name("k_uni_small", Ds, D), s, 8.0 * Ds * (double)n);
        uint64_t* o_small = L.nb == 1 ? dst : parts.as<uint64_t>();
        const MailboxSig sig = L.nb == 1 && last ? mailbox_take(dst) : MailboxSig{};
        auto go = [&](auto mp) {
          k_uni_small<Ds, D, decltype(mp)::value><<<(unsigned)L.nb, RT, 0, s>>>(c, n, pw, o_small, sig);
        };
        switch (mj) {
          case 1: go(std::integral_constant<int, 1>{}); break;
          case 2: go(std::integral_constant<int, 2>{}); break;
          case 3: go(std::integral_constant<int, 3>{}); break;
          default: go(std::integral_constant<int, 4>{}); break;
        }
        WHIR_HIP(hipGetLastError());
      }
      if (L.nb > 1) finish_reduce<D, UNI_MP>(parts.as<uint64_t>(), (unsigned)L.nb, dst, s, last);
      if (dst != o) WHIR_HIP(hipMemcpyAsync(o, dst, mj * D * 8, hipMemcpyDeviceToDevice, s));
      continue;
    }
    k_uni_tables<D><<<blocks_for(L.entries, RT), RT, 0, s>>>(buf.as<uint64_t>(), L, pw);
    WHIR_HIP(hipGetLastError());
    {
      ProfScope prof(kernel_name("k_uni", Ds, D), s, 8.0 * Ds * (double)n);
      k_uni<Ds, D><<<(unsigned)L.nb, RT, 0, s>>>(c, n, buf.as<uint64_t>(), L, (int)mj, parts.as<uint64_t>());
      WHIR_HIP(hipGetLastError());
    }
    finish_reduce<D, UNI_MP>(parts.as<uint64_t>(), (unsigned)L.nb, dst, s, last);
    if (dst != o) WHIR_HIP(hipMemcpyAsync(o, dst, mj * D * 8, hipMemcpyDeviceToDevice, s));
  }
}

void univariate_eval_many_dev(int Ds, int D, const uint64_t* c, size_t n, const uint64_t* xs, size_t m,
                              uint64_t* out, hipStream_t s, bool padded) {
  if (!m) return;
  if (Ds == 1) {
    switch (D) {
      case 1: uni_t<1, 1>(c, n, xs, m, out, s, padded); return;
      case 2: uni_t<1, 2>(c, n, xs, m, out, s, padded); return;
      case 3: uni_t<1, 3>(c, n, xs, m, out, s, padded); return;
    }
  } else if (Ds == D) {
    switch (D) {
      case 2: uni_t<2, 2>(c, n, xs, m, out, s, padded); return;
      case 3: uni_t<3, 3>(c, n, xs, m, out, s, padded); return;
    }
  }
  throw Error("univariate_evaluate: unsupported degrees");
}

// ------------------------------------------------------------------ batched univariate evaluation
// Evaluations f_e(x_e) = sum_i c_e[i] x_e^i of many base-field vectors (e < count),
// each at its own point, in three launches for the whole batch: tables (x_e^q for
// q < S and x_e^(b S) per segment, from host-computed squarings), segment sums (one
// block per segment of S = RT * UNI_K coefficients, found by a binary search over the
// jobs' first segments), and one reducing block per job.
struct UniJobDev {
  const uint64_t* c;
  uint64_t n, nb, blk0, P, B;
};

template <int D>
__global__ void k_uni_many_tables(uint64_t* buf, const UniJobDev* jobs, const uint64_t* pw, size_t count,
                                  size_t entries_per_job_p) {
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  const size_t S = (size_t)RT * UNI_K;
  // entries: count * S powers x^q, then every job's segment offsets x^(b S)
  size_t j, i;
  uint64_t ex;
  uint64_t* dst;
  if (e < count * S) {
    j = e / S;
    i = e % S;
    ex = i;
    dst = buf + jobs[j].P + i * D;
  } else {
    size_t r = e - count * S;
    // job of segment entry r: jobs are laid out with their segment counts in order
    size_t lo = 0, hi = count;
    while (hi - lo > 1) {
      const size_t mid = (lo + hi) / 2;
      if (jobs[mid].blk0 <= r) lo = mid; else hi = mid;
    }
    j = lo;
    if (r >= jobs[j].blk0 + jobs[j].nb) return;
    i = r - jobs[j].blk0;
    ex = (uint64_t)i * S;
    dst = buf + jobs[j].B + i * D;
  }
  (void)entries_per_job_p;
  Ext<D> acc = gl::ext_base<D>(1);
  for (int b = 0; ex; b++, ex >>= 1)
    if (ex & 1) acc = gl::emul<D>(acc, ld<D>(pw, (j * UNI_PW + b)));
  st<D>(dst, 0, acc);
}

template <int D>
__global__ __launch_bounds__(RT) void k_uni_many(const UniJobDev* jobs, size_t count, const uint64_t* buf,
                                                 uint64_t* parts) {
  const size_t b = blockIdx.x;
  size_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const size_t mid = (lo + hi) / 2;
    if (jobs[mid].blk0 <= b) lo = mid; else hi = mid;
  }
  const UniJobDev J = jobs[lo];
  const size_t seg = b - J.blk0;
  const size_t S = (size_t)RT * UNI_K;
  const size_t i0 = seg * S + threadIdx.x;
  Ext<D> acc[1] = {gl::ext_zero<D>()};
#pragma unroll
  for (int k = 0; k < UNI_K; k++) {
    const size_t i = i0 + (size_t)k * RT;
    if (i >= J.n) break;
    const size_t q = (size_t)k * RT + threadIdx.x;
    acc[0] = gl::eadd<D>(acc[0], gl::escale<D>(ld<D>(buf + J.P, q), J.c[i]));
  }
  acc[0] = gl::emul<D>(acc[0], ld<D>(buf + J.B, seg));
  block_reduce_store<D, 1>(acc, parts + b * D);
}

template <int D>
__global__ __launch_bounds__(RT) void k_uni_many_reduce(const UniJobDev* jobs, const uint64_t* parts,
                                                        uint64_t* out) {
  const UniJobDev J = jobs[blockIdx.x];
  Ext<D> v[1] = {gl::ext_zero<D>()};
  for (size_t p = threadIdx.x; p < J.nb; p += blockDim.x) v[0] = gl::eadd<D>(v[0], ld<D>(parts, J.blk0 + p));
  block_reduce_store<D, 1>(v, out + blockIdx.x * D);
}

template <int D>
static void uni_batch_t(const UniJob* jobs, size_t count, uint64_t* out, hipStream_t s) {
  const size_t S = (size_t)RT * UNI_K;
  std::vector<UniJobDev> dj(count);
  size_t blk = 0, words = 0;
  for (size_t j = 0; j < count; j++) {
    WHIR_REQUIRE(jobs[j].n > 0, "univariate batch: empty vector");
    dj[j].c = jobs[j].coeffs;
    dj[j].n = jobs[j].n;
    dj[j].nb = (jobs[j].n + S - 1) / S;
    WHIR_REQUIRE((uint64_t)(dj[j].nb - 1) * S < ((uint64_t)1 << UNI_PW), "univariate batch: vector too long");
    dj[j].blk0 = blk;
    blk += dj[j].nb;
  }
  for (size_t j = 0; j < count; j++) {
    dj[j].P = words;
    words += S * D;
  }
  for (size_t j = 0; j < count; j++) {
    dj[j].B = words;
    words += dj[j].nb * D;
  }
  // squarings x^(2^b), b < UNI_PW, per job
  std::vector<uint64_t> pw(count * UNI_PW * D);
  for (size_t j = 0; j < count; j++) {
    Ext<D> p;
    for (int d = 0; d < D; d++) p.c[d] = jobs[j].x[d];
    for (int b = 0; b < UNI_PW; b++) {
      for (int d = 0; d < D; d++) pw[(j * UNI_PW + b) * D + d] = p.c[d];
      p = gl::emul<D>(p, p);
    }
  }
  const size_t job_words = (count * sizeof(UniJobDev) + 7) / 8;
  DevBuf in(job_words * 8 + pw.size() * 8, s), buf(words * 8, s), parts(blk * D * 8, s);
  {
    std::vector<uint64_t> host(job_words + pw.size());
    memcpy(host.data(), dj.data(), count * sizeof(UniJobDev));
    std::copy(pw.begin(), pw.end(), host.begin() + job_words);
    upload_async(in.ptr, host.data(), host.size() * 8, s);
  }
  const UniJobDev* djobs = (const UniJobDev*)in.ptr;
  const uint64_t* dpw = in.as<uint64_t>() + job_words;
  const size_t entries = count * S + blk;
  k_uni_many_tables<D><<<blocks_for(entries, RT), RT, 0, s>>>(buf.as<uint64_t>(), djobs, dpw, count, S);
  WHIR_HIP(hipGetLastError());
  {
    ProfScope prof(kernel_name("k_uni_many", 1, D), s, 8.0 * (double)blk * S);
    k_uni_many<D><<<(unsigned)blk, RT, 0, s>>>(djobs, count, buf.as<uint64_t>(), parts.as<uint64_t>());
    WHIR_HIP(hipGetLastError());
  }
  k_uni_many_reduce<D><<<(unsigned)count, RT, 0, s>>>(djobs, parts.as<uint64_t>(), out);
  WHIR_HIP(hipGetLastError());
}

void univariate_eval_batch_dev(int D, const UniJob* jobs, size_t count, uint64_t* out, hipStream_t s) {
  if (!count) return;
  switch (D) {
    case 1: uni_batch_t<1>(jobs, count, out, s); return;
    case 2: uni_batch_t<2>(jobs, count, out, s); return;
    case 3: uni_batch_t<3>(jobs, count, out, s); return;
  }
  throw Error("univariate batch: extension degree must be 1, 2 or 3");
}

void univariate_eval_dev(int Ds, int D, const uint64_t* c, size_t n, const uint64_t* x, uint64_t* out,
                         hipStream_t s) {
  univariate_eval_many_dev(Ds, D, c, n, x, 1, out, s);
}

// ------------------------------------------------------------------ lift, axpy, dot, gather
template <int D>
__global__ void k_lift(const uint64_t* src, size_t n, uint64_t* dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    st<D>(dst, i, gl::ext_base<D>(src[i]));
}

void lift_dev(int D, const uint64_t* src, size_t n, uint64_t* dst, hipStream_t s) {
  if (!n) return;
  ProfScope prof(kernel_name("k_lift", D), s, 8.0 * (1.0 + D) * (double)n);
  switch (D) {
    case 1: k_lift<1><<<grid_for(n), RT, 0, s>>>(src, n, dst); break;
    case 2: k_lift<2><<<grid_for(n), RT, 0, s>>>(src, n, dst); break;
    case 3: k_lift<3><<<grid_for(n), RT, 0, s>>>(src, n, dst); break;
    default: throw Error("lift: extension degree must be 1, 2 or 3");
  }
  WHIR_HIP(hipGetLastError());
}

template <int D>
__global__ void k_axpy(uint64_t* acc, Ext<D> w, const uint64_t* v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    st<D>(acc, i, gl::eadd<D>(ld<D>(acc, i), gl::escale<D>(w, v[i])));
}

void mixed_scalar_mul_add_dev(int D, uint64_t* acc, const uint64_t* w, const uint64_t* v, size_t n, hipStream_t s) {
  if (!n) return;
  switch (D) {
    case 1: k_axpy<1><<<grid_for(n), RT, 0, s>>>(acc, host_ext<1>(w), v, n); break;
    case 2: k_axpy<2><<<grid_for(n), RT, 0, s>>>(acc, host_ext<2>(w), v, n); break;
    case 3: k_axpy<3><<<grid_for(n), RT, 0, s>>>(acc, host_ext<3>(w), v, n); break;
    default: throw Error("mixed_scalar_mul_add: extension degree must be 1, 2 or 3");
  }
  WHIR_HIP(hipGetLastError());
}

template <int D>
__global__ void k_axpy_ext(uint64_t* acc, Ext<D> w, const uint64_t* v, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    st<D>(acc, i, gl::eadd<D>(ld<D>(acc, i), gl::emul<D>(w, ld<D>(v, i))));
}

void scalar_mul_add_dev(int D, uint64_t* acc, const uint64_t* w, const uint64_t* v, size_t n, hipStream_t s) {
  if (!n) return;
  switch (D) {
    case 1: k_axpy_ext<1><<<grid_for(n), RT, 0, s>>>(acc, host_ext<1>(w), v, n); break;
    case 2: k_axpy_ext<2><<<grid_for(n), RT, 0, s>>>(acc, host_ext<2>(w), v, n); break;
    case 3: k_axpy_ext<3><<<grid_for(n), RT, 0, s>>>(acc, host_ext<3>(w), v, n); break;
    default: throw Error("scalar_mul_add: extension degree must be 1, 2 or 3");
  }
  WHIR_HIP(hipGetLastError());
}

template <int D>
__global__ __launch_bounds__(RT) void k_dot(const uint64_t* a, const uint64_t* b, size_t n, uint64_t* parts) {
  Ext<D> acc[1] = {gl::ext_zero<D>()};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    acc[0] = gl::eadd<D>(acc[0], gl::emul<D>(ld<D>(a, i), ld<D>(b, i)));
  block_reduce_store<D, 1>(acc, parts + (size_t)blockIdx.x * D);
}

template <int D>
static void dot_t(const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out, hipStream_t s) {
  unsigned g = grid_for(n ? n : 1);
  DevBuf parts((size_t)g * D * 8, s);
  k_dot<D><<<g, RT, 0, s>>>(a, b, n, parts.as<uint64_t>());
  WHIR_HIP(hipGetLastError());
  finish_reduce<D, 1>(parts.as<uint64_t>(), g, out, s);
}

void dot_dev(int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out, hipStream_t s) {
  switch (D) {
    case 1: dot_t<1>(a, b, n, out, s); break;
    case 2: dot_t<2>(a, b, n, out, s); break;
    case 3: dot_t<3>(a, b, n, out, s); break;
    default: throw Error("dot: extension degree must be 1, 2 or 3");
  }
}

template <int D>
__global__ void k_gather(const uint64_t* cw, size_t N, size_t M, const uint64_t* idx, size_t k, uint64_t* out) {
  size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= k * M) return;
  size_t q = e / M, j = e % M;
  size_t r = idx[q];
  if (r >= N) return;
  st<D>(out, q * M + j, ld<D>(cw, j * N + r));
}

// Opening gather in one launch: rows (out_rows[q*M + j] = cw[j*N + idx[q]], Ds words
// each) and Merkle sibling hashes (out_hashes[i] = nodes[sib[i]]); idx_sib holds the
// k row indices followed by the ns node indices.  With D > 0 it also sums each opened
// row against per-column weights of degree D (wts[j], D words):
// out_vals[q] = sum_j wts[j] cw[j*N + idx[q]].  When M is a power of two up to 64 the
// M lanes that gather a row each weight their element and sum across the row with
// XOR shuffles (the row's lanes are aligned within one wave), so no lane waits on M
// loads in turn; other widths sum each row on one lane.
template <int Ds, int D>
__global__ void k_gather_open(const uint64_t* __restrict__ cw, size_t N, size_t M, const uint64_t* __restrict__ idx_sib,
                              size_t k, const uint4* __restrict__ nodes, size_t ns, uint64_t* out_rows,
                              uint4* out_hashes, const uint64_t* __restrict__ wts, uint64_t* out_vals,
                              bool lane_sums) {
  size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e < k * M) {
    const size_t q = e / M, j = e % M, r = idx_sib[q];
    const Ext<Ds> x = r < N ? ld<Ds>(cw, j * N + r) : gl::ext_zero<Ds>();
    if (r < N) st<Ds>(out_rows, e, x);
    if constexpr (D > 0) {
      if (lane_sums) {
        Ext<D> p;
        if constexpr (Ds == 1) p = gl::escale<D>(ld<D>(wts, j), x.c[0]);
        else p = gl::emul<D>(ld<D>(wts, j), x);
        for (unsigned o = 1; o < (unsigned)M; o <<= 1) {
          Ext<D> t;
#pragma unroll
          for (int d = 0; d < D; d++) t.c[d] = __shfl_xor(p.c[d], (int)o);
          p = gl::eadd<D>(p, t);
        }
        if (j == 0) st<D>(out_vals, q, p);  // zero for an index past the codeword
      }
    }
    return;
  }
  e -= k * M;
  if (e < 2 * ns) {
    out_hashes[e] = nodes[2 * idx_sib[k + (e >> 1)] + (e & 1)];
    return;
  }
  if constexpr (D > 0) {
    e -= 2 * ns;
    if (!lane_sums && e < k) {
      const size_t r = idx_sib[e];
      Ext<D> acc = gl::ext_zero<D>();
      if (r < N) {
#pragma unroll 8
        for (size_t j = 0; j < M; j++) {
          if constexpr (Ds == 1) acc = gl::eadd<D>(acc, gl::escale<D>(ld<D>(wts, j), cw[j * N + r]));
          else acc = gl::eadd<D>(acc, gl::emul<D>(ld<D>(wts, j), ld<D>(cw, j * N + r)));
        }
      }
      st<D>(out_vals, e, acc);
    }
  }
}

template <int Ds, int D>
static void gather_open_t(const uint64_t* cw, size_t N, size_t M, const uint64_t* idx_sib, size_t k,
                          const uint8_t* nodes, size_t ns, uint64_t* out_rows, uint8_t* out_hashes,
                          const uint64_t* wts, uint64_t* out_vals, hipStream_t s) {
  const bool lane_sums = D > 0 && M >= 1 && M <= 64 && (M & (M - 1)) == 0;
  const size_t total = k * M + 2 * ns + (D > 0 && !lane_sums ? k : 0);
  if (!total) return;
  // Rows start at multiples of M within each block (RT % M == 0), so a row's lanes
  // are aligned within one wave.
  static_assert(RT % 64 == 0, "block size must be a multiple of the wave size");
  k_gather_open<Ds, D><<<blocks_for(total, RT), RT, 0, s>>>(cw, N, M, idx_sib, k, (const uint4*)nodes, ns, out_rows,
                                                              (uint4*)out_hashes, wts, out_vals, lane_sums);
  WHIR_HIP(hipGetLastError());
}

void gather_open_dev(int Ds, const uint64_t* cw, size_t N, size_t M, const uint64_t* idx_sib, size_t k,
                     const uint8_t* nodes, size_t ns, uint64_t* out_rows, uint8_t* out_hashes, hipStream_t s, int D,
                     const uint64_t* wts, uint64_t* out_vals) {
  WHIR_REQUIRE(ns == 0 || ((uintptr_t)out_hashes & 15) == 0, "gather_open: hash output must be 16-byte aligned");
  WHIR_REQUIRE(D == 0 || (wts && out_vals), "gather_open: row sums need weights and an output");
  const int key = Ds * 10 + D;
  switch (key) {
    case 10: gather_open_t<1, 0>(cw, N, M, idx_sib, k, nodes, ns, out_rows, out_hashes, wts, out_vals, s); return;
    case 20: gather_open_t<2, 0>(cw, N, M, idx_sib, k, nodes, ns, out_rows, out_hashes, wts, out_vals, s); return;
    case 30: gather_open_t<3, 0>(cw, N, M, idx_sib, k, nodes, ns, out_rows, out_hashes, wts, out_vals, s); return;
    case 11: gather_open_t<1, 1>(cw, N, M, idx_sib, k, nodes, ns, out_rows, out_hashes, wts, out_vals, s); return;
    case 12: gather_open_t<1, 2>(cw, N, M, idx_sib, k, nodes, ns, out_rows, out_hashes, wts, out_vals, s); return;
    case 13: gather_open_t<1, 3>(cw, N, M, idx_sib, k, nodes, ns, out_rows, out_hashes, wts, out_vals, s); return;
    case 22: gather_open_t<2, 2>(cw, N, M, idx_sib, k, nodes, ns, out_rows, out_hashes, wts, out_vals, s); return;
    case 33: gather_open_t<3, 3>(cw, N, M, idx_sib, k, nodes, ns, out_rows, out_hashes, wts, out_vals, s); return;
    default: throw Error("gather: unsupported row / weight degrees");
  }
}

void gather_rows_dev(int D, const uint64_t* cw, size_t N, size_t M, const uint64_t* idx, size_t k, uint64_t* out,
                     hipStream_t s) {
  if (!k || !M) return;
  unsigned g = blocks_for(k * M, RT);
  switch (D) {
    case 1: k_gather<1><<<g, RT, 0, s>>>(cw, N, M, idx, k, out); break;
    case 2: k_gather<2><<<g, RT, 0, s>>>(cw, N, M, idx, k, out); break;
    case 3: k_gather<3><<<g, RT, 0, s>>>(cw, N, M, idx, k, out); break;
    default: throw Error("gather: extension degree must be 1, 2 or 3");
  }
  WHIR_HIP(hipGetLastError());
}

}  // namespace whir
