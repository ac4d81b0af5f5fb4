// Sumcheck prover round and covector builders on CDNA4.
//
// All of these are streaming passes over HBM with exact Goldilocks arithmetic,
// so their order of summation does not change any result.  Reductions run as
// grid-stride partial sums per block (LDS tree) followed by one small block.
#include "algebra.h"
#include "common.h"
#include "gl.h"

#include <algorithm>
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
template <int D>
__device__ __forceinline__ void st(uint64_t* p, size_t i, const Ext<D>& v) {
#pragma unroll
  for (int d = 0; d < D; d++) p[i * D + d] = v.c[d];
}
template <int D>
__device__ __forceinline__ Ext<D> ld_host(const uint64_t* const* unused, const Ext<D>& v) { return v; }

// Sum NV extension values over the block; thread 0 writes them to out.
template <int D, int NV>
__device__ void block_reduce_store(Ext<D> v[NV], uint64_t* out) {
  __shared__ uint64_t red[RT * NV * D];
#pragma unroll
  for (int k = 0; k < NV; k++)
#pragma unroll
    for (int d = 0; d < D; d++) red[(threadIdx.x * NV + k) * D + d] = v[k].c[d];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k < NV; k++)
#pragma unroll
        for (int d = 0; d < D; d++) {
          size_t a = (threadIdx.x * NV + k) * D + d, b = ((threadIdx.x + s) * NV + k) * D + d;
          red[a] = gl::add(red[a], red[b]);
        }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < NV * D; k++) out[k] = red[k];
}

// Final reduction of `nparts` partials of NV*D words each.
template <int D, int NV>
__global__ __launch_bounds__(RT) void k_reduce_parts(const uint64_t* parts, size_t nparts, uint64_t* out) {
  Ext<D> v[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) v[k] = gl::ext_zero<D>();
  for (size_t p = threadIdx.x; p < nparts; p += blockDim.x)
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = gl::eadd<D>(v[k], ld<D>(parts + p * NV * D, k));
  block_reduce_store<D, NV>(v, out);
}

static unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>(blocks_for(n, RT), MAX_BLOCKS); }

template <int D, int NV>
static void finish_reduce(const uint64_t* parts, unsigned nparts, uint64_t* out, hipStream_t s) {
  k_reduce_parts<D, NV><<<1, RT, 0, s>>>(parts, nparts, out);
  WHIR_HIP(hipGetLastError());
}

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
template <int D>
__global__ __launch_bounds__(RT) void k_sc_poly(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                                                size_t half, size_t nhi, uint64_t* parts) {
  Ext<D> acc[2] = {gl::ext_zero<D>(), gl::ext_zero<D>()};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < half; i += (size_t)gridDim.x * blockDim.x) {
    Ext<D> a0 = ld<D>(a, i), b0 = ld<D>(b, i);
    Ext<D> p0 = gl::emul<D>(a0, b0);
    acc[0] = gl::eadd<D>(acc[0], p0);
    if (i < nhi) {
      Ext<D> da = gl::esub<D>(ld<D>(a, half + i), a0), db = gl::esub<D>(ld<D>(b, half + i), b0);
      acc[1] = gl::eadd<D>(acc[1], gl::emul<D>(da, db));
    } else {
      acc[1] = gl::eadd<D>(acc[1], p0);  // implicit zero padding: (0 - a)(0 - b) = a b
    }
  }
  block_reduce_store<D, 2>(acc, parts + blockIdx.x * 2 * D);
}

template <int D>
__global__ void k_sc_single(const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out) {
  if (threadIdx.x != 0) return;
  Ext<D> c0 = n ? gl::emul<D>(ld<D>(a, 0), ld<D>(b, 0)) : gl::ext_zero<D>();
  st<D>(out, 0, c0);
  st<D>(out, 1, gl::ext_zero<D>());
}

template <int D>
static void sc_poly(const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out, hipStream_t s) {
  if (n <= 1) {
    k_sc_single<D><<<1, 64, 0, s>>>(a, b, n, out);
    WHIR_HIP(hipGetLastError());
    return;
  }
  size_t half = next_pow2(n) >> 1, nhi = n - half;
  unsigned g = grid_for(half);
  DevBuf parts((size_t)g * 2 * D * 8, s);
  ProfScope prof("sumcheck_poly", s, 16.0 * D * (double)n);
  k_sc_poly<D><<<g, RT, 0, s>>>(a, b, half, nhi, parts.as<uint64_t>());
  WHIR_HIP(hipGetLastError());
  finish_reduce<D, 2>(parts.as<uint64_t>(), g, out, s);
}

void sumcheck_poly_dev(int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out, hipStream_t s) {
  switch (D) {
    case 1: sc_poly<1>(a, b, n, out, s); break;
    case 2: sc_poly<2>(a, b, n, out, s); break;
    case 3: sc_poly<3>(a, b, n, out, s); break;
    default: throw Error("sumcheck: extension degree must be 1, 2 or 3");
  }
}

// ------------------------------------------------------------------ fold
template <int D>
__global__ __launch_bounds__(RT) void k_fold(uint64_t* v, size_t half, size_t nhi, Ext<D> w, Ext<D> one_minus_w) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < half; i += (size_t)gridDim.x * blockDim.x) {
    Ext<D> lo = ld<D>(v, i);
    if (i < nhi) {
      Ext<D> hi = ld<D>(v, half + i);
      st<D>(v, i, gl::eadd<D>(lo, gl::emul<D>(gl::esub<D>(hi, lo), w)));
    } else {
      st<D>(v, i, gl::emul<D>(lo, one_minus_w));
    }
  }
}

template <int D>
static size_t fold_t(uint64_t* v, size_t n, const uint64_t* w, hipStream_t s) {
  if (n <= 1) return n;
  size_t half = next_pow2(n) >> 1, nhi = n - half;
  Ext<D> ww = host_ext<D>(w);
  Ext<D> om = gl::esub<D>(gl::ext_base<D>(1), ww);
  ProfScope prof("fold", s, 8.0 * D * ((double)n + (double)half));
  k_fold<D><<<grid_for(half), RT, 0, s>>>(v, half, nhi, ww, om);
  WHIR_HIP(hipGetLastError());
  return half;
}

size_t fold_dev(int D, uint64_t* v, size_t n, const uint64_t* w, hipStream_t s) {
  switch (D) {
    case 1: return fold_t<1>(v, n, w, s);
    case 2: return fold_t<2>(v, n, w, s);
    case 3: return fold_t<3>(v, n, w, s);
  }
  throw Error("fold: extension degree must be 1, 2 or 3");
}

// ------------------------------------------------------------------ fused fold + polynomial
// Folded length h = next_pow2(n)/2 (a power of two >= 2); new pairs (i, h/2 + i).
template <int D>
__device__ __forceinline__ Ext<D> fold_at(const uint64_t* v, size_t i, size_t h, size_t nhi, const Ext<D>& w,
                                          const Ext<D>& om) {
  Ext<D> lo = ld<D>(v, i);
  if (i < nhi) return gl::eadd<D>(lo, gl::emul<D>(gl::esub<D>(ld<D>(v, h + i), lo), w));
  return gl::emul<D>(lo, om);
}

template <int D>
__global__ __launch_bounds__(RT) void k_fold_poly(uint64_t* a, uint64_t* b, size_t h, size_t nhi, Ext<D> w,
                                                  Ext<D> om, uint64_t* parts) {
  Ext<D> acc[2] = {gl::ext_zero<D>(), gl::ext_zero<D>()};
  const size_t q = h >> 1;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < q; i += (size_t)gridDim.x * blockDim.x) {
    Ext<D> a0 = fold_at<D>(a, i, h, nhi, w, om), a1 = fold_at<D>(a, q + i, h, nhi, w, om);
    Ext<D> b0 = fold_at<D>(b, i, h, nhi, w, om), b1 = fold_at<D>(b, q + i, h, nhi, w, om);
    st<D>(a, i, a0); st<D>(a, q + i, a1);
    st<D>(b, i, b0); st<D>(b, q + i, b1);
    acc[0] = gl::eadd<D>(acc[0], gl::emul<D>(a0, b0));
    acc[1] = gl::eadd<D>(acc[1], gl::emul<D>(gl::esub<D>(a1, a0), gl::esub<D>(b1, b0)));
  }
  block_reduce_store<D, 2>(acc, parts + blockIdx.x * 2 * D);
}

template <int D>
static size_t fold_poly_t(uint64_t* a, uint64_t* b, size_t n, const uint64_t* w, uint64_t* out, hipStream_t s) {
  if (n <= 2) {
    // folded length <= 1: fold then the one-element polynomial
    size_t m = fold_t<D>(a, n, w, s);
    fold_t<D>(b, n, w, s);
    sc_poly<D>(a, b, m, out, s);
    return m;
  }
  size_t h = next_pow2(n) >> 1, nhi = n - h;
  Ext<D> ww = host_ext<D>(w);
  Ext<D> om = gl::esub<D>(gl::ext_base<D>(1), ww);
  unsigned g = grid_for(h >> 1);
  DevBuf parts((size_t)g * 2 * D * 8, s);
  ProfScope prof("fold_sumcheck_poly", s, 8.0 * D * (2.0 * n + 2.0 * h));
  k_fold_poly<D><<<g, RT, 0, s>>>(a, b, h, nhi, ww, om, parts.as<uint64_t>());
  WHIR_HIP(hipGetLastError());
  finish_reduce<D, 2>(parts.as<uint64_t>(), g, out, s);
  return h;
}

size_t fold_and_poly_dev(int D, uint64_t* a, uint64_t* b, size_t n, const uint64_t* w, uint64_t* out,
                         hipStream_t s) {
  switch (D) {
    case 1: return fold_poly_t<1>(a, b, n, w, out, s);
    case 2: return fold_poly_t<2>(a, b, n, w, out, s);
    case 3: return fold_poly_t<3>(a, b, n, w, out, s);
  }
  throw Error("fold: extension degree must be 1, 2 or 3");
}

// ------------------------------------------------------------------ eval_eq
// eq(z, x) = eq_hi(x >> k2) * eq_lo(x & (2^k2 - 1)); the tables are built on the
// host (2^k1 + 2^k2 entries) and the scalar is folded into eq_hi.
template <int D>
__global__ __launch_bounds__(RT) void k_eval_eq(uint64_t* acc, size_t n, const uint64_t* hi, const uint64_t* lo,
                                                int k2) {
  const size_t mask = ((size_t)1 << k2) - 1;
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) {
    Ext<D> e = gl::emul<D>(ld<D>(hi, x >> k2), ld<D>(lo, x & mask));
    st<D>(acc, x, gl::eadd<D>(ld<D>(acc, x), e));
  }
}

// Host table: t[x] = s * prod_{i<k} (x_i ? z_i : 1 - z_i), x_0 = MSB of x.
template <int D>
static std::vector<uint64_t> eq_table(const uint64_t* z, size_t k, Ext<D> s) {
  std::vector<uint64_t> t(((size_t)1 << k) * D);
  std::vector<Ext<D>> cur{s};
  for (size_t i = 0; i < k; i++) {
    Ext<D> zi = host_ext<D>(z + i * D);
    Ext<D> om = gl::esub<D>(gl::ext_base<D>(1), zi);
    std::vector<Ext<D>> nxt(cur.size() * 2);
    for (size_t x = 0; x < cur.size(); x++) {
      nxt[2 * x] = gl::emul<D>(cur[x], om);
      nxt[2 * x + 1] = gl::emul<D>(cur[x], zi);
    }
    cur.swap(nxt);
  }
  for (size_t x = 0; x < cur.size(); x++)
    for (int d = 0; d < D; d++) t[x * D + d] = cur[x].c[d];
  return t;
}

template <int D>
static void eval_eq_t(uint64_t* acc, const uint64_t* z, size_t k, const uint64_t* scalar, hipStream_t s) {
  size_t k1 = k / 2, k2 = k - k1;
  auto hi = eq_table<D>(z, k1, host_ext<D>(scalar));
  auto lo = eq_table<D>(z + k1 * D, k2, gl::ext_base<D>(1));
  DevBuf dh(hi.size() * 8, s), dl(lo.size() * 8, s);
  upload_async(dh.ptr, hi.data(), hi.size() * 8, s);
  upload_async(dl.ptr, lo.data(), lo.size() * 8, s);
  size_t n = (size_t)1 << k;
  ProfScope prof("eval_eq", s, 16.0 * D * (double)n);
  k_eval_eq<D><<<grid_for(n), RT, 0, s>>>(acc, n, dh.as<uint64_t>(), dl.as<uint64_t>(), (int)k2);
  WHIR_HIP(hipGetLastError());
}

void eval_eq_dev(int D, uint64_t* acc, const uint64_t* point, size_t k, const uint64_t* scalar, hipStream_t s) {
  switch (D) {
    case 1: eval_eq_t<1>(acc, point, k, scalar, s); break;
    case 2: eval_eq_t<2>(acc, point, k, scalar, s); break;
    case 3: eval_eq_t<3>(acc, point, k, scalar, s); break;
    default: throw Error("eval_eq: extension degree must be 1, 2 or 3");
  }
}

// ------------------------------------------------------------------ geometric accumulate
// Block b covers [b*RT*K, (b+1)*RT*K); lane t owns i = b*RT*K + t + k*RT.
// For each point j: p = s_j * x_j^(b*RT*K) * x_j^t, then acc_k += p, p *= x_j^RT.
// xb[j*nb + b] = s_j x_j^(b RT K), xt[j*RT + t] = x_j^t, xs[j] = x_j^RT (host tables).
static constexpr int GA_K = 16;

// Points are split into extension points (first m_e of xb; tables xt_e, xs_e in
// the extension) and base-field points (next m_b; tables xt_b, xs_b in the base
// field, stepped with D base multiplications instead of a full extension product).
// The in-domain STIR points are all base-field elements.
template <int D>
__global__ __launch_bounds__(RT) void k_geom(uint64_t* acc, size_t n, const uint64_t* xb, const uint64_t* xt_e,
                                             const uint64_t* xs_e, const uint64_t* xt_b, const uint64_t* xs_b,
                                             size_t m_e, size_t m_b, size_t nb) {
  const size_t b = blockIdx.x;
  const size_t i0 = b * (size_t)RT * GA_K + threadIdx.x;
  Ext<D> sum[GA_K];
#pragma unroll
  for (int k = 0; k < GA_K; k++) sum[k] = gl::ext_zero<D>();
  for (size_t j = 0; j < m_e; j++) {
    Ext<D> p = gl::emul<D>(ld<D>(xb, j * nb + b), ld<D>(xt_e, j * RT + threadIdx.x));
    Ext<D> step = ld<D>(xs_e, j);
#pragma unroll
    for (int k = 0; k < GA_K; k++) {
      sum[k] = gl::eadd<D>(sum[k], p);
      p = gl::emul<D>(p, step);
    }
  }
  for (size_t j = 0; j < m_b; j++) {
    Ext<D> p = gl::escale<D>(ld<D>(xb, (m_e + j) * nb + b), xt_b[j * RT + threadIdx.x]);
    const uint64_t step = xs_b[j];
#pragma unroll
    for (int k = 0; k < GA_K; k++) {
      sum[k] = gl::eadd<D>(sum[k], p);
      p = gl::escale<D>(p, step);
    }
  }
#pragma unroll
  for (int k = 0; k < GA_K; k++) {
    size_t i = i0 + (size_t)k * RT;
    if (i < n) st<D>(acc, i, gl::eadd<D>(ld<D>(acc, i), sum[k]));
  }
}

template <int D>
static void geom_t(uint64_t* acc, size_t n, const uint64_t* scalars, const uint64_t* points, size_t m,
                   hipStream_t s) {
  if (!n || !m) return;
  size_t span = (size_t)RT * GA_K;
  size_t nb = (n + span - 1) / span;
  // Extension points first, then points that lie in the base field.
  std::vector<size_t> order_e, order_b;
  for (size_t j = 0; j < m; j++) {
    bool base = true;
    for (int d = 1; d < D; d++) base = base && points[j * D + d] == 0;
    (base && D > 1 ? order_b : order_e).push_back(j);
  }
  const size_t m_e = order_e.size(), m_b = order_b.size();
  std::vector<uint64_t> xb(m * nb * D), xt_e(std::max<size_t>(m_e, 1) * RT * D), xs_e(std::max<size_t>(m_e, 1) * D);
  std::vector<uint64_t> xt_b(std::max<size_t>(m_b, 1) * RT), xs_b(std::max<size_t>(m_b, 1));
  size_t slot = 0;
  for (size_t q = 0; q < m_e + m_b; q++, slot++) {
    const bool is_base = q >= m_e;
    const size_t j = is_base ? order_b[q - m_e] : order_e[q];
    Ext<D> x = host_ext<D>(points + j * D);
    Ext<D> sj = host_ext<D>(scalars + j * D);
    Ext<D> xspan = gl::epow<D>(x, span);
    Ext<D> cur = sj;
    for (size_t b = 0; b < nb; b++) {
      for (int d = 0; d < D; d++) xb[(slot * nb + b) * D + d] = cur.c[d];
      cur = gl::emul<D>(cur, xspan);
    }
    Ext<D> pw = gl::ext_base<D>(1);
    for (int t = 0; t < RT; t++) {
      if (is_base) xt_b[(q - m_e) * RT + t] = pw.c[0];
      else
        for (int d = 0; d < D; d++) xt_e[(q * RT + t) * D + d] = pw.c[d];
      pw = gl::emul<D>(pw, x);
    }
    if (is_base) xs_b[q - m_e] = pw.c[0];  // x^RT
    else
      for (int d = 0; d < D; d++) xs_e[q * D + d] = pw.c[d];
  }
  DevBuf db(xb.size() * 8, s), dte(xt_e.size() * 8, s), dse(xs_e.size() * 8, s), dtb(xt_b.size() * 8, s),
      dsb(xs_b.size() * 8, s);
  upload_async(db.ptr, xb.data(), xb.size() * 8, s);
  upload_async(dte.ptr, xt_e.data(), xt_e.size() * 8, s);
  upload_async(dse.ptr, xs_e.data(), xs_e.size() * 8, s);
  upload_async(dtb.ptr, xt_b.data(), xt_b.size() * 8, s);
  upload_async(dsb.ptr, xs_b.data(), xs_b.size() * 8, s);
  ProfScope prof("geometric_accumulate", s, 16.0 * D * (double)n);
  k_geom<D><<<(unsigned)nb, RT, 0, s>>>(acc, n, db.as<uint64_t>(), dte.as<uint64_t>(), dse.as<uint64_t>(),
                                        dtb.as<uint64_t>(), dsb.as<uint64_t>(), m_e, m_b, nb);
  WHIR_HIP(hipGetLastError());
}

void geometric_accumulate_dev(int D, uint64_t* acc, size_t n, const uint64_t* scalars, const uint64_t* points,
                              size_t m, hipStream_t s) {
  switch (D) {
    case 1: geom_t<1>(acc, n, scalars, points, m, s); break;
    case 2: geom_t<2>(acc, n, scalars, points, m, s); break;
    case 3: geom_t<3>(acc, n, scalars, points, m, s); break;
    default: throw Error("geometric_accumulate: extension degree must be 1, 2 or 3");
  }
}

// ------------------------------------------------------------------ univariate evaluation
// Same lane layout as geometric_accumulate: lane partial = sum_k c_i x^i.
template <int Ds, int D>
__global__ __launch_bounds__(RT) void k_uni(const uint64_t* c, size_t n, Ext<D> xbase_step, const uint64_t* xb,
                                            const uint64_t* xt, Ext<D> step, uint64_t* parts) {
  const size_t b = blockIdx.x;
  const size_t i0 = b * (size_t)RT * GA_K + threadIdx.x;
  Ext<D> p = gl::emul<D>(ld<D>(xb, b), ld<D>(xt, threadIdx.x));
  Ext<D> acc[1] = {gl::ext_zero<D>()};
#pragma unroll
  for (int k = 0; k < GA_K; k++) {
    size_t i = i0 + (size_t)k * RT;
    if (i < n) {
      if (Ds == 1) {
        acc[0] = gl::eadd<D>(acc[0], gl::escale<D>(p, c[i]));
      } else {
        acc[0] = gl::eadd<D>(acc[0], gl::emul<D>(p, ld<D>(c, i)));
      }
    }
    p = gl::emul<D>(p, step);
  }
  (void)xbase_step;
  block_reduce_store<D, 1>(acc, parts + b * D);
}

template <int Ds, int D>
static void uni_t(const uint64_t* c, size_t n, const uint64_t* x, uint64_t* out, hipStream_t s) {
  if (!n) {
    std::vector<uint64_t> z(D, 0);
    upload_async(out, z.data(), D * 8, s);
    return;
  }
  size_t span = (size_t)RT * GA_K;
  size_t nb = (n + span - 1) / span;
  Ext<D> xx = host_ext<D>(x);
  Ext<D> xspan = gl::epow<D>(xx, span);
  std::vector<uint64_t> xb(nb * D), xt(RT * D);
  Ext<D> cur = gl::ext_base<D>(1);
  for (size_t b = 0; b < nb; b++) {
    for (int d = 0; d < D; d++) xb[b * D + d] = cur.c[d];
    cur = gl::emul<D>(cur, xspan);
  }
  Ext<D> pw = gl::ext_base<D>(1);
  for (int t = 0; t < RT; t++) {
    for (int d = 0; d < D; d++) xt[t * D + d] = pw.c[d];
    pw = gl::emul<D>(pw, xx);
  }
  DevBuf db(xb.size() * 8, s), dt(xt.size() * 8, s), parts(nb * D * 8, s);
  upload_async(db.ptr, xb.data(), xb.size() * 8, s);
  upload_async(dt.ptr, xt.data(), xt.size() * 8, s);
  {
    ProfScope prof("univariate_eval", s, 8.0 * Ds * (double)n);
    k_uni<Ds, D><<<(unsigned)nb, RT, 0, s>>>(c, n, xspan, db.as<uint64_t>(), dt.as<uint64_t>(), pw,
                                             parts.as<uint64_t>());
    WHIR_HIP(hipGetLastError());
  }
  finish_reduce<D, 1>(parts.as<uint64_t>(), (unsigned)nb, out, s);
}

void univariate_eval_dev(int Ds, int D, const uint64_t* c, size_t n, const uint64_t* x, uint64_t* out,
                         hipStream_t s) {
  if (Ds == 1) {
    switch (D) {
      case 1: uni_t<1, 1>(c, n, x, out, s); return;
      case 2: uni_t<1, 2>(c, n, x, out, s); return;
      case 3: uni_t<1, 3>(c, n, x, out, s); return;
    }
  } else if (Ds == D) {
    switch (D) {
      case 2: uni_t<2, 2>(c, n, x, out, s); return;
      case 3: uni_t<3, 3>(c, n, x, out, s); return;
    }
  }
  throw Error("univariate_evaluate: unsupported degrees");
}

// ------------------------------------------------------------------ lift, axpy, dot, gather
template <int D>
__global__ void k_lift(const uint64_t* src, size_t n, uint64_t* dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    st<D>(dst, i, gl::ext_base<D>(src[i]));
}

void lift_dev(int D, const uint64_t* src, size_t n, uint64_t* dst, hipStream_t s) {
  if (!n) return;
  ProfScope prof("lift", s, 8.0 * (1.0 + D) * (double)n);
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
__global__ __launch_bounds__(RT) void k_dot(const uint64_t* a, const uint64_t* b, size_t n, uint64_t* parts) {
  Ext<D> acc[1] = {gl::ext_zero<D>()};
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    acc[0] = gl::eadd<D>(acc[0], gl::emul<D>(ld<D>(a, i), ld<D>(b, i)));
  block_reduce_store<D, 1>(acc, parts + blockIdx.x * D);
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
