// Reed-Solomon expansion (interleaved coset NTT) for CDNA4.
//
// Replaces NttEngine::interleaved_encode (src/algebra/ntt/cooley_tukey.rs:411-477).
// For message j (length L, zero padded to the coset size C = next_order(L)) and
// codeword index r = c*C + t (coset-major, the reference's default order):
//
//     M[r][j] = sum_i m_j[i] * w_N^{(c + t*Q) i} = DFT_C( m_j[i] * w_N^{c i} )[t],
//
// with w_N = 7^((p-1)/N) and Q = N/C.  The device layout of the codeword is
// column-major (message-major): cw[j*N + r].  Leaf hashing reads it in tiles of
// whole rows, so the row-major matrix only has to be formed for host callers.
//
// DFT_C of length C = 2^n is a four-step transform over C = C1 * C2:
//   pass 1, per column i1: Y[t2][i1] = w_C^{i1 t2} * sum_i2 x[i1 + C1 i2] w_C2^{i2 t2}
//   pass 2, per row t2:    X[t2 + C2 t1] = sum_i1 Y[t2][i1] w_C1^{i1 t1}
// Each pass stages a tile of whole columns/rows in LDS and runs an in-place
// radix-2 decimation-in-frequency transform there; the bit-reversed result is
// read back in natural order on the way out.  The coset twist w_N^{c i} is fused
// into the loads of pass 1 (or of the single pass when C <= 2^11).
#include "common.h"
#include "gl.h"
#include "ntt.h"

#include <map>
#include <mutex>
#include <vector>

namespace whir {

using gl::Ext;

static constexpr int NTT_THREADS = 256;
static constexpr int TILE_WORDS = 4096;   // u64 words of LDS per tile (32 KiB); a tile
                                          // of one 2^11 Fp3 column (6144 words) still fits 64 KiB
static constexpr int MAX_PASS_LOG = 11;   // largest DFT staged in one pass

// --------------------------------------------------------------------------
// Twiddle tables: for n = 2^k, lo[j] = w_n^j (j < 2^h), hi[j] = w_n^{j 2^h},
// h = ceil(k/2).  w_n^e = lo[e & (2^h-1)] * hi[e >> h].  Also a flat table
// w_S^j, j < S/2, for the in-LDS transforms.
// --------------------------------------------------------------------------
struct TwTable {
  const uint64_t* lo = nullptr;
  const uint64_t* hi = nullptr;
  int h = 0;
};

namespace {
std::mutex g_tw_mu;
std::map<std::pair<int, int>, TwTable> g_tw;                    // (device, log n)
std::map<std::pair<int, int>, const uint64_t*> g_flat;          // (device, log S)

const uint64_t* upload_persistent(const std::vector<uint64_t>& v) {
  void* p = nullptr;
  WHIR_HIP(hipMalloc(&p, std::max<size_t>(v.size(), 1) * 8));
  WHIR_HIP(hipMemcpy(p, v.data(), v.size() * 8, hipMemcpyHostToDevice));
  return (const uint64_t*)p;
}
}  // namespace

TwTable twiddles(int logn) {
  int dev = 0;
  WHIR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_tw_mu);
  auto key = std::make_pair(dev, logn);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  int h = (logn + 1) / 2;
  uint64_t w = gl::root_of_unity(1ull << logn);
  std::vector<uint64_t> lo(1ull << h), hi(1ull << (logn - h));
  uint64_t x = 1;
  for (auto& v : lo) { v = x; x = gl::mul(x, w); }
  uint64_t wh = gl::pow(w, 1ull << h);
  x = 1;
  for (auto& v : hi) { v = x; x = gl::mul(x, wh); }
  TwTable t;
  t.lo = upload_persistent(lo);
  t.hi = upload_persistent(hi);
  t.h = h;
  g_tw[key] = t;
  return t;
}

static const uint64_t* flat_twiddles(int logS) {
  int dev = 0;
  WHIR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_tw_mu);
  auto key = std::make_pair(dev, logS);
  auto it = g_flat.find(key);
  if (it != g_flat.end()) return it->second;
  size_t half = logS ? (1ull << (logS - 1)) : 1;
  std::vector<uint64_t> v(half);
  uint64_t w = gl::root_of_unity(1ull << logS);
  uint64_t x = 1;
  for (auto& e : v) { e = x; x = gl::mul(x, w); }
  const uint64_t* p = upload_persistent(v);
  g_flat[key] = p;
  return p;
}

__device__ __forceinline__ uint64_t tw_at(const TwTable& t, uint64_t e) {
  uint64_t mask = (1ull << t.h) - 1;
  return gl::mul(t.lo[e & mask], t.hi[e >> t.h]);
}

// --------------------------------------------------------------------------
// In-LDS radix-2 DIF over `cols` columns of length S = 2^logS stored
// s[col*S + idx]; output X[k] lands at index bitrev(k).  tw[j] = w_S^j.
// --------------------------------------------------------------------------
template <int D>
__device__ void lds_dif(Ext<D>* s, const uint64_t* tw, int logS, int cols) {
  const int S = 1 << logS;
  const int nb = cols * (S >> 1);
  for (int lh = logS - 1; lh >= 0; lh--) {
    const int h = 1 << lh;
    const int tstep = S >> (lh + 1);  // twiddle stride: w_{2h}^k = w_S^{k * S/(2h)}
    for (int q = threadIdx.x; q < nb; q += blockDim.x) {
      int col = q >> (logS - 1);
      int p = q & ((S >> 1) - 1);
      int k = p & (h - 1);
      int i = ((p >> lh) << (lh + 1)) + k;
      Ext<D>* base = s + col * S;
      Ext<D> u = base[i], v = base[i + h];
      base[i] = gl::eadd<D>(u, v);
      base[i + h] = gl::escale<D>(gl::esub<D>(u, v), tw[k * tstep]);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ int bitrev(int x, int bits) {
  return bits ? (int)(__brev((unsigned)x) >> (32 - bits)) : 0;
}

// Coefficient i of the masked message: message || mask || zeros.
template <int D>
__device__ __forceinline__ Ext<D> coeff_at(const uint64_t* msg, size_t L, const uint64_t* mask,
                                           size_t ml, size_t i) {
  Ext<D> r;
  if (i < L) {
#pragma unroll
    for (int d = 0; d < D; d++) r.c[d] = msg[i * D + d];
  } else if (i < L + ml) {
#pragma unroll
    for (int d = 0; d < D; d++) r.c[d] = mask[(i - L) * D + d];
  } else {
    r = gl::ext_zero<D>();
  }
  return r;
}

// --------------------------------------------------------------------------
// Single pass: C = 2^logC <= 2^11.  Each block transforms G sequences.
// --------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_single(RsArgs a, TwTable twN, const uint64_t* twS,
                                                            int G) {
  extern __shared__ uint64_t smem[];
  const int logC = a.logC;
  const int C = 1 << logC;
  Ext<D>* s = (Ext<D>*)smem;
  uint64_t* tws = smem + (size_t)G * C * D;
  for (int k = threadIdx.x; k < (C >> 1); k += blockDim.x) tws[k] = twS[k];
  const size_t nseq = (size_t)a.M * a.Q;
  const size_t seq0 = (size_t)blockIdx.x * G;
  for (int e = threadIdx.x; e < G * C; e += blockDim.x) {
    int g = e >> logC, i = e & (C - 1);
    size_t seq = seq0 + g;
    Ext<D> v = gl::ext_zero<D>();
    if (seq < nseq) {
      size_t j = seq / a.Q, c = seq % a.Q;
      v = coeff_at<D>(a.msgs + j * a.msg_stride * D, a.L, a.masks ? a.masks + j * a.mask_len * D : nullptr,
                      a.mask_len, i);
      uint64_t ex = ((uint64_t)c * (uint64_t)i) & (a.N - 1);
      if (ex) v = gl::escale<D>(v, tw_at(twN, ex));
    }
    s[e] = v;
  }
  __syncthreads();
  lds_dif<D>(s, tws, logC, G);
  for (int e = threadIdx.x; e < G * C; e += blockDim.x) {
    int g = e >> logC, t = e & (C - 1);
    size_t seq = seq0 + g;
    if (seq >= nseq) continue;
    size_t j = seq / a.Q, c = seq % a.Q;
    Ext<D> v = s[(g << logC) + bitrev(t, logC)];
    uint64_t* o = a.out + (j * a.N + c * C + t) * D;
#pragma unroll
    for (int d = 0; d < D; d++) o[d] = v.c[d];
  }
}

// --------------------------------------------------------------------------
// Pass 1: columns.  Block = (b consecutive i1, coset c, message j); the Q blocks
// of one column tile read the same input, which L2 / MALL serve after the first.
// Writes Y (same shape as the codeword) at [j][c*C + t2*C1 + i1].
// --------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_pass1(RsArgs a, TwTable twN, TwTable twC,
                                                           const uint64_t* twS, uint64_t* Y, int logC1, int b) {
  extern __shared__ uint64_t smem[];
  const int logC2 = a.logC - logC1;
  const int C1 = 1 << logC1, C2 = 1 << logC2;
  const size_t C = (size_t)1 << a.logC;
  Ext<D>* s = (Ext<D>*)smem;
  uint64_t* tws = smem + (size_t)b * C2 * D;
  for (int k = threadIdx.x; k < (C2 >> 1); k += blockDim.x) tws[k] = twS[k];
  const int i1_0 = blockIdx.x * b;
  const size_t c = blockIdx.y, j = blockIdx.z;
  const uint64_t* msg = a.msgs + j * a.msg_stride * D;
  const uint64_t* mask = a.masks ? a.masks + j * a.mask_len * D : nullptr;
  const int tile = b * C2;
  {
    for (int e = threadIdx.x; e < tile; e += blockDim.x) {
      int col = e % b, i2 = e / b;
      size_t i = (size_t)(i1_0 + col) + (size_t)C1 * i2;
      Ext<D> v = coeff_at<D>(msg, a.L, mask, a.mask_len, i);
      uint64_t ex = ((uint64_t)c * (uint64_t)i) & (a.N - 1);
      if (ex) v = gl::escale<D>(v, tw_at(twN, ex));
      s[col * C2 + i2] = v;
    }
    __syncthreads();
    lds_dif<D>(s, tws, logC2, b);
    uint64_t* yj = Y + (j * a.N + c * C) * D;
    for (int e = threadIdx.x; e < tile; e += blockDim.x) {
      int col = e % b, t2 = e / b;
      int i1 = i1_0 + col;
      Ext<D> v = s[col * C2 + bitrev(t2, logC2)];
      uint64_t ex = (uint64_t)i1 * (uint64_t)t2;
      if (ex) v = gl::escale<D>(v, tw_at(twC, ex));
      uint64_t* o = yj + ((size_t)t2 * C1 + i1) * D;
#pragma unroll
      for (int d = 0; d < D; d++) o[d] = v.c[d];
    }
  }
}

// --------------------------------------------------------------------------
// Pass 2: rows.  Block = (b consecutive t2, coset c, message j).
// --------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_pass2(RsArgs a, const uint64_t* twS, const uint64_t* Y,
                                                           int logC1, int b) {
  extern __shared__ uint64_t smem[];
  const int logC2 = a.logC - logC1;
  const int C1 = 1 << logC1, C2 = 1 << logC2;
  const size_t C = (size_t)1 << a.logC;
  Ext<D>* s = (Ext<D>*)smem;
  uint64_t* tws = smem + (size_t)b * C1 * D;
  for (int k = threadIdx.x; k < (C1 >> 1); k += blockDim.x) tws[k] = twS[k];
  const int t2_0 = blockIdx.x * b;
  const size_t c = blockIdx.y, j = blockIdx.z;
  const uint64_t* yj = Y + (j * a.N + c * C) * D;
  const int tile = b * C1;
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    int row = e >> logC1, i1 = e & (C1 - 1);
    const uint64_t* src = yj + ((size_t)(t2_0 + row) * C1 + i1) * D;
    Ext<D> v;
#pragma unroll
    for (int d = 0; d < D; d++) v.c[d] = src[d];
    s[e] = v;
  }
  __syncthreads();
  lds_dif<D>(s, tws, logC1, b);
  uint64_t* oj = a.out + (j * a.N + c * C) * D;
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    int row = e % b, t1 = e / b;
    Ext<D> v = s[row * C1 + bitrev(t1, logC1)];
    uint64_t* o = oj + ((size_t)(t2_0 + row) + (size_t)C2 * t1) * D;
#pragma unroll
    for (int d = 0; d < D; d++) o[d] = v.c[d];
  }
}

// Column-major [M][N] -> row-major [N][M] (host-facing layout of interleaved_encode).
template <int D>
__global__ void k_cols_to_rows(const uint64_t* __restrict__ cw, uint64_t* __restrict__ out, size_t N, size_t M) {
  __shared__ uint64_t tile[32][33 * D];
  size_t r0 = (size_t)blockIdx.x * 32, j0 = (size_t)blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    size_t j = j0 + k, r = r0 + tx;
    if (j < M && r < N)
      for (int d = 0; d < D; d++) tile[k][tx * D + d] = cw[(j * N + r) * D + d];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    size_t r = r0 + k, j = j0 + tx;
    if (j < M && r < N)
      for (int d = 0; d < D; d++) out[(r * M + j) * D + d] = tile[tx][k * D + d];
  }
}

// --------------------------------------------------------------------------
// Host drivers.
// --------------------------------------------------------------------------
static int log2_exact(uint64_t n) {
  int k = 0;
  while ((1ull << k) < n) k++;
  return (1ull << k) == n ? k : -1;
}

static int floor_pow2(int x) {
  int p = 1;
  while (p * 2 <= x) p *= 2;
  return p;
}

template <int D>
static void launch(RsArgs a, uint64_t* scratch, hipStream_t st) {
  TwTable twN = twiddles(log2_exact(a.N));
  const int logC = a.logC;
  const size_t nseq = (size_t)a.M * a.Q;
  if (logC <= MAX_PASS_LOG) {
    int C = 1 << logC;
    int G = std::max(1, floor_pow2(TILE_WORDS / (C * D)));
    const uint64_t* twS = flat_twiddles(logC);
    size_t smem = ((size_t)G * C * D + std::max(1, C / 2)) * 8;
    unsigned blocks = (unsigned)((nseq + G - 1) / G);
    ProfScope prof("ntt_single", st, 8.0 * D * ((double)a.M * a.L + (double)a.M * a.N));
    k_ntt_single<D><<<blocks, NTT_THREADS, smem, st>>>(a, twN, twS, G);
    WHIR_HIP(hipGetLastError());
    return;
  }
  WHIR_REQUIRE(logC <= 2 * MAX_PASS_LOG, "interleaved_encode: coset size above 2^22 is not supported");
  const int logC1 = (logC + 1) / 2, logC2 = logC - logC1;
  const int C1 = 1 << logC1, C2 = 1 << logC2;
  TwTable twC = twiddles(logC);
  // pass 1
  {
    int b = std::min(C1, std::max(1, floor_pow2(TILE_WORDS / (C2 * D))));
    size_t smem = ((size_t)b * C2 * D + C2 / 2) * 8;
    dim3 grid(C1 / b, a.Q, a.M);
    ProfScope prof("ntt_pass1", st, 8.0 * D * ((double)a.M * a.L + (double)a.M * a.N));
    k_ntt_pass1<D><<<grid, NTT_THREADS, smem, st>>>(a, twN, twC, flat_twiddles(logC2), scratch, logC1, b);
    WHIR_HIP(hipGetLastError());
  }
  // pass 2
  {
    int b = std::min(C2, std::max(1, floor_pow2(TILE_WORDS / (C1 * D))));
    size_t smem = ((size_t)b * C1 * D + C1 / 2) * 8;
    dim3 grid(C2 / b, a.Q, a.M);
    ProfScope prof("ntt_pass2", st, 16.0 * D * (double)a.M * a.N);
    k_ntt_pass2<D><<<grid, NTT_THREADS, smem, st>>>(a, flat_twiddles(logC1), scratch, logC1, b);
    WHIR_HIP(hipGetLastError());
  }
}

size_t rs_scratch_words(int D, uint64_t M, uint64_t N, uint64_t C) {
  int logC = log2_exact(C);
  return logC > MAX_PASS_LOG ? (size_t)M * N * D : 0;
}

void rs_encode_cols(int D, RsArgs a, uint64_t* scratch, hipStream_t st) {
  WHIR_REQUIRE(a.M > 0, "rs_encode: no messages");
  WHIR_REQUIRE(log2_exact(a.N) >= 0, "rs_encode: the device path needs a power-of-two codeword length");
  WHIR_REQUIRE(log2_exact(1ull << a.logC) >= 0 && ((a.N >> a.logC) << a.logC) == a.N,
               "rs_encode: coset size must divide the codeword length");
  switch (D) {
    case 1: launch<1>(a, scratch, st); break;
    case 2: launch<2>(a, scratch, st); break;
    case 3: launch<3>(a, scratch, st); break;
    default: throw Error("rs_encode: extension degree must be 1, 2 or 3");
  }
}

void cols_to_rows(int D, const uint64_t* cw, uint64_t* out, size_t N, size_t M, hipStream_t st) {
  dim3 grid((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32));
  switch (D) {
    case 1: k_cols_to_rows<1><<<grid, 256, 0, st>>>(cw, out, N, M); break;
    case 2: k_cols_to_rows<2><<<grid, 256, 0, st>>>(cw, out, N, M); break;
    case 3: k_cols_to_rows<3><<<grid, 256, 0, st>>>(cw, out, N, M); break;
  }
  WHIR_HIP(hipGetLastError());
}

}  // namespace whir
