// Device Reed-Solomon encoder interface (implemented in ntt.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "gl.h"

namespace whir {

// One interleaved encode: M messages, message j at msgs + j*msg_stride*D
// (length L elements of degree D), optional masks (mask_len per message),
// codeword length N = Q * 2^logC, written column-major to out[j*N + r].
//
// Several instances in one launch (batched commits): with msg_tab set, message j is
// read from msg_tab[j / mpi] + (j % mpi) * msg_stride * D, and with out_tab set its
// codeword goes to out_tab[j / mpi] + (j % mpi) * N * D (device arrays of one pointer
// per instance, mpi messages each); msgs / out are then unused.
struct RsArgs {
  const uint64_t* msgs;
  size_t msg_stride;
  size_t L;
  const uint64_t* masks;
  size_t mask_len;
  uint64_t* out;
  uint64_t N;
  uint64_t Q;
  int logC;
  uint64_t M;
  const uint64_t* const* msg_tab = nullptr;
  uint64_t* const* out_tab = nullptr;
  uint64_t mpi = 1;
};

// Powers of w_n, the order-n root of unity of the reference's NTT (n = 2^logn), as two
// device tables: w_n^e = lo[e & (2^h - 1)] * hi[e >> h] for e < n.  Built on first use
// per (device, logn), kept for the life of the process; not during a stream capture.
struct TwTable {
  const uint64_t* lo = nullptr;  // w_n^j, j < 2^h
  const uint64_t* hi = nullptr;  // w_n^(j 2^h)
  int h = 0;
};
TwTable twiddles(int logn);
#if defined(__HIP__)
__device__ __forceinline__ uint64_t tw_at(const TwTable& t, uint64_t e) {
  const uint64_t mask = (1ull << t.h) - 1;
  return gl::mul(t.lo[e & mask], t.hi[e >> t.h]);
}
#endif

// Words of device scratch rs_encode_cols needs for these sizes.
size_t rs_scratch_words(int D, uint64_t M, uint64_t N, uint64_t C);
void rs_encode_cols(int D, RsArgs a, uint64_t* scratch, hipStream_t st);
// While set on the calling thread (a stream capture is in progress), rs_encode_cols
// neither tunes nor builds twiddle tables: it uses recorded or built-in schedules and
// requires the tables of its shapes to exist (rs_prepare builds them beforehand).
void rs_set_capturing(bool on);
void rs_prepare(int D, uint64_t N, int logC, uint64_t M, hipStream_t st);
void cols_to_rows(int D, const uint64_t* cw, uint64_t* out, size_t N, size_t M, hipStream_t st);

// Schedule autotuner of rs_encode_cols (see ntt.hip): tune new shapes on first use,
// pin a schedule for a shape, forget all shapes, and a JSON list of the table.
void ntt_autotune(bool on);
void ntt_set_schedule(int D, int logN, int logC, uint64_t M, int logC1, int v1, int tw1, int v2, int tw2);
void ntt_clear_schedules();
std::string ntt_schedules_json();

}  // namespace whir
