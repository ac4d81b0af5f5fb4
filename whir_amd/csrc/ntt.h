// Device Reed-Solomon encoder interface (implemented in ntt.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <string>

namespace whir {

// One interleaved encode: M messages, message j at msgs + j*msg_stride*D
// (length L elements of degree D), optional masks (mask_len per message),
// codeword length N = Q * 2^logC, written column-major to out[j*N + r].
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
};

// Words of device scratch rs_encode_cols needs for these sizes.
size_t rs_scratch_words(int D, uint64_t M, uint64_t N, uint64_t C);
void rs_encode_cols(int D, RsArgs a, uint64_t* scratch, hipStream_t st);
void cols_to_rows(int D, const uint64_t* cw, uint64_t* out, size_t N, size_t M, hipStream_t st);

// Schedule autotuner of rs_encode_cols (see ntt.hip): tune new shapes on first use,
// pin a schedule for a shape, forget all shapes, and a JSON list of the table.
void ntt_autotune(bool on);
void ntt_set_schedule(int D, int logN, int logC, uint64_t M, int logC1, int v1, int tw1, int v2, int tw2);
void ntt_clear_schedules();
std::string ntt_schedules_json();

}  // namespace whir
