// Device Reed-Solomon encoder interface (implemented in ntt.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

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

}  // namespace whir
