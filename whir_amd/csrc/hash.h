// Device hash engines, leaf hashing and Merkle tree construction (hash.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace whir {

// Engine numbering of the C ABI; the reference's registry is src/hash/mod.rs:28-36.
enum Engine : int { ENG_COPY = 0, ENG_SHA2 = 1, ENG_KECCAK = 2, ENG_SHA3 = 3, ENG_BLAKE3 = 4 };

bool engine_supports_size(int engine, size_t size);
size_t engine_preferred_batch(int engine);

// count messages of `size` bytes each (contiguous) -> count 32-byte digests.
void hash_many_dev(int engine, size_t size, const uint8_t* in, size_t count, uint8_t* out, hipStream_t st);

// Leaf hashes of a column-major codeword cw[j*N + r] (M columns of degree-D
// elements): leaf r = H(canonical LE bytes of row r).
void hash_rows_cols_dev(int engine, int D, const uint64_t* cw, size_t N, size_t M, uint8_t* leaves,
                        hipStream_t st);

// Leaf hashes and the whole Merkle tree of a column-major codeword into nodes
// (merkle_num_nodes(N) hashes, laid out as merkle_build_dev leaves them).
// The root also goes to root_out (32 bytes, device or mapped host memory) when given.
void commit_cols_dev(int engine, int D, const uint64_t* cw, size_t N, size_t M, uint8_t* nodes, hipStream_t st,
                     uint8_t* root_out = nullptr);
// commit_cols_dev of a codeword whose N rows are all equal (N a power of two), such as
// the codeword of messages of length one: the same nodes, made from row 0 alone (each
// level's node is the hash of the level below paired with itself).
void commit_const_rows_dev(int engine, int D, const uint64_t* cw, size_t N, size_t M, uint8_t* nodes,
                           hipStream_t st, uint8_t* root_out = nullptr);
// The same commit with the root computed on the host (BLAKE3 rows of at most one chunk,
// see const_rows_on_host): root receives it when the call returns; the node array is
// filled by a launch on st.
bool const_rows_on_host(int engine, size_t row_bytes, size_t N);
// commit_cols_dev with the top of the tree (at most 64 nodes wide) hashed on the host,
// for the Keccak engines (host_top_ok): root receives the root when the call returns.
bool host_top_ok(int engine);
void commit_cols_host_top(int engine, int D, const uint64_t* cw, size_t N, size_t M, uint8_t* nodes, hipStream_t st,
                          uint8_t root[32]);
void commit_const_rows_host(int D, const uint64_t* cw, size_t N, size_t M, uint8_t* nodes, hipStream_t st,
                            uint8_t root[32]);

// commit_cols_dev of `count` codewords of the same shape in shared launches: codeword
// i at cw_tab[i], its nodes at nodes_tab[i] (device arrays of pointers), its root to
// roots + 32 i (device memory).  N >= 8, a power of two.  Launches only (no
// allocation, no synchronization), so a stream capture may record it.
void commit_cols_many_dev(int engine, int D, const uint64_t* const* cw_tab, uint8_t* const* nodes_tab, size_t count,
                          size_t N, size_t M, uint8_t* roots, hipStream_t st);

// Leaf hashes of a row-major matrix m[r*M + j].
void hash_rows_dev(int engine, int D, const uint64_t* m, size_t rows, size_t M, uint8_t* leaves,
                   hipStream_t st);

// Merkle nodes over the first num_leaves leaf hashes already stored in
// nodes[0 .. num_leaves); nodes holds (2^(layers+1) - 1) hashes.
size_t merkle_layers(size_t num_leaves);
// out[k] = nodes[idx[k]] (32-byte hashes), idx a device array of k node indices.
void gather_hashes_dev(const uint8_t* nodes, const uint64_t* idx, size_t k, uint8_t* out, hipStream_t st);
size_t merkle_num_nodes(size_t num_leaves);
void merkle_build_dev(int engine, uint8_t* nodes, size_t num_leaves, hipStream_t st, uint8_t* root_out = nullptr);

// Proof of work: the minimal nonce in [first, first + count) with
// u64_le(H(chal || nonce_le || 0^24)[0..8]) <= threshold; false if there is none.
// first and count are multiples of pow_chunk().  One bounded launch.
bool pow_grind_dev(int engine, const uint8_t challenge[32], uint64_t threshold, uint64_t first, uint64_t count,
                   uint64_t* nonce, hipStream_t st);
uint64_t pow_chunk();

}  // namespace whir
