// Device sumcheck and covector kernels (algebra.hip).  Vectors are arrays of
// degree-D extension elements (D u64 words each, canonical, lowest coefficient
// first).  Scalars passed by value are host arrays of D words.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace whir {

// compute_sumcheck_polynomial (src/algebra/sumcheck.rs:12-63) over a[0..n), b[0..n);
// result (c0, c2) written to dev_out[0..2D) on the device.
void sumcheck_poly_dev(int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* dev_out, hipStream_t st);

// fold (src/algebra/sumcheck.rs:69-104), in place; returns the new length.
size_t fold_dev(int D, uint64_t* v, size_t n, const uint64_t* w, hipStream_t st);

// fold a and b by w, then compute_sumcheck_polynomial of the folded vectors, in one
// pass (src/algebra/sumcheck.rs:106-111).  Returns the new length.
size_t fold_and_poly_dev(int D, uint64_t* a, uint64_t* b, size_t n, const uint64_t* w, uint64_t* dev_out,
                         hipStream_t st);

// fold_and_poly_dev with the block partials in a caller-provided device area of at
// least fold_poly_parts_words(D, n) words (no allocation; parts may be null when that
// is 0).  The partial count is capped by the kernel's grid (at most 2048 blocks of
// 2 D words), whatever n is.
size_t fold_poly_parts_words(int D, size_t n);
size_t fold_and_poly_ws_dev(int D, uint64_t* a, uint64_t* b, size_t n, const uint64_t* w, uint64_t* dev_out,
                            uint64_t* parts, hipStream_t st);

// The same three with a first operand of Ds words per element (Ds = 1: the
// base-field vector of the initial sumcheck round, read without lifting it).
// Folding writes the folded a to a_out (Ds = D: may be a_in itself).
void sumcheck_poly_mixed_dev(int Ds, int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* dev_out,
                             hipStream_t st);
size_t fold_into_dev(int Ds, int D, const uint64_t* a_in, uint64_t* a_out, size_t n, const uint64_t* w,
                     hipStream_t st);
// fold_into_dev of a and fold_dev of b by the same w in one launch.
size_t fold_both_dev(int Ds, int D, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t n, const uint64_t* w,
                     hipStream_t st);
size_t fold_and_poly_mixed_dev(int Ds, int D, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t n,
                               const uint64_t* w, uint64_t* dev_out, hipStream_t st);

// Two sumcheck rounds per pass over vectors of length L (a power of two >= 4; a of Ds
// words, b of D words).  mode 0: sums of round k and the coefficients of round k+1 in
// r_k (8 values of D words: c0, c2, S0, S1, S2, T0, T1, T2) over (a_in, b); mode 1:
// first folds (a_in, b) of length 4L by r1 then r2 into (a_out, b), then the same 8
// sums over the folded vectors; mode 2: the fold, then (c0, c2) only; mode 3: the fold
// alone.  r1, r2: host arrays of D words (unused in mode 0).
void sumcheck_pair_dev(int mode, int Ds, int D, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t L,
                       const uint64_t* r1, const uint64_t* r2, uint64_t* dev_out, hipStream_t st);

// A covector of length n = 2^nv held as nt product terms instead of n values:
//   y(i) = sum_k hi_k[i >> sh] * lo_k[i & (2^sh - 1)].
// eq terms factor this way (eq(z, i) is the eq of z's first nv - sh coordinates at the
// high bits of i times the eq of the rest at the low bits; the term's scalar goes into
// hi), and so do geometric terms (t x^i = t x^(i_hi 2^sh) * x^(i_lo)).  A sumcheck round
// folds the top variable, a bit of i_hi while hi has bits left, and the same lo factor
// multiplies both halves of every pair, so folding y folds the hi tables alone.
struct CovView {
  const uint64_t* hi = nullptr;  // nt tables of nhi elements (D words each), table k at hi + k nhi D
  const uint64_t* lo = nullptr;  // nt tables of 2^sh elements, table k at lo + k 2^sh D
  uint32_t nt = 0;
  uint32_t sh = 0;
  uint64_t nhi = 0;
};
// Device words covview_build_dev needs for np eq terms and m geometric terms over 2^nv
// elements.
size_t covview_words(int D, size_t np, size_t m, size_t nv);
// The view of covector_dev's sum (no accumulate) for n = 2^nv, split at sh = nv - nv/2:
// np eq terms (points of nv elements, scalars), then m geometric terms; tables in buf.
CovView covview_build_dev(int D, uint64_t* buf, size_t nv, const uint64_t* eq_points, const uint64_t* eq_scalars,
                          size_t np, const uint64_t* scalars, const uint64_t* points, size_t m, hipStream_t st);
// The view of y folded by r1 then r2 (sumcheck_pair_dev's double fold): hi tables of
// nhi / 4 elements written to out_hi (nt nhi / 4 elements).  Needs nhi >= 4.
CovView covview_fold2_dev(int D, const CovView& v, const uint64_t* r1, const uint64_t* r2, uint64_t* out_hi,
                          hipStream_t st);
// y(i) for i < n into out.
void covview_write_dev(int D, const CovView& v, uint64_t* out, size_t n, hipStream_t st);
// sumcheck_pair_dev with b given by a view: mode 0 reads y from v instead of b; modes
// 1-3 take v as the view of the covector ALREADY folded by (r1, r2) (covview_fold2_dev)
// and write its values into b (length L), so b holds the folded covector afterwards.
void sumcheck_pair_view_dev(int mode, int Ds, int D, const uint64_t* a_in, uint64_t* a_out, uint64_t* b, size_t L,
                            const uint64_t* r1, const uint64_t* r2, const CovView& v, uint64_t* dev_out,
                            hipStream_t st);

// eval_eq (src/algebra/multilinear.rs:170-190): acc[x] += s * prod_i eq(z_i, x_i), x_0 = MSB.
void eval_eq_dev(int D, uint64_t* acc, const uint64_t* point, size_t k, const uint64_t* scalar, hipStream_t st);

// geometric_accumulate (src/algebra/mod.rs:154-177): acc[i] += sum_j s_j x_j^i.
void geometric_accumulate_dev(int D, uint64_t* acc, size_t n, const uint64_t* scalars, const uint64_t* points,
                              size_t m, hipStream_t st);

// Both of the above, and any number of each, in one pass:
//   acc[i] = (accumulate ? acc[i] : 0) + sum_p eq_scalars_p eq(eq_points_p, i) + sum_j scalars_j points_j^i
// for i < n (n = 2^nv when np > 0).  Points and scalars are host arrays.
void covector_dev(int D, uint64_t* acc, size_t n, bool accumulate, const uint64_t* eq_points,
                  const uint64_t* eq_scalars, size_t np, size_t nv, const uint64_t* scalars, const uint64_t* points,
                  size_t m, hipStream_t st);
// covector_dev's power terms for a round's STIR constraints: m_e points given as field
// elements (the OOD samples, D words each), then m_b in-domain points x_j = w_N^exps[j]
// (points[j] holds the same values as base-field words), N = 2^logN; scalars (m_e + m_b
// elements, D words each) in that order.  Many terms over a long covector take a kernel
// factored by i = 256 h + l (per-element work: D 64x64 products per in-domain term).
void covector_stir_dev(int D, uint64_t* acc, size_t n, bool accumulate, const uint64_t* scalars,
                       const uint64_t* ext_points, size_t m_e, const uint64_t* exps, const uint64_t* points, size_t m_b,
                       int logN, hipStream_t st);

// mixed_univariate_evaluate at m host points x[0..m) in shared passes; results to
// dev_out[0..m D).
// padded: dev_out has room for m rounded up to a multiple of 4 values (the kernel
// then writes its whole groups there directly).
void univariate_eval_many_dev(int Ds, int D, const uint64_t* coeffs, size_t n, const uint64_t* x, size_t m,
                              uint64_t* dev_out, hipStream_t st, bool padded = false);

// mixed_univariate_evaluate of many base-field vectors, each at its own degree-D
// point, in shared launches: out[e*D .. (e+1)*D) = sum_i jobs[e].coeffs[i] x_e^i.
struct UniJob {
  const uint64_t* coeffs;
  size_t n;
  uint64_t x[3];
};
void univariate_eval_batch_dev(int D, const UniJob* jobs, size_t count, uint64_t* dev_out, hipStream_t st);

// mixed_univariate_evaluate (src/algebra/mod.rs:105-129): sum_i c_i x^i with c in the
// base field (Ds = 1) or in the extension (Ds = D); result to dev_out[0..D).
void univariate_eval_dev(int Ds, int D, const uint64_t* coeffs, size_t n, const uint64_t* x, uint64_t* dev_out,
                         hipStream_t st);

// lift (src/algebra/mod.rs:51-59): dst[i] = embed(src[i]).
void lift_dev(int D, const uint64_t* src, size_t n, uint64_t* dst, hipStream_t st);

// mixed_scalar_mul_add (src/algebra/mod.rs:88-98): acc[i] += w * v[i], v in the base field.
void mixed_scalar_mul_add_dev(int D, uint64_t* acc, const uint64_t* w, const uint64_t* v, size_t n, hipStream_t st);

// scalar_mul_add (src/algebra/mod.rs:76-83): acc[i] += w * v[i], v in the extension.
void scalar_mul_add_dev(int D, uint64_t* acc, const uint64_t* w, const uint64_t* v, size_t n, hipStream_t s);

// dot (src/algebra/mod.rs:36) to dev_out[0..D).
void dot_dev(int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* dev_out, hipStream_t st);

// out[k] = v[idx[k]] rows of a column-major matrix (gather for openings):
// out[k*M + j] = cw[j*N + idx[k]] (degree-D elements).
// Rows and Merkle sibling hashes of an opening in one launch: idx_sib = k row
// indices then ns node indices (device); out_hashes 16-byte aligned.  D > 0 also
// writes out_vals[q] = sum_j wts[j] row_q[j] (D words each; wts: M weights of D words
// in device memory; rows of Ds words embedded in the degree-D extension).
void gather_open_dev(int Ds, const uint64_t* cw, size_t N, size_t M, const uint64_t* idx_sib, size_t k,
                     const uint8_t* nodes, size_t ns, uint64_t* out_rows, uint8_t* out_hashes, hipStream_t st,
                     int D = 0, const uint64_t* wts = nullptr, uint64_t* out_vals = nullptr);

void gather_rows_dev(int D, const uint64_t* cw, size_t N, size_t M, const uint64_t* idx, size_t k, uint64_t* out,
                     hipStream_t st);

}  // namespace whir
