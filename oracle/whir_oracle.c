/*
 * whir_oracle.c -- CPU restatement of the WHIR prover hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle that the
 * parity tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use as
 * the checker.  The product (whir_amd/, libwhir_amd.so) never links, loads or
 * calls it.
 *
 * Every routine restates one function of the reference crate (WizardOfMenlo/whir,
 * Rust) and cites it as path:line.  Field arithmetic is deliberately the most
 * obvious form (unsigned __int128 and `%`), so that it shares no code or trick
 * with the device arithmetic it checks.
 *
 * Element layout: an element of the degree-D extension (D = 1, 2, 3) is D
 * consecutive canonical uint64 coefficients, lowest degree first (the order the
 * reference's ArkFieldEncoder writes them, src/protocols/matrix_commit.rs:59-63).
 *
 * Parity status per function: see the header of oracle/oracle.py.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint8_t u8;
typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------- */
/* Goldilocks field and its extensions (src/algebra/fields.rs:85-143).       */
/* ------------------------------------------------------------------------- */

#define GL_P 0xFFFFFFFF00000001ULL

/* Threads of the OpenMP loops below (the large loops split into independent
 * chunks; every partial sum is exact field arithmetic, so the results do not
 * depend on the thread count).  n <= 0 keeps the current setting. */
int oracle_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
}

static inline u64 f_add(u64 a, u64 b) { return (u64)(((u128)a + b) % GL_P); }
static inline u64 f_sub(u64 a, u64 b) { return (u64)(((u128)a + GL_P - b) % GL_P); }
static inline u64 f_mul(u64 a, u64 b) { return (u64)(((u128)a * b) % GL_P); }
static inline u64 f_neg(u64 a) { return a ? GL_P - a : 0; }

static u64 f_pow(u64 a, u64 e) {
    u64 r = 1;
    while (e) {
        if (e & 1) r = f_mul(r, a);
        a = f_mul(a, a);
        e >>= 1;
    }
    return r;
}
static u64 f_inv(u64 a) { return f_pow(a, GL_P - 2); }

/* Non-residue of the extension: Field64_2 = Fp[u]/(u^2 - 7) (fields.rs:93-98),
 * Field64_3 = Fp[u]/(u^3 - 2) (fields.rs:108-114). */
static u64 ext_nonresidue(int D) { return D == 2 ? 7 : (D == 3 ? 2 : 0); }

static void e_add(int D, u64 *r, const u64 *a, const u64 *b) {
    for (int i = 0; i < D; i++) r[i] = f_add(a[i], b[i]);
}
static void e_sub(int D, u64 *r, const u64 *a, const u64 *b) {
    for (int i = 0; i < D; i++) r[i] = f_sub(a[i], b[i]);
}
/* Schoolbook product, then reduce u^D = W. */
static void e_mul(int D, u64 *r, const u64 *a, const u64 *b) {
    u64 t[5] = {0, 0, 0, 0, 0};
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) t[i + j] = f_add(t[i + j], f_mul(a[i], b[j]));
    u64 w = ext_nonresidue(D);
    for (int k = 2 * D - 2; k >= D; k--) t[k - D] = f_add(t[k - D], f_mul(t[k], w));
    for (int i = 0; i < D; i++) r[i] = t[i];
}
static void e_scale(int D, u64 *r, const u64 *a, u64 s) {
    for (int i = 0; i < D; i++) r[i] = f_mul(a[i], s);
}
static void e_set_base(int D, u64 *r, u64 s) {
    r[0] = s;
    for (int i = 1; i < D; i++) r[i] = 0;
}
static void e_pow(int D, u64 *r, const u64 *a, u64 e) {
    u64 acc[3], base[3];
    e_set_base(D, acc, 1);
    memcpy(base, a, sizeof(u64) * D);
    while (e) {
        if (e & 1) e_mul(D, acc, acc, base);
        e_mul(D, base, base, base);
        e >>= 1;
    }
    memcpy(r, acc, sizeof(u64) * D);
}

void oracle_field_op(int D, int op, const u64 *a, const u64 *b, u64 *out, size_t n) {
    /* op: 0 add, 1 sub, 2 mul, 3 pow (b[0] is the exponent for every element), 4 inverse (D=1) */
    for (size_t i = 0; i < n; i++) {
        const u64 *x = a + i * D, *y = b + i * D;
        u64 *z = out + i * D;
        switch (op) {
        case 0: e_add(D, z, x, y); break;
        case 1: e_sub(D, z, x, y); break;
        case 2: e_mul(D, z, x, y); break;
        case 3: e_pow(D, z, x, b[0]); break;
        case 4: z[0] = f_inv(x[0]); break;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* NTT engine (src/algebra/ntt/cooley_tukey.rs).                              */
/* ------------------------------------------------------------------------- */

/* cooley_tukey.rs:53-79 with Field64's FftField constants: the supported
 * subgroup has order 3 * 2^32 and omega_order = 7^((p-1)/order), pinned by the
 * reference test cooley_tukey.rs:582-595. */
#define NTT_ORDER (3ULL << 32)

static u64 ntt_omega_order(void) { return f_pow(7, (GL_P - 1) / NTT_ORDER); }

/* root(order), cooley_tukey.rs:171-183 */
int oracle_ntt_root(u64 order, u64 *out) {
    if (order == 0) { *out = 1; return 0; }
    if (NTT_ORDER % order) return -1;
    *out = f_pow(ntt_omega_order(), NTT_ORDER / order);
    return 0;
}

/* divisors(order, [2, 3]) then binary search, ntt/utils.rs:69-84 and
 * cooley_tukey.rs:364-368.  Divisors of 3*2^32 are 2^a and 3*2^a, a <= 32. */
int oracle_next_order(u64 size, u64 *out) {
    u64 best = 0;
    for (int a = 0; a <= 32; a++) {
        u64 c1 = 1ULL << a, c3 = 3ULL << a;
        if (c1 >= size && (best == 0 || c1 < best)) best = c1;
        if (c3 >= size && (best == 0 || c3 < best)) best = c3;
    }
    if (best == 0) return -1;
    *out = best;
    return 0;
}

/* generator(N) = omega_order^(order/N), cooley_tukey.rs:370-373 */
int oracle_generator(u64 N, u64 *out) {
    if (N == 0 || NTT_ORDER % N) return -1;
    *out = f_pow(ntt_omega_order(), NTT_ORDER / N);
    return 0;
}

/* Coset size selection shared by evaluation_points and interleaved_encode,
 * cooley_tukey.rs:387-390 and 436-439. */
static int coset_size_for(u64 masked_len, u64 N, u64 *C) {
    u64 c;
    if (oracle_next_order(masked_len, &c)) return -1;
    while (N % c) {
        if (oracle_next_order(c + 1, &c)) return -1;
    }
    *C = c;
    return 0;
}

/* evaluation_points with the default (permuted) order:
 * cooley_tukey.rs:375-402, transpose_permute at ntt/transpose.rs:10-15. */
int oracle_evaluation_points(u64 masked_len, u64 N, const u64 *idx, size_t n, u64 *out) {
    if (masked_len > N || NTT_ORDER % N) return -1;
    u64 gen, C;
    oracle_generator(N, &gen);
    if (coset_size_for(masked_len, N, &C)) return -1;
    u64 Q = N / C;
    for (size_t k = 0; k < n; k++) {
        if (idx[k] >= N) return -1;
        u64 row = idx[k] / C, col = idx[k] % C;
        out[k] = f_pow(gen, row + col * Q);
    }
    return 0;
}

/* transpose of a batch of rows x cols matrices (value semantics of
 * ntt/transpose.rs:21-58). */
static void transpose_batch(int D, u64 *v, size_t len, size_t rows, size_t cols) {
    size_t sz = rows * cols;
    if (sz == 0 || len == 0) return;
    size_t nm = len / sz;
    if (nm == 1) { /* one large matrix: split its rows */
        u64 *tmp = (u64 *)malloc(sizeof(u64) * D * sz);
#pragma omp parallel for schedule(static) if (sz > 65536)
        for (size_t i = 0; i < rows; i++)
            for (size_t j = 0; j < cols; j++)
                memcpy(tmp + (j * rows + i) * D, v + (i * cols + j) * D, sizeof(u64) * D);
        memcpy(v, tmp, sizeof(u64) * D * sz);
        free(tmp);
        return;
    }
#pragma omp parallel if (len > 65536)
    {
        u64 *tmp = (u64 *)malloc(sizeof(u64) * D * sz);
#pragma omp for schedule(static)
        for (size_t m = 0; m < nm; m++) {
            u64 *mat = v + m * sz * D;
            for (size_t i = 0; i < rows; i++)
                for (size_t j = 0; j < cols; j++)
                    memcpy(tmp + (j * rows + i) * D, mat + (i * cols + j) * D, sizeof(u64) * D);
            memcpy(mat, tmp, sizeof(u64) * D * sz);
        }
        free(tmp);
    }
}

/* apply_twiddles, cooley_tukey.rs:526-572: values[i][j] *= roots[(i*step*j) % len]
 * where the roots table has `rlen` entries (powers of omega_rlen). */
static void apply_twiddles(int D, u64 *v, size_t len, const u64 *roots, size_t rlen,
                           size_t rows, size_t cols) {
    size_t size = rows * cols;
    size_t step = rlen / size;
#pragma omp parallel for collapse(2) schedule(static) if (len > 65536)
    for (size_t m = 0; m < len / size; m++) {
        for (size_t i = 1; i < rows; i++) {
            u64 *mat = v + m * size * D;
            size_t st = (i * step) % rlen;
            size_t index = st;
            for (size_t j = 1; j < cols; j++) {
                index %= rlen;
                e_scale(D, mat + (i * cols + j) * D, mat + (i * cols + j) * D, roots[index]);
                index += st;
            }
        }
    }
}

/* sqrt_factor, ntt/utils.rs:3-51 */
static size_t sqrt_factor(size_t n) {
    int twos = __builtin_ctzll(n);
    size_t base = n >> twos;
    if (base == 1) return (size_t)1 << (twos / 2);
    if (base == 3) {
        if (twos == 0) return 1;
        return (twos % 2 == 0) ? ((size_t)3 << ((twos - 1) / 2)) : ((size_t)2 << (twos / 2));
    }
    if (base == 9) {
        if (twos == 1) return 3;
        return (twos % 2 == 0) ? ((size_t)3 << (twos / 2)) : ((size_t)4 << (twos / 2));
    }
    return 0;
}

typedef struct {
    u64 h3p, h3m, w4, w8_1, w8_3, w16_1, w16_3, w16_9;
} ntt_consts;

/* cooley_tukey.rs:107-125 */
static void make_consts(ntt_consts *k) {
    u64 w3;
    oracle_ntt_root(3, &w3);
    u64 w3_2 = f_mul(w3, w3);
    u64 half = f_inv(2);
    k->h3m = f_mul(f_sub(w3, w3_2), half);
    k->h3p = f_mul(f_add(w3, w3_2), half);
    oracle_ntt_root(4, &k->w4);
    oracle_ntt_root(8, &k->w8_1);
    k->w8_3 = f_pow(k->w8_1, 3);
    oracle_ntt_root(16, &k->w16_1);
    k->w16_3 = f_pow(k->w16_1, 3);
    k->w16_9 = f_pow(k->w16_1, 9);
}

#define EL(v, i) ((v) + (size_t)(i) * D)
static void bfly(int D, u64 *x, u64 *y) { /* (x, y) = (x + y, x - y) */
    u64 s[3], d[3];
    e_add(D, s, x, y);
    e_sub(D, d, x, y);
    memcpy(x, s, sizeof(u64) * D);
    memcpy(y, d, sizeof(u64) * D);
}
static void swp(int D, u64 *x, u64 *y) {
    u64 t[3];
    memcpy(t, x, sizeof(u64) * D);
    memcpy(x, y, sizeof(u64) * D);
    memcpy(y, t, sizeof(u64) * D);
}
static void scl(int D, u64 *x, u64 s) { e_scale(D, x, x, s); }

static void ntt_dispatch(int D, const ntt_consts *k, u64 *v, size_t len, const u64 *roots,
                         size_t rlen, size_t size);

/* ntt_recurse, cooley_tukey.rs:240-253 */
static void ntt_recurse(int D, const ntt_consts *k, u64 *v, size_t len, const u64 *roots,
                        size_t rlen, size_t size) {
    size_t n1 = sqrt_factor(size);
    size_t n2 = size / n1;
    transpose_batch(D, v, len, n1, n2);
    ntt_dispatch(D, k, v, len, roots, rlen, n1);
    transpose_batch(D, v, len, n2, n1);
    apply_twiddles(D, v, len, roots, rlen, n1, n2);
    ntt_dispatch(D, k, v, len, roots, rlen, n2);
    transpose_batch(D, v, len, n1, n2);
}

/* ntt_dispatch base cases, cooley_tukey.rs:255-360 */
static void ntt_dispatch(int D, const ntt_consts *k, u64 *v0, size_t len, const u64 *roots,
                         size_t rlen, size_t size) {
    switch (size) {
    case 0:
    case 1:
        return;
    case 2:
#pragma omp parallel for schedule(static) if (len > 65536)
        for (size_t o = 0; o < len; o += 2) bfly(D, EL(v0, o), EL(v0, o + 1));
        return;
    case 3:
#pragma omp parallel for schedule(static) if (len > 65536)
        for (size_t o = 0; o < len; o += 3) {
            u64 *v = EL(v0, o);
            u64 a0[3];
            memcpy(a0, EL(v, 0), sizeof(u64) * D);
            bfly(D, EL(v, 1), EL(v, 2));
            e_add(D, EL(v, 0), EL(v, 0), EL(v, 1));
            scl(D, EL(v, 1), k->h3p);
            scl(D, EL(v, 2), k->h3m);
            e_add(D, EL(v, 1), EL(v, 1), a0);
            bfly(D, EL(v, 1), EL(v, 2));
        }
        return;
    case 4:
#pragma omp parallel for schedule(static) if (len > 65536)
        for (size_t o = 0; o < len; o += 4) {
            u64 *v = EL(v0, o);
            bfly(D, EL(v, 0), EL(v, 2));
            bfly(D, EL(v, 1), EL(v, 3));
            scl(D, EL(v, 3), k->w4);
            bfly(D, EL(v, 0), EL(v, 1));
            bfly(D, EL(v, 2), EL(v, 3));
            swp(D, EL(v, 1), EL(v, 2));
        }
        return;
    case 8:
#pragma omp parallel for schedule(static) if (len > 65536)
        for (size_t o = 0; o < len; o += 8) {
            u64 *v = EL(v0, o);
            bfly(D, EL(v, 0), EL(v, 4));
            bfly(D, EL(v, 1), EL(v, 5));
            bfly(D, EL(v, 2), EL(v, 6));
            bfly(D, EL(v, 3), EL(v, 7));
            scl(D, EL(v, 5), k->w8_1);
            scl(D, EL(v, 6), k->w4);
            scl(D, EL(v, 7), k->w8_3);
            bfly(D, EL(v, 0), EL(v, 2));
            bfly(D, EL(v, 1), EL(v, 3));
            scl(D, EL(v, 3), k->w4);
            bfly(D, EL(v, 0), EL(v, 1));
            bfly(D, EL(v, 2), EL(v, 3));
            bfly(D, EL(v, 4), EL(v, 6));
            bfly(D, EL(v, 5), EL(v, 7));
            scl(D, EL(v, 7), k->w4);
            bfly(D, EL(v, 4), EL(v, 5));
            bfly(D, EL(v, 6), EL(v, 7));
            swp(D, EL(v, 1), EL(v, 4));
            swp(D, EL(v, 3), EL(v, 6));
        }
        return;
    case 16:
#pragma omp parallel for schedule(static) if (len > 65536)
        for (size_t o = 0; o < len; o += 16) {
            u64 *v = EL(v0, o);
            for (int i = 0; i < 4; i++) {
                u64 *w = EL(v, i);
                bfly(D, EL(w, 0), EL(w, 8));
                bfly(D, EL(w, 4), EL(w, 12));
                scl(D, EL(w, 12), k->w4);
                bfly(D, EL(w, 0), EL(w, 4));
                bfly(D, EL(w, 8), EL(w, 12));
                swp(D, EL(w, 4), EL(w, 8));
            }
            scl(D, EL(v, 5), k->w16_1);
            scl(D, EL(v, 6), k->w8_1);
            scl(D, EL(v, 7), k->w16_3);
            scl(D, EL(v, 9), k->w8_1);
            scl(D, EL(v, 10), k->w4);
            scl(D, EL(v, 11), k->w8_3);
            scl(D, EL(v, 13), k->w16_3);
            scl(D, EL(v, 14), k->w8_3);
            scl(D, EL(v, 15), k->w16_9);
            for (int i = 0; i < 4; i++) {
                u64 *w = EL(v, i * 4);
                bfly(D, EL(w, 0), EL(w, 2));
                bfly(D, EL(w, 1), EL(w, 3));
                scl(D, EL(w, 3), k->w4);
                bfly(D, EL(w, 0), EL(w, 1));
                bfly(D, EL(w, 2), EL(w, 3));
                swp(D, EL(w, 1), EL(w, 2));
            }
            swp(D, EL(v, 1), EL(v, 4));
            swp(D, EL(v, 2), EL(v, 8));
            swp(D, EL(v, 3), EL(v, 12));
            swp(D, EL(v, 6), EL(v, 9));
            swp(D, EL(v, 7), EL(v, 13));
            swp(D, EL(v, 11), EL(v, 14));
        }
        return;
    default:
        ntt_recurse(D, k, v0, len, roots, rlen, size);
    }
}

/* roots_table(order): powers of root(order), cooley_tukey.rs:186-236 */
static u64 *roots_table(size_t order) {
    u64 w;
    oracle_ntt_root(order, &w);
    u64 *r = (u64 *)malloc(sizeof(u64) * order);
    u64 x = 1;
    for (size_t i = 0; i < order; i++) {
        r[i] = x;
        x = f_mul(x, w);
    }
    return r;
}

/* ntt_batch, cooley_tukey.rs:133-140: `len` elements, NTT of `size` on each
 * contiguous chunk, natural order in and out. */
int oracle_ntt_batch(int D, u64 *values, size_t len, size_t size) {
    if (size == 0 || len % size) return -1;
    if (size <= 1) return 0;
    if (NTT_ORDER % size) return -1;
    ntt_consts k;
    make_consts(&k);
    u64 *roots = roots_table(size);
    ntt_dispatch(D, &k, values, len, roots, size, size);
    free(roots);
    return 0;
}

/* Direct O(n^2) DFT: out[k] = sum_j in[j] * root(n)^(jk).  Used to pin the
 * engine restatement the way cooley_tukey.rs:842-1068 pins the reference. */
int oracle_dft_direct(int D, const u64 *in, u64 *out, size_t n) {
    u64 w;
    if (oracle_ntt_root(n, &w)) return -1;
    for (size_t k2 = 0; k2 < n; k2++) {
        u64 acc[3] = {0, 0, 0};
        u64 wk = f_pow(w, k2), x = 1;
        for (size_t j = 0; j < n; j++) {
            u64 t[3];
            e_scale(D, t, in + j * D, x);
            e_add(D, acc, acc, t);
            x = f_mul(x, wk);
        }
        memcpy(out + k2 * D, acc, sizeof(u64) * D);
    }
    return 0;
}

/* interleaved_encode, cooley_tukey.rs:411-477 (default order, not rs_in_order).
 * messages: num_messages x L elements, contiguous; masks: num_messages x
 * mask_len elements or NULL; out: N x num_messages elements, row-major. */
int oracle_interleaved_encode(int D, const u64 *messages, size_t num_messages, size_t L,
                              const u64 *masks, size_t mask_len, size_t N, u64 *out) {
    if (N == 0 || NTT_ORDER % N) return -1;
    if (num_messages == 0) return 0;
    size_t masked = L + mask_len;
    if (masked > N) return -1;
    u64 C;
    if (coset_size_for(masked, N, &C)) return -1;
    size_t Q = N / C;
    size_t pad = C - masked;
    size_t total = num_messages * N;
    u64 *res = (u64 *)calloc(total * D, sizeof(u64));
    size_t pos = 0;
    for (size_t m = 0; m < num_messages; m++) {
        for (size_t c = 0; c < Q; c++) {
            memcpy(res + pos * D, messages + m * L * D, sizeof(u64) * D * L);
            pos += L;
            if (mask_len) memcpy(res + pos * D, masks + m * mask_len * D, sizeof(u64) * D * mask_len);
            pos += mask_len;
            pos += pad; /* zeros from calloc */
        }
    }
    u64 *roots = roots_table(N);
    apply_twiddles(D, res, total, roots, N, Q, C);
    free(roots);
    oracle_ntt_batch(D, res, total, C);
    transpose_batch(D, res, total, num_messages, N);
    memcpy(out, res, sizeof(u64) * D * total);
    free(res);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Hashes.                                                                    */
/* ------------------------------------------------------------------------- */

/* BLAKE3, restated from the public BLAKE3 specification (the reference calls
 * blake3 1.8.3, src/hash/blake3_engine.rs:88-160).  Full tree mode so that the
 * general hasher used for unsupported sizes (blake3_engine.rs:89-95) is covered. */
static const u32 B3_IV[8] = {0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
                             0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19};
static const int B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum { B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8 };

static inline u32 rotr32(u32 x, int n) { return (x >> n) | (x << (32 - n)); }
static void b3_g(u32 *s, int a, int b, int c, int d, u32 x, u32 y) {
    s[a] = s[a] + s[b] + x;
    s[d] = rotr32(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr32(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + y;
    s[d] = rotr32(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr32(s[b] ^ s[c], 7);
}
static void b3_compress(const u32 cv[8], const u8 block[64], u32 block_len, u64 counter,
                        u32 flags, u32 out[16]) {
    u32 m[16], s[16];
    for (int i = 0; i < 16; i++)
        m[i] = (u32)block[4 * i] | ((u32)block[4 * i + 1] << 8) | ((u32)block[4 * i + 2] << 16) |
               ((u32)block[4 * i + 3] << 24);
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    for (int i = 0; i < 4; i++) s[8 + i] = B3_IV[i];
    s[12] = (u32)counter;
    s[13] = (u32)(counter >> 32);
    s[14] = block_len;
    s[15] = flags;
    for (int r = 0; r < 7; r++) {
        b3_g(s, 0, 4, 8, 12, m[0], m[1]);
        b3_g(s, 1, 5, 9, 13, m[2], m[3]);
        b3_g(s, 2, 6, 10, 14, m[4], m[5]);
        b3_g(s, 3, 7, 11, 15, m[6], m[7]);
        b3_g(s, 0, 5, 10, 15, m[8], m[9]);
        b3_g(s, 1, 6, 11, 12, m[10], m[11]);
        b3_g(s, 2, 7, 8, 13, m[12], m[13]);
        b3_g(s, 3, 4, 9, 14, m[14], m[15]);
        u32 t[16];
        for (int i = 0; i < 16; i++) t[i] = m[B3_PERM[i]];
        memcpy(m, t, sizeof m);
    }
    for (int i = 0; i < 8; i++) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}
/* Chaining value of one chunk (<= 1024 bytes) with chunk counter `ctr`;
 * `root` marks the final block as ROOT (only for single-chunk inputs). */
static void b3_chunk(const u8 *in, size_t len, u64 ctr, int root, u32 cv_out[8]) {
    u32 cv[8], out[16];
    memcpy(cv, B3_IV, sizeof cv);
    size_t nblocks = len == 0 ? 1 : (len + 63) / 64;
    for (size_t b = 0; b < nblocks; b++) {
        u8 block[64];
        memset(block, 0, 64);
        size_t off = b * 64, bl = len - off < 64 ? len - off : 64;
        if (len == 0) bl = 0;
        memcpy(block, in + off, bl);
        u32 flags = 0;
        if (b == 0) flags |= B3_CHUNK_START;
        if (b == nblocks - 1) flags |= B3_CHUNK_END | (root ? B3_ROOT : 0);
        b3_compress(cv, block, (u32)bl, ctr, flags, out);
        memcpy(cv, out, 32);
    }
    memcpy(cv_out, cv, 32);
}
static void b3_parent(const u32 l[8], const u32 r[8], int root, u32 out_cv[8]) {
    u8 block[64];
    for (int i = 0; i < 8; i++) {
        for (int b = 0; b < 4; b++) {
            block[4 * i + b] = (u8)(l[i] >> (8 * b));
            block[32 + 4 * i + b] = (u8)(r[i] >> (8 * b));
        }
    }
    u32 out[16];
    b3_compress(B3_IV, block, 64, 0, B3_PARENT | (root ? B3_ROOT : 0), out);
    memcpy(out_cv, out, 32);
}
/* Subtree over chunks [c0, c0 + n), n a power of two handled by the spec's
 * left-largest-power-of-two split. */
static void b3_subtree(const u8 *in, size_t len, u64 c0, int root, u32 cv[8]) {
    size_t nchunks = len <= 1024 ? 1 : (len + 1023) / 1024;
    if (nchunks == 1) {
        b3_chunk(in, len, c0, root, cv);
        return;
    }
    size_t left_chunks = 1;
    while (left_chunks * 2 < nchunks) left_chunks *= 2;
    size_t left_len = left_chunks * 1024;
    u32 l[8], r[8];
    b3_subtree(in, left_len, c0, 0, l);
    b3_subtree(in + left_len, len - left_len, c0 + left_chunks, 0, r);
    b3_parent(l, r, root, cv);
}
void oracle_blake3(const u8 *in, size_t len, u8 out[32]) {
    u32 cv[8];
    b3_subtree(in, len, 0, 1, cv);
    for (int i = 0; i < 8; i++)
        for (int b = 0; b < 4; b++) out[4 * i + b] = (u8)(cv[i] >> (8 * b));
}

/* Keccak-f[1600] (FIPS 202), with SHA3-256 and Keccak-256 padding.  The
 * reference uses sha3 0.10.8 through DigestEngine (src/hash/digest_engine.rs:22-23). */
static const u64 KRC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL, 0x8000000080008000ULL,
    0x000000000000808bULL, 0x0000000080000001ULL, 0x8000000080008081ULL, 0x8000000000008009ULL,
    0x000000000000008aULL, 0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL, 0x8000000000008003ULL,
    0x8000000000008002ULL, 0x8000000000000080ULL, 0x000000000000800aULL, 0x800000008000000aULL,
    0x8000000080008081ULL, 0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL};
static const int KROT[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43,
                             25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};
static inline u64 rotl64(u64 x, int n) { return n ? (x << n) | (x >> (64 - n)) : x; }
void oracle_keccak_f(u64 A[25]) {
    for (int r = 0; r < 24; r++) {
        u64 C[5], Dd[5], B[25];
        for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        for (int x = 0; x < 5; x++) Dd[x] = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) A[i] ^= Dd[i % 5];
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++) B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(A[x + 5 * y], KROT[x + 5 * y]);
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++)
                A[x + 5 * y] = B[x + 5 * y] ^ ((~B[(x + 1) % 5 + 5 * y]) & B[(x + 2) % 5 + 5 * y]);
        A[0] ^= KRC[r];
    }
}
static void keccak256_pad(const u8 *in, size_t len, u8 pad, u8 out[32]) {
    u64 A[25];
    memset(A, 0, sizeof A);
    const size_t rate = 136;
    u8 block[136];
    size_t off = 0;
    for (;;) {
        size_t take = len - off < rate ? len - off : rate;
        memset(block, 0, rate);
        memcpy(block, in + off, take);
        int last = take < rate;
        if (last) {
            block[take] ^= pad;
            block[rate - 1] ^= 0x80;
        }
        for (size_t i = 0; i < rate / 8; i++) {
            u64 w = 0;
            for (int b = 0; b < 8; b++) w |= (u64)block[8 * i + b] << (8 * b);
            A[i] ^= w;
        }
        oracle_keccak_f(A);
        off += take;
        if (last) break;
    }
    for (int i = 0; i < 4; i++)
        for (int b = 0; b < 8; b++) out[8 * i + b] = (u8)(A[i] >> (8 * b));
}
void oracle_sha3_256(const u8 *in, size_t len, u8 out[32]) { keccak256_pad(in, len, 0x06, out); }
void oracle_keccak256(const u8 *in, size_t len, u8 out[32]) { keccak256_pad(in, len, 0x01, out); }

/* SHA-256 (FIPS 180-4), the reference's `Sha2` engine (digest_engine.rs:21). */
static const u32 SHA_K[64] = {
    0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
    0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
    0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
    0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
    0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
    0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
    0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
    0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
static void sha256_block(u32 H[8], const u8 *blk) {
    u32 W[64];
    for (int i = 0; i < 16; i++)
        W[i] = ((u32)blk[4 * i] << 24) | ((u32)blk[4 * i + 1] << 16) | ((u32)blk[4 * i + 2] << 8) | blk[4 * i + 3];
    for (int i = 16; i < 64; i++) {
        u32 s0 = rotr32(W[i - 15], 7) ^ rotr32(W[i - 15], 18) ^ (W[i - 15] >> 3);
        u32 s1 = rotr32(W[i - 2], 17) ^ rotr32(W[i - 2], 19) ^ (W[i - 2] >> 10);
        W[i] = W[i - 16] + s0 + W[i - 7] + s1;
    }
    u32 a = H[0], b = H[1], c = H[2], d = H[3], e = H[4], f = H[5], g = H[6], h = H[7];
    for (int i = 0; i < 64; i++) {
        u32 S1 = rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25);
        u32 ch = (e & f) ^ (~e & g);
        u32 t1 = h + S1 + ch + SHA_K[i] + W[i];
        u32 S0 = rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22);
        u32 mj = (a & b) ^ (a & c) ^ (b & c);
        u32 t2 = S0 + mj;
        h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    H[0] += a; H[1] += b; H[2] += c; H[3] += d; H[4] += e; H[5] += f; H[6] += g; H[7] += h;
}
void oracle_sha256(const u8 *in, size_t len, u8 out[32]) {
    u32 H[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a, 0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    size_t full = len / 64;
    for (size_t i = 0; i < full; i++) sha256_block(H, in + 64 * i);
    u8 tail[128];
    memset(tail, 0, sizeof tail);
    size_t rem = len - full * 64;
    memcpy(tail, in + full * 64, rem);
    tail[rem] = 0x80;
    size_t tl = rem + 9 <= 64 ? 64 : 128;
    u64 bits = (u64)len * 8;
    for (int i = 0; i < 8; i++) tail[tl - 1 - i] = (u8)(bits >> (8 * i));
    sha256_block(H, tail);
    if (tl == 128) sha256_block(H, tail + 64);
    for (int i = 0; i < 8; i++)
        for (int b = 0; b < 4; b++) out[4 * i + b] = (u8)(H[i] >> (24 - 8 * b));
}

/* Engine numbering shared with the product's C ABI: 0 Copy, 1 Sha2, 2 Keccak,
 * 3 Sha3, 4 Blake3 (the registry of src/hash/mod.rs:28-36). */
static int hash_one(int engine, const u8 *in, size_t len, u8 out[32]) {
    switch (engine) {
    case 0: /* Copy engine, src/hash/copy_engine.rs:25-60: zero-padded copy of <= 32 bytes */
        if (len > 32) return -1;
        memset(out, 0, 32);
        memcpy(out, in, len);
        return 0;
    case 1: oracle_sha256(in, len, out); return 0;
    case 2: oracle_keccak256(in, len, out); return 0;
    case 3: oracle_sha3_256(in, len, out); return 0;
    case 4: oracle_blake3(in, len, out); return 0;
    }
    return -1;
}

/* HashEngine::hash_many, src/hash/mod.rs:79-85: `count` messages of `size`
 * bytes each, concatenated. */
int oracle_hash_many(int engine, size_t size, const u8 *in, size_t count, u8 *out) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad) if (count > 256)
    for (size_t i = 0; i < count; i++)
        bad |= hash_one(engine, in + i * size, size, out + 32 * i) != 0;
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* Merkle tree and matrix commitment.                                         */
/* ------------------------------------------------------------------------- */

/* layers_for_size, src/protocols/merkle_tree.rs:281-283 */
size_t oracle_layers_for_size(size_t n) {
    size_t p = 1, l = 0;
    while (p < n) { p <<= 1; l++; }
    return l;
}

/* merkle_tree::Config::commit, merkle_tree.rs:82-127.  nodes_out holds
 * (2^(layers+1) - 1) * 32 bytes: [padded leaves][next layer]...[root]. */
int oracle_merkle_commit(int engine, const u8 *leaves, size_t num_leaves, u8 *nodes_out) {
    size_t layers = oracle_layers_for_size(num_leaves);
    size_t width = (size_t)1 << layers;
    size_t num_nodes = ((size_t)1 << (layers + 1)) - 1;
    memset(nodes_out, 0, 32 * num_nodes);
    memcpy(nodes_out, leaves, 32 * num_leaves);
    size_t prev = 0;
    size_t cur = width;
    for (size_t w = width; w > 1; w >>= 1) {
        if (oracle_hash_many(engine, 64, nodes_out + 32 * prev, w / 2, nodes_out + 32 * cur)) return -1;
        prev = cur;
        cur += w / 2;
    }
    return 0;
}

static int cmp_size(const void *a, const void *b) {
    size_t x = *(const size_t *)a, y = *(const size_t *)b;
    return x < y ? -1 : (x > y);
}

/* merkle_tree::Config::open, merkle_tree.rs:145-186: sibling hashes in the
 * order they are written as hints.  Returns the number of hints written. */
long oracle_merkle_open(const u8 *nodes, size_t num_leaves, const size_t *indices, size_t n,
                        u8 *hints_out) {
    size_t layers = oracle_layers_for_size(num_leaves);
    size_t *idx = (size_t *)malloc(sizeof(size_t) * (n ? n : 1));
    memcpy(idx, indices, sizeof(size_t) * n);
    qsort(idx, n, sizeof(size_t), cmp_size);
    size_t m = 0;
    for (size_t i = 0; i < n; i++)
        if (m == 0 || idx[m - 1] != idx[i]) idx[m++] = idx[i];
    size_t layer_off = 0, width = (size_t)1 << layers;
    long nh = 0;
    while (width > 1) {
        size_t k = 0, nm = 0;
        while (k < m) {
            size_t a = idx[k];
            if (k + 1 < m && idx[k + 1] == (a ^ 1)) {
                idx[nm++] = a >> 1;
                k += 2;
            } else {
                memcpy(hints_out + 32 * nh, nodes + 32 * (layer_off + (a ^ 1)), 32);
                nh++;
                idx[nm++] = a >> 1;
                k += 1;
            }
        }
        m = nm;
        layer_off += width;
        width >>= 1;
    }
    free(idx);
    return nh;
}

/* ArkFieldEncoder + hash_rows, src/protocols/matrix_commit.rs:130-156, 334-368:
 * leaf r = H(canonical LE bytes of row r). */
int oracle_hash_rows(int engine, int D, const u64 *matrix, size_t rows, size_t cols, u8 *out) {
    size_t msg = 8 * (size_t)D * cols;
    int bad = 0;
#pragma omp parallel reduction(| : bad) if (rows > 256)
    {
        u8 *buf = (u8 *)malloc(msg ? msg : 1);
#pragma omp for schedule(static)
        for (size_t r = 0; r < rows; r++) {
            const u64 *row = matrix + r * cols * D;
            for (size_t e = 0; e < cols * (size_t)D; e++)
                for (int b = 0; b < 8; b++) buf[8 * e + b] = (u8)(row[e] >> (8 * b));
            bad |= hash_one(engine, buf, msg, out + 32 * r) != 0;
        }
        free(buf);
    }
    return bad ? -1 : 0;
}

/* ------------------------------------------------------------------------- */
/* Sumcheck and covector kernels.                                             */
/* ------------------------------------------------------------------------- */

static size_t next_pow2(size_t n) {
    size_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

/* compute_sumcheck_polynomial, src/algebra/sumcheck.rs:12-63 */
void oracle_sumcheck_poly(int D, const u64 *a, size_t na, const u64 *b, size_t nb, u64 *c0, u64 *c2) {
    size_t n = na < nb ? na : nb;
    e_set_base(D, c0, 0);
    e_set_base(D, c2, 0);
    if (n == 0) return;
    if (n == 1) {
        e_mul(D, c0, a, b);
        return;
    }
    size_t half = next_pow2(n) >> 1;
    size_t nhi = n - half;
    /* partial sums per thread, added in any order (exact arithmetic) */
#pragma omp parallel if (half > 4096)
    {
        u64 p0[3] = {0, 0, 0}, p2[3] = {0, 0, 0}, acc[3] = {0, 0, 0}, t[3], u[3], v[3];
#pragma omp for schedule(static) nowait
        for (size_t i = 0; i < nhi; i++) {
            e_mul(D, t, a + i * D, b + i * D);
            e_add(D, p0, p0, t);
            e_sub(D, u, a + (half + i) * D, a + i * D);
            e_sub(D, v, b + (half + i) * D, b + i * D);
            e_mul(D, t, u, v);
            e_add(D, p2, p2, t);
        }
#pragma omp for schedule(static) nowait
        for (size_t i = nhi; i < half; i++) {
            e_mul(D, t, a + i * D, b + i * D);
            e_add(D, acc, acc, t);
        }
        e_add(D, p0, p0, acc);
        e_add(D, p2, p2, acc);
#pragma omp critical
        {
            e_add(D, c0, c0, p0);
            e_add(D, c2, c2, p2);
        }
    }
}

/* fold, src/algebra/sumcheck.rs:69-104.  Returns the new length. */
size_t oracle_fold(int D, u64 *v, size_t n, const u64 *w) {
    if (n <= 1) return n;
    size_t half = next_pow2(n) >> 1;
    size_t nhi = n - half;
#pragma omp parallel for schedule(static) if (nhi > 4096)
    for (size_t i = 0; i < nhi; i++) {
        u64 t[3], u[3];
        e_sub(D, u, v + (half + i) * D, v + i * D);
        e_mul(D, t, u, w);
        e_add(D, v + i * D, v + i * D, t);
    }
    u64 one_minus[3], one[3];
    e_set_base(D, one, 1);
    e_sub(D, one_minus, one, w);
    for (size_t i = nhi; i < half; i++) e_mul(D, v + i * D, v + i * D, one_minus);
    return half;
}

/* eval_eq, src/algebra/multilinear.rs:170-190 (point[0] selects the high half). */
static void eval_eq_rec(int D, u64 *acc, const u64 *point, size_t k, const u64 *s) {
    if (k == 0) {
        e_add(D, acc, acc, s);
        return;
    }
    u64 s1[3], s0[3];
    e_mul(D, s1, s, point);
    e_sub(D, s0, s, s1);
    size_t half = (size_t)1 << (k - 1);
    eval_eq_rec(D, acc, point + D, k - 1, s0);
    eval_eq_rec(D, acc + half * D, point + D, k - 1, s1);
}
void oracle_eval_eq(int D, u64 *acc, const u64 *point, size_t k, const u64 *scalar) {
    /* the 2^t subtrees below the first t variables are independent */
    size_t t = k < 6 ? 0 : 6;
    size_t sub = (size_t)1 << (k - t);
#pragma omp parallel for schedule(dynamic) if (t > 0)
    for (size_t h = 0; h < ((size_t)1 << t); h++) {
        u64 s[3];
        memcpy(s, scalar, sizeof(u64) * D);
        for (size_t q = 0; q < t; q++) {
            u64 s1[3], s0[3];
            e_mul(D, s1, s, point + q * D);
            e_sub(D, s0, s, s1);
            memcpy(s, ((h >> (t - 1 - q)) & 1) ? s1 : s0, sizeof(u64) * D);
        }
        eval_eq_rec(D, acc + h * sub * D, point + t * D, k - t, s);
    }
}

/* multilinear_extend (exact size), src/algebra/multilinear.rs:6-151 */
void oracle_multilinear_extend(int D, const u64 *evals, size_t k, const u64 *point, u64 *out) {
    size_t n = (size_t)1 << k;
    u64 *t = (u64 *)malloc(sizeof(u64) * D * n);
    memcpy(t, evals, sizeof(u64) * D * n);
    for (size_t j = 0; j < k; j++) {
        /* fold on the first (most significant) variable */
        oracle_fold(D, t, n >> j, point + j * D);
    }
    memcpy(out, t, sizeof(u64) * D);
    free(t);
}

/* geometric_accumulate, src/algebra/mod.rs:154-177:
 * acc[i] += sum_j scalars[j] * points[j]^i. */
void oracle_geometric_accumulate(int D, u64 *acc, size_t n, const u64 *scalars, const u64 *points,
                                 size_t m) {
    const size_t CH = 4096;
#pragma omp parallel for schedule(dynamic) if (n > CH)
    for (size_t i0 = 0; i0 < n; i0 += CH) {
        size_t i1 = i0 + CH < n ? i0 + CH : n;
        u64 *s = (u64 *)malloc(sizeof(u64) * D * (m ? m : 1));
        for (size_t j = 0; j < m; j++) { /* s_j = scalar_j * x_j^i0 */
            u64 xp[3];
            e_pow(D, xp, points + j * D, i0);
            e_mul(D, s + j * D, scalars + j * D, xp);
        }
        for (size_t i = i0; i < i1; i++) {
            for (size_t j = 0; j < m; j++) {
                e_add(D, acc + i * D, acc + i * D, s + j * D);
                e_mul(D, s + j * D, s + j * D, points + j * D);
            }
        }
        free(s);
    }
}

/* mixed_univariate_evaluate (Basefield embedding), src/algebra/mod.rs:105-129:
 * sum_i coeffs[i] * x^i with coeffs in the base field (Ds = 1) or the same
 * extension (Ds = D), x in the degree-D extension. */
void oracle_univariate_evaluate(int Ds, int D, const u64 *coeffs, size_t n, const u64 *x, u64 *out) {
    /* Horner per chunk of 2^16 coefficients, chunk b scaled by x^(b 2^16) */
    const size_t CH = (size_t)1 << 16;
    u64 total[3];
    e_set_base(D, total, 0);
#pragma omp parallel for schedule(dynamic) if (n > CH)
    for (size_t i0 = 0; i0 < n; i0 += CH) {
        size_t i1 = i0 + CH < n ? i0 + CH : n;
        u64 acc[3];
        e_set_base(D, acc, 0);
        for (size_t i = i1; i-- > i0;) {
            e_mul(D, acc, acc, x);
            u64 c[3];
            if (Ds == 1) e_set_base(D, c, coeffs[i]);
            else memcpy(c, coeffs + i * D, sizeof(u64) * D);
            e_add(D, acc, acc, c);
        }
        u64 xp[3];
        e_pow(D, xp, x, i0);
        e_mul(D, acc, acc, xp);
#pragma omp critical
        e_add(D, total, total, acc);
    }
    memcpy(out, total, sizeof(u64) * D);
}

/* lift (Basefield embedding map), src/algebra/mod.rs:51-59 */
void oracle_lift(int D, const u64 *src, size_t n, u64 *dst) {
    for (size_t i = 0; i < n; i++) e_set_base(D, dst + i * D, src[i]);
}

/* mixed_scalar_mul_add, src/algebra/mod.rs:88-98: acc[i] += w * v[i] (v base) */
void oracle_mixed_scalar_mul_add(int D, u64 *acc, const u64 *w, const u64 *v, size_t n) {
    for (size_t i = 0; i < n; i++) {
        u64 t[3];
        e_scale(D, t, w, v[i]);
        e_add(D, acc + i * D, acc + i * D, t);
    }
}

/* dot, src/algebra/mod.rs:36 */
void oracle_dot(int D, const u64 *a, const u64 *b, size_t n, u64 *out) {
    u64 total[3];
    e_set_base(D, total, 0);
#pragma omp parallel if (n > 4096)
    {
        u64 acc[3], t[3];
        e_set_base(D, acc, 0);
#pragma omp for schedule(static) nowait
        for (size_t i = 0; i < n; i++) {
            e_mul(D, t, a + i * D, b + i * D);
            e_add(D, acc, acc, t);
        }
#pragma omp critical
        e_add(D, total, total, acc);
    }
    memcpy(out, total, sizeof(u64) * D);
}

/* ------------------------------------------------------------------------- */
/* Proof of work and challenge indices.                                      */
/* ------------------------------------------------------------------------- */

/* The minimal nonce of proof_of_work::Config::prove, src/protocols/proof_of_work.rs:68-166:
 * u64_le(H(challenge || nonce_le || 0^24)[0..8]) <= threshold.  Scans nonces
 * [start, start + limit); returns 1 and the nonce if found. */
int oracle_pow_search(int engine, const u8 challenge[32], u64 threshold, u64 start, u64 limit,
                      u64 *nonce_out) {
    u8 in[64], h[32];
    memset(in, 0, 64);
    memcpy(in, challenge, 32);
    for (u64 n = start; n - start < limit; n++) {
        for (int b = 0; b < 8; b++) in[32 + b] = (u8)(n >> (8 * b));
        hash_one(engine, in, 64, h);
        u64 v = 0;
        for (int b = 0; b < 8; b++) v |= (u64)h[b] << (8 * b);
        if (v <= threshold) {
            *nonce_out = n;
            return 1;
        }
    }
    return 0;
}

/* challenge_indices byte-to-index map, src/protocols/challenge_indices.rs:29-42 */
void oracle_bytes_to_indices(const u8 *entropy, size_t count, size_t num_leaves, size_t *out) {
    size_t sb = 0;
    size_t lg = 0;
    while (((size_t)1 << lg) < num_leaves) lg++;
    sb = (lg + 7) / 8;
    for (size_t i = 0; i < count; i++) {
        size_t acc = 0;
        for (size_t b = 0; b < sb; b++) acc = (acc << 8) | entropy[i * sb + b];
        out[i] = acc % num_leaves;
    }
}

/* proof_of_work::threshold, src/protocols/proof_of_work.rs:29-39 (f64 exp2, ceil). */
#include <math.h>
u64 oracle_pow_threshold(double bits) {
    double t = ceil(exp2(64.0 - bits));
    if (t >= 18446744073709551615.0) return ~0ULL;
    return (u64)t;
}
