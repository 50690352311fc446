/*
 * CPU oracle for the brute-force L2 kNN path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (the HIP library, the C++
 * Index API, the Python package) links, loads or calls this code. Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark use it,
 * and only as the checker / the reported CPU baseline.
 *
 * Three oracles live here:
 *   1. oracle_search_canonical: a CPU restatement of the reference's
 *      searchL2Distance decomposition (src/internal/search.cu:79-216) in a
 *      fixed, documented FP32 operation order (the "canonical order"):
 *        norm(v)  = fmaf chain  n = fmaf(v[j], v[j], n), j = 0..d-1, n0 = 0
 *        dot(q,x) = fmaf chain  s = fmaf(q[j], x[j], s), j = 0..d-1, s0 = 0
 *        rel      = fmaf(-2, dot, ||x||^2)        (== fl(||x||^2 - 2 dot))
 *        select the k smallest rel, ordered by (rel, index) lexicographically
 *        D        = fl(rel + ||q||^2)              (exact mode)
 *      The reference's own FP32 order is unobservable (closed cuBLAS SGEMM and
 *      compiler-contracted norms), so this order *defines* plain-path parity.
 *   2. oracle_search_fp64: sum_j (q_j - x_j)^2 in double over the float
 *      inputs, k smallest by (distance, index). Judges recall and distance
 *      error for every mode.
 *   3. Encoders: bf16 round-to-nearest-even and OCP FP8 e4m3 (FN) with
 *      saturation, written from the IEEE-754 / OCP OFP8 rules.
 */
#ifndef RPLSH_ORACLE_H
#define RPLSH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Canonical FP32 squared norms: follows src/internal/norm.cu:191-202
 * (computeL2Norm with squared=true) with the canonical fmaf chain. */
void oracle_norms_canonical(const float* x, int64_t n, int d, float* out);

/* Canonical dot product of two length-d vectors (fmaf chain). */
float oracle_dot_canonical(const float* a, const float* b, int d);

/* Canonical FP32 restatement of searchL2Distance (search.cu:79-216):
 * rel = ||x||^2 - 2 q.x (matrix_ops.cu:56-58 with alpha=-2, select.cu:189-224),
 * k smallest by (rel, index), then + ||q||^2 when exact != 0
 * (broadcast.cu:15-38, index_bf.cpp:68-69). Empty index or k > n pads with
 * FLT_MAX / -1 (search.cu:93-100, index.h:42-44). xnorms may be NULL.
 * nthreads <= 0 means all available. */
void oracle_search_canonical(const float* x, int64_t n, int d, const float* xnorms,
                             const float* q, int64_t nq, int k, int exact,
                             int32_t* out_idx, float* out_dist, int nthreads);

/* FP64 brute force: D64 = sum_j (q_j - x_j)^2 accumulated in double,
 * k smallest by (D64, index). */
void oracle_search_fp64(const float* x, int64_t n, int d, const float* q, int64_t nq,
                        int k, int32_t* out_idx, double* out_dist, int nthreads);

/* FP64 distances of given (query, index) pairs: out[i*k+j] for idx[i*k+j]. */
void oracle_pair_dist_fp64(const float* x, int d, const float* q, int64_t nq, int k,
                           const int32_t* idx, double* out);

/* recall@k = mean over queries of |R ∩ G| / k, as Dataset::evaluate
 * (include/dataset.h:135-152) but without mutating the inputs. -1 entries
 * never match. */
double oracle_recall(const int32_t* result, const int32_t* truth, int64_t nq, int k);

/* bf16 round-to-nearest-even of a float (NaN kept quiet). */
uint16_t oracle_bf16_rne(float v);
float oracle_bf16_to_float(uint16_t h);

/* OCP FP8 e4m3 (FN variant: no infinities, max 448, NaN = S.1111.111),
 * round-to-nearest-even with saturation to +-448. */
uint8_t oracle_e4m3_rne_sat(float v);
float oracle_e4m3_to_float(uint8_t b);

/* Random-projection LSH front end (the library's IndexRPLSH; the reference
 * has no LSH step, only the Gaussian initialiser of matrix_ops.cu:71-76).
 * Codes: bit b of table t set iff oracle_dot_canonical(W[t*K+b], x) >=
 * off[t*K+b] (off NULL: 0); codes[t*n + i], margins[i*L*K + t*K + b] =
 * dot - off when not NULL. */
void oracle_lsh_hash(const float* x, int64_t n, int d, const float* w, const float* off, int L,
                     int K, uint32_t* codes, float* margins);

/* BF16 dot product of a and b (length d) in the LSH pre-rank's piece/tree
 * order: P = ceil(d/8) zero-padded pieces of 8, G = min(64, pow2 >= P)
 * partial fmaf chains (chain p: pieces p, p+G, ...), then s[p] += s[p+m] for
 * m = G/2 .. 1; returns s[0]. */
float oracle_bf16_tree_dot(const float* a, const float* b, int d);

/* LSH search: tables = rows stably sorted by code; per query, buckets of
 * (probe p = 0..probes, table t) in that order -- p = 0 the query's code,
 * p >= 1 the code with the bit of the p-th smallest (|dot - off| bits, bit)
 * key flipped -- concatenated and cut after `budget` entries, de-duplicated,
 * then rel = fmaf(-2, dot, ||x||^2) over the distinct rows, k smallest by
 * (rel, index), D = rel + ||q||^2, padded with -1 / FLT_MAX. prerank f > 0
 * (when f*k < budget): first keep the f*k distinct rows with the smallest
 * (rel_b, index), rel_b = fmaf(-2, dot_b, ||x||^2) with dot_b =
 * oracle_bf16_tree_dot(q, x, d). raw_total /
 * distinct_total (may be NULL): entries taken and distinct rows, summed. */
void oracle_lsh_search(const float* x, int64_t n, int d, const float* w, const float* off, int L,
                       int K, int probes, int budget, int prerank, const float* q, int64_t nq, int k,
                       int32_t* out_idx, float* out_dist, int64_t* raw_total,
                       int64_t* distinct_total);

#ifdef __cplusplus
}
#endif
#endif
