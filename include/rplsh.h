/*
 * rplsh.h -- C ABI of the MI355X brute-force L2 kNN library (libcurplsh_amd.so).
 *
 * This is the boundary the C++ Index/IndexBF host classes (include/curplsh/)
 * call, and what any other binding (ctypes, cgo, JNI) binds. Plain pointers,
 * sizes and status codes only; no HIP or C++ types cross it. Streams are
 * passed as opaque `void*` (a hipStream_t; NULL = the null stream).
 *
 * Conventions
 *  - Every entry point returns rplsh_status; RPLSH_OK == 0. On failure a
 *    thread-local message is available from rplsh_last_error().
 *  - Pointers are caller-owned unless a function says it copies.
 *  - Functions named *_dev take device pointers only. rplsh_bf_add and
 *    rplsh_bf_search accept host or device pointers, like the reference
 *    IndexBF (src/index_bf.cpp:26-74).
 *  - Indices are int32, distances float32, row-major [rows][cols].
 *  - Not thread-safe per handle: one rplsh_resources / rplsh_bf per thread, or
 *    external locking (device_resources.h:113-165 is unlocked in the reference
 *    too).
 */
#ifndef RPLSH_H
#define RPLSH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rplsh_status {
  RPLSH_OK = 0,
  RPLSH_ERR_INVALID_ARG = 1,  /* precondition violated (reference: host_assert -> abort) */
  RPLSH_ERR_HIP = 2,          /* HIP runtime error (reference: checkCudaErrors -> exit) */
  RPLSH_ERR_OUT_OF_MEMORY = 3,
  RPLSH_ERR_STATE = 4,        /* call not valid in the object's current state */
  RPLSH_ERR_UNSUPPORTED = 5,
  RPLSH_ERR_BUDGET = 6,       /* HBM budget of a serving loop exceeded */
  RPLSH_ERR_NO_DEVICE = 7
} rplsh_status;

/* Message of the last failed call on this thread ("" if none). */
const char* rplsh_last_error(void);
const char* rplsh_version(void);

/* ------------------------------------------------------------ devices -- */
/* getNumDevices / getCurrentDevice / setCurrentDevice
 * (include/internal/cuda_utils.h:1230-1300, used by index.cpp:14). */
rplsh_status rplsh_device_count(int* count);
rplsh_status rplsh_get_device(int* device);
rplsh_status rplsh_set_device(int device);
/* getDeviceForAddress (cuda_utils.h:1369-1388): device of a pointer, -1 for
 * host memory. */
rplsh_status rplsh_pointer_device(const void* ptr, int* device);

/* ---------------------------------------------------------- resources -- */
/* DeviceResources (include/device_resources.h:60-263): per-device default
 * stream, async-copy stream and two alternate streams, created lazily. There
 * is no BLAS or RNG handle: the contraction is this library's own kernel. */
typedef struct rplsh_resources rplsh_resources;

rplsh_status rplsh_resources_create(rplsh_resources** out);
rplsh_status rplsh_resources_destroy(rplsh_resources* res);
/* initializeForDevice (device_resources.h:113-165). */
rplsh_status rplsh_resources_init_device(rplsh_resources* res, int device);
/* getDefaultStream (device_resources.h:181-184). */
rplsh_status rplsh_resources_default_stream(rplsh_resources* res, int device, void** stream);
/* setDefaultStream: a user stream replaces the created one
 * (userDefaultStreams_, device_resources.h:248). NULL restores it. */
rplsh_status rplsh_resources_set_default_stream(rplsh_resources* res, int device, void* stream);
/* getAlternateStreams (device_resources.h:187-190); writes up to *count. */
rplsh_status rplsh_resources_alternate_streams(rplsh_resources* res, int device,
                                               void** streams, int* count);
/* getAsyncCopyStream (device_resources.h:199-202). */
rplsh_status rplsh_resources_async_copy_stream(rplsh_resources* res, int device, void** stream);
/* syncDefaultStream (device_resources.h:220-223). */
rplsh_status rplsh_resources_sync_default_stream(rplsh_resources* res, int device);

/* ------------------------------------------------------------- memory -- */
/* allocMemory / copyMemory (include/memory_space.h:26-41). memory_space:
 * 0 = Device (hipMalloc), 1 = Unified (hipMallocManaged). */
rplsh_status rplsh_malloc(int device, size_t bytes, int memory_space, void** out);
rplsh_status rplsh_free(void* ptr);
/* Stream-ordered copy in any direction (hipMemcpyDefault semantics). */
rplsh_status rplsh_memcpy_async(void* dst, const void* src, size_t bytes, void* stream);
rplsh_status rplsh_memset_async(void* dst, int value, size_t bytes, void* stream);
rplsh_status rplsh_stream_sync(void* stream);
rplsh_status rplsh_device_sync(void);
/* Pinned host memory (hipHostMalloc) for fast staging. */
rplsh_status rplsh_host_alloc(size_t bytes, void** out);
rplsh_status rplsh_host_free(void* ptr);
rplsh_status rplsh_mem_info(int device, size_t* free_bytes, size_t* total_bytes);

/* Events for timing on a given stream. */
typedef struct rplsh_event rplsh_event;
rplsh_status rplsh_event_create(rplsh_event** out);
rplsh_status rplsh_event_destroy(rplsh_event* ev);
rplsh_status rplsh_event_record(rplsh_event* ev, void* stream);
rplsh_status rplsh_event_elapsed_ms(rplsh_event* start, rplsh_event* stop, float* ms);
rplsh_status rplsh_event_sync(rplsh_event* ev);

/* ------------------------------------------- internal kernels (device) -- */
/* computeL2Norm (include/internal/norm.h:7-8, src/internal/norm.cu:191-202):
 * out[i] = sum_j x[i][j]^2 (squared != 0) or its sqrt, in the canonical
 * order n = fmaf(x[j], x[j], n), j = 0..d-1. Device pointers. */
rplsh_status rplsh_l2norm_dev(const float* x, int64_t n, int d, float* out, int squared,
                              void* stream);

/* Encoders used by the precision forms. Device pointers.
 * bf16: round-to-nearest-even, row-major [n][ld] with zero padding d..ld-1.
 * e4m3: per-row FP32 scale s = the smallest power of two >= max_j|x_j| / 448
 * (1 for an all-zero row; a power of two so the MFMA applies it exactly as an
 * E8M0 exponent), code = e4m3_rne_sat(x_j / s); scales[i] = s. */
rplsh_status rplsh_encode_bf16_dev(const float* x, int64_t n, int d, int ld, uint16_t* out,
                                   void* stream);
rplsh_status rplsh_encode_e4m3_dev(const float* x, int64_t n, int d, int ld, uint8_t* out,
                                   float* scales, void* stream);

/* Precision of the candidate contraction. */
typedef enum rplsh_precision {
  RPLSH_PREC_FP32 = 0, /* plain path: FP32 MFMA candidates + canonical FP32 re-rank */
  RPLSH_PREC_BF16 = 1, /* BF16 MFMA candidates, over-fetch, FP32 + FP64-boundary re-rank */
  RPLSH_PREC_FP8 = 2   /* FP8 e4m3 (per-row scale) candidates, same re-rank */
} rplsh_precision;

/* searchL2Distance (include/internal/search.h:8-11, search.cu:79-216) on
 * device pointers, plain FP32 path: indices/distances [nq][k] ascending by
 * (distance, index). xnorms may be NULL (computed, as search.cu:102-108).
 * exact != 0 adds ||q||^2 (broadcast.cu:15-38). k > n pads with -1/FLT_MAX;
 * n == 0 fills -1/FLT_MAX (search.cu:93-100). Synchronises `stream`. */
rplsh_status rplsh_search_l2_dev(const float* x, int64_t n, int d, const float* xnorms,
                                 const float* q, int64_t nq, int k, int32_t* indices,
                                 float* distances, int exact, void* stream);

/* No reference counterpart: rplsh_search_l2_dev for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, semantics and results;
 * every temporary comes from a fixed device scratch array of
 * rplsh_search_l2_noalloc_scratch_bytes() (64 MiB) per device that the
 * library holds until the process ends. No hipMalloc/hipFree, pinned-host
 * allocation or HIP event creation. Sizes whose temporaries do not fit
 * (depth-0 workspace plus a retry of up to 64 queries; norms when xnorms is
 * NULL; padded rows when x is not 16-byte aligned with d % 4 == 0) are
 * rejected with RPLSH_ERR_INVALID_ARG before any work; a deeper retry that
 * runs out of scratch fails with RPLSH_ERR_OUT_OF_MEMORY. Calls that overlap
 * on one device race on the shared scratch: serialise them. */
rplsh_status rplsh_search_l2_dev_noalloc(const float* x, int64_t n, int d, const float* xnorms,
                                         const float* q, int64_t nq, int k, int32_t* indices,
                                         float* distances, int exact, void* stream);
size_t rplsh_search_l2_noalloc_scratch_bytes(void);

/* l2Select (include/internal/select.h:29-34, src/internal/select.cu:189-224):
 * per row r of products [rows][n] (row stride ldp), the k smallest
 * v = base_norms[i] + products[r][i] (one FP32 add), ascending, ties by the
 * lower i; outputs [rows][k]; k > n pads FLT_MAX / -1. k <= 4096. Device
 * pointers; returns when the results are in place. */
rplsh_status rplsh_l2_select_dev(const float* products, int64_t rows, int64_t n, int64_t ldp,
                                 const float* base_norms, int k, float* out_dist, int32_t* out_idx,
                                 void* stream);
/* blockSelect, both overloads (include/internal/select.h:8-24): per row of
 * in [rows][n] the k smallest values (select_max == 0, ascending) or the k
 * largest (descending), labelled by the column (labels == NULL) or by
 * labels[r][i]; ties by the lower label; k > n pads FLT_MAX (-FLT_MAX for the
 * largest) / -1. k <= 4096. Device pointers; returns when done. */
rplsh_status rplsh_block_select_dev(const float* in, const int32_t* labels, int64_t rows, int64_t n,
                                    int k, int select_max, float* out_k, int32_t* out_v, void* stream);
/* broadcastAlongRowsSum (include/internal/broadcast.h:7-8,
 * src/internal/broadcast.cu:15-74): m[r][c] += v[r]. Device pointers,
 * enqueued on `stream`. */
rplsh_status rplsh_broadcast_rows_sum_dev(const float* v, float* m, int64_t rows, int64_t cols,
                                          void* stream);

/* ------------------------------------------------------------ IndexBF -- */
/* IndexBFConfig + IndexConfig (include/index_bf.h:8-19, include/index.h:8-16)
 * plus the precision-form knobs of this library. */
typedef struct rplsh_bf_config {
  int device;                  /* IndexConfig::device */
  int memory_space;            /* IndexConfig::memorySpace: 0 Device, 1 Unified */
  int store_transposed;        /* IndexBFConfig::storeTransposed (accepted; layout is
                                  always the row-major + packed copies below) */
  int compute_exact_distances; /* IndexBFConfig::computeExactDistances; IndexBF::search
                                  always returns exact distances (index_bf.cpp:68-69) */
  int precision;               /* rplsh_precision */
  int kprime_factor;           /* over-fetch k' = kprime_factor * k for BF16/FP8 (0 -> 4) */
  int fp64_boundary;           /* FP64 re-rank of the boundary band (BF16/FP8), default 1 */
  int persistent_max_queries;  /* BF16/FP8 searches of up to this many queries (k' <= 512)
                                  run in the persistent kernel (one launch per step: prep,
                                  each filter level and its select, re-rank; the same
                                  results); 0: never. Default 32 (one query tile: there
                                  the persistent filter streams the copy fastest) */
  int residual_clusters;       /* BF16/FP8 on clustered rows: > 0 builds a second candidate
                                  copy holding each row minus the nearest of this many
                                  centroids (farthest-point seeding + Lloyd on a sample),
                                  in tiles of one centroid; the full filter level then
                                  contracts (q - c).(x - c) with the exact FP32 terms
                                  ||x - c||^2 and q.c - ||c||^2/2 in the C operand, which
                                  keeps the low-precision error at the scale of a
                                  cluster's spread. 0: off (default). Searches with it
                                  run on the launched pipeline. */
} rplsh_bf_config;

/* Fills the defaults: device 0, Device memory, FP32, k' factor 4, boundary on,
 * persistent kernel up to 256 queries. */
void rplsh_bf_config_default(rplsh_bf_config* cfg);

typedef struct rplsh_bf rplsh_bf;

/* IndexBF::IndexBF (index_bf.cpp:8-11) / Index::Index (index.cpp:8-27):
 * dim > 0 and device < count, else RPLSH_ERR_INVALID_ARG. `res` is not
 * owned and must outlive the index (index.h:75). */
rplsh_status rplsh_bf_create(rplsh_resources* res, int dim, const rplsh_bf_config* cfg,
                             rplsh_bf** out);
rplsh_status rplsh_bf_destroy(rplsh_bf* index);
/* IndexBF::train: a no-op, the index needs no training (index_bf.cpp:13-15). */
rplsh_status rplsh_bf_train(rplsh_bf* index, int64_t n, const float* x);
/* IndexBF::add (index_bf.cpp:26-44): copies x (host or device) into an owned
 * device buffer, computes the squared norms and the low-precision copy of the
 * configured precision. Unlike the reference's one-shot add, repeated adds
 * append (labels continue sequentially). n <= 0 is a no-op. */
rplsh_status rplsh_bf_add(rplsh_bf* index, int64_t n, const float* x);
/* IndexBF::search (index_bf.cpp:46-74): k nearest by exact squared L2,
 * ascending, ties by lower index; host or device pointers for queries and
 * outputs; k > n pads with -1 / FLT_MAX. nq <= 0 is a no-op. Returns when the
 * results are in place. */
rplsh_status rplsh_bf_search(rplsh_bf* index, int64_t nq, const float* queries, int k,
                             int32_t* indices, float* distances);
/* IndexBF::reset (index_bf.cpp:17-24). */
rplsh_status rplsh_bf_reset(rplsh_bf* index);
/* Index files (no reference counterpart: the reference lists save/load as
 * missing). save: configuration, installed tuning entries and the FP32 rows;
 * load: a new index on `device` (< 0: the saved one) with the same rows,
 * configuration and tuning; norms and candidate copies are rebuilt. */
rplsh_status rplsh_bf_save(rplsh_bf* index, const char* path);
rplsh_status rplsh_bf_load(rplsh_resources* res, const char* path, int device, rplsh_bf** out);
rplsh_status rplsh_bf_num_data(const rplsh_bf* index, int64_t* n);
rplsh_status rplsh_bf_dimension(const rplsh_bf* index, int* d);
rplsh_status rplsh_bf_device(const rplsh_bf* index, int* device);
rplsh_status rplsh_bf_set_precision(rplsh_bf* index, int precision, int kprime_factor);
/* Largest batch IndexBF::search runs in the persistent kernel (0: never);
 * results do not depend on it. */
rplsh_status rplsh_bf_set_persistent(rplsh_bf* index, int max_queries);
/* Device pointers of the resident data (row-major [n][d] f32, norms [n]). */
rplsh_status rplsh_bf_data_dev(const rplsh_bf* index, const float** data, const float** norms);

/* Counters of the last search on this index. */
typedef struct rplsh_search_stats {
  int64_t queries;          /* queries in the call */
  int64_t candidates;       /* total candidates kept for re-rank */
  int64_t survivors;        /* total survivors of the final filter level */
  int32_t levels;           /* filter levels run (sample passes + full pass) */
  int32_t retried_queries;  /* queries re-run because a threshold was too tight or
                               too loose, or (FP32) the candidate margin was not
                               certified */
  int32_t boundary_fp64;    /* candidates re-ranked in FP64 */
  int32_t kprime;           /* candidates per query re-ranked */
  float ms_total;           /* device time of the call */
  float ms_filter_full;     /* device time of the full-database filter launch
                               (HIP events on the search stream) */
  float ms_filter_samples;  /* device time of the sample-level filter launches */
  float ms_select;          /* device time of the last-level top-k' select */
  float ms_rerank;          /* device time of the re-rank */
  int64_t rows_filter_full; /* database rows the full filter launch scores: all
                               rows, less the tiles of a carried last sample
                               level (packed copies, LDS filter variants) */
  int32_t filter_variant;   /* packed copies: filter variant used (rplsh_tune_result) */
  int32_t target_wgs;       /* workgroups the filter grid aimed for */
  int32_t persistent;       /* 1: the search ran in the persistent kernel (its filter
                               scores every row of a level: rows_filter_full = n) */
} rplsh_search_stats;
rplsh_status rplsh_bf_last_stats(const rplsh_bf* index, rplsh_search_stats* out);

/* --------------------------------------------------------- autotuner -- */
/* Per-index tuning of the search for a given k and batch size (no reference
 * counterpart; the reference's chooseTileSize, search.cu:55-76, is a fixed
 * memory-budget heuristic):
 *  - kprime_factor (BF16/FP8): the smallest k'/k in {1,2,3,4,6,8,12,16}
 *    whose recall@k against the plain path's exact result on the sample
 *    reaches target_recall;
 *  - target_wgs: workgroups the filter grid aims for (query tiles x database
 *    slices), from {4096, 8192, 16384}, by measured batch time;
 *  - (c_factor, rho): the last level keeps ~c_factor * k' rows per query (its
 *    candidate queue; capacity 4x that); its threshold is the rho-th smallest
 *    key of a sample with stride c_factor * k' / rho. Pairs from
 *    {(3,auto),(4,16),(2,32),(3,32),(2,48),(2,64),(8,8)}, by time plus the
 *    expected cost of re-running queries that keep fewer than k' rows
 *    (probability P(Gamma(rho) < rho / c) per query, one extra pass per
 *    failing batch); the four fastest (grid, c, rho, variant) candidates of
 *    the sweep are timed again in five interleaved rounds and the lowest
 *    total wins (one timing cannot order candidates within a few percent);
 *  - variant (BF16/FP8): the filter's data movement and tile shape, by time,
 *    among LDS wide (0), LDS narrow (1) and ring wide (3); direct (2), ring
 *    narrow (4) and LDS wide with 8-wave workgroups (5) can be installed.
 *    With LDS wide or narrow chosen, the sample levels are then timed with
 *    the other of the two (same survivors) and run with it if faster;
 *  - the full level in L2-sized slices (variant bit 16): timed against the
 *    chosen grid in five interleaved rounds and kept unless more than 0.5 %
 *    slower (the same speed with fewer bytes over HBM).
 * target_recall <= 0 keeps the configured k' factor and tunes the rest.
 * The result is installed on the index and used by later searches with the
 * same k (the entry with the nearest batch size). */
typedef struct rplsh_tune_result {
  int k;
  int64_t batch_nq;
  int kprime_factor;  /* 0 for the plain path */
  int target_wgs;
  int c_factor;
  int variant;        /* bits 0-7: the full level's filter variant: 0 LDS wide, 1 LDS
                         narrow, 2 direct, 3 ring wide, 4 ring narrow, 5 LDS wide with
                         8-wave workgroups (BF16/FP8 only); bits 8-15: 1 + the sample
                         levels' variant (0: the same as the full level); bit 16: the
                         full level in slices of at most 576 KiB of rows, which keeps
                         the query blocks' re-reads of a slice in L2 (launched
                         searches; the persistent kernel ignores it) */
  int rho;            /* rank of the last threshold within its sample (0: automatic) */
  float recall;       /* recall@k of the chosen k' on the sample (1 for plain) */
  float ms_default;   /* batch time with the built-in defaults */
  float ms_tuned;     /* batch time with the chosen parameters */
} rplsh_tune_result;

/* sample: nsample queries (host or device); batch_nq: batch size to time
 * (the sample is repeated to fill it). */
rplsh_status rplsh_bf_autotune(rplsh_bf* index, const float* sample, int64_t nsample, int k,
                               int64_t batch_nq, float target_recall, rplsh_tune_result* out);
/* Installs a result (for example one loaded from a cache file). */
rplsh_status rplsh_bf_set_tuning(rplsh_bf* index, const rplsh_tune_result* t);
rplsh_status rplsh_bf_clear_tuning(rplsh_bf* index);

/* --------------------------------------------------------- IndexRPLSH -- */
/* Random-projection LSH front end before the exact re-rank: the "RPLSH" step
 * of the reference's name, which the reference does not implement (its only
 * piece is the Gaussian projection initialiser matrixRandomInitialize,
 * src/internal/matrix_ops.cu:71-76). Approximate: a row is a candidate only
 * if it shares a probed bucket with the query.
 *  - projections W [tables*bits][dim]: standard normal, from the library's
 *    seeded generator (rplsh_gen_gaussian(W, tables*bits, dim, seed, 0));
 *  - code of x in table t: bit b set iff the canonical FP32 dot (fmaf chain
 *    over j, as the distance dot) of W[t*bits+b] and x is >= the offset
 *    o[t*bits+b]: 0 (hyperplanes through the origin), or with `center` the
 *    projection's median over the first min(n, 65536) rows of the first add
 *    (element m/2 of the m sorted dots), set again after a reset;
 *  - per table, rows stably sorted by code (a bucket is a range, rows
 *    ascending within it);
 *  - a query probes, in each table, its own bucket and then, for p = 1..probes,
 *    the bucket with the bit of its p-th smallest |dot - offset| flipped
 *    (ties: lower bit); buckets are concatenated in (probe, table) order, cut after
 *    max_candidates entries and de-duplicated;
 *  - the distinct candidates (or, with `prerank`, the prerank*k of them with
 *    the smallest BF16 distance) are re-ranked exactly (canonical FP32, as the
 *    plain path): k smallest by (distance, index), padded with -1 / FLT_MAX. */
typedef struct rplsh_lsh_config {
  int device;
  int tables;          /* hash tables, >= 1 */
  int bits;            /* bits per table code, 1..24 */
  int probes;          /* extra single-bit-flip buckets per table, 0..bits;
                          (probes + 1) * tables <= 1024 */
  int max_candidates;  /* bucket entries taken per query before de-duplication, 1..8192;
                          above 4096 only with a pre-rank (the exact re-rank takes at
                          most 4096 per query: a search whose min(prerank*k,
                          max_candidates) exceeds that fails with UNSUPPORTED) */
  uint64_t seed;       /* projection seed */
  int center;          /* 1: median offsets (L2 proximity), 0: through the origin (angle) */
  int prerank;         /* 0: exact re-rank of every distinct candidate; f > 0 (when
                          f*k < max_candidates): a BF16 pre-rank keeps the f*k smallest
                          (rel_b, index) for the exact re-rank, rel_b = fmaf(-2, dot_b,
                          ||x||^2) with dot_b the product of the RNE-encoded row and
                          query in piece/tree order: P = ceil(dim/8) zero-padded pieces
                          of 8, G = P rounded up to a power of two partial sums (sum p:
                          fmaf over piece p's elements in order, from +0; sums past P
                          are +0), then s[p] += s[p+m] for m = G/2 .. 1; result s[0].
                          Keeps a BF16 copy of the rows (2*dim bytes per row); 0..64 */
} rplsh_lsh_config;

/* Fills the defaults: device 0, 16 tables x 16 bits, no extra probes,
 * 1024 candidates, seed 1234, median offsets, no pre-rank. */
void rplsh_lsh_config_default(rplsh_lsh_config* cfg);

typedef struct rplsh_lsh rplsh_lsh;

typedef struct rplsh_lsh_stats {
  int64_t queries;
  int64_t raw_candidates;       /* bucket entries taken (sum over queries) */
  int64_t distinct_candidates;  /* after de-duplication: rows re-ranked */
  float ms_total;               /* device time of the call */
  float ms_candidates;          /* query hashing and candidate gathering */
  float ms_rerank;              /* exact re-rank */
} rplsh_lsh_stats;

rplsh_status rplsh_lsh_create(rplsh_resources* res, int dim, const rplsh_lsh_config* cfg,
                              rplsh_lsh** out);
rplsh_status rplsh_lsh_destroy(rplsh_lsh* index);
/* Appends rows (host or device pointer), hashes them and rebuilds the tables. */
rplsh_status rplsh_lsh_add(rplsh_lsh* index, int64_t n, const float* x);
rplsh_status rplsh_lsh_reset(rplsh_lsh* index);
/* k approximate nearest neighbours with exact squared L2 distances,
 * ascending, ties by lower index; host or device pointers. */
rplsh_status rplsh_lsh_search(rplsh_lsh* index, int64_t nq, const float* queries, int k,
                              int32_t* indices, float* distances);
rplsh_status rplsh_lsh_last_stats(const rplsh_lsh* index, rplsh_lsh_stats* out);
rplsh_status rplsh_lsh_num_data(const rplsh_lsh* index, int64_t* n);
rplsh_status rplsh_lsh_get_config(const rplsh_lsh* index, rplsh_lsh_config* out);
/* W [tables*bits][dim] into host memory. */
rplsh_status rplsh_lsh_projections(const rplsh_lsh* index, float* out);
/* Offsets o [tables*bits] into host memory. */
rplsh_status rplsh_lsh_offsets(const rplsh_lsh* index, float* out);
/* Host copies [tables][n] of the codes in row order, the sorted codes and the
 * rows in bucket order; any of the three may be NULL. */
rplsh_status rplsh_lsh_tables(const rplsh_lsh* index, uint32_t* codes, uint32_t* sorted,
                              int32_t* perm);
/* Codes [tables][nq] and (if not NULL) margins dot - offset
 * [nq][tables*bits] of queries (host or device; outputs host or device). */
rplsh_status rplsh_lsh_hash_queries(rplsh_lsh* index, int64_t nq, const float* queries,
                                    uint32_t* codes, float* margins);

/* --------------------------------------------------------- HBM budget -- */
/* DeviceMemory / DeviceMemoryReservation (include/device_memory.h:12-89):
 * a byte budget on one device with reusable blocks. rplsh_memory_reserve
 * returns a block of at least `bytes` (a cached free block of at most twice
 * the size when there is one, else a new hipMalloc), charged to the budget
 * until released; it waits up to timeout_s (< 0: without limit, 0: not at
 * all) while other reservations hold the budget and fails with
 * RPLSH_ERR_BUDGET when the bytes never or not yet fit. Cached blocks count
 * against the budget and are freed when a new allocation needs the room.
 * rplsh_memory_release waits for the stream the block was reserved on, then
 * keeps the block for reuse. A serving loop admits its batches against such a
 * budget (its own, or one passed in rplsh_server_config.memory), so
 * reservations taken from it and the loop's batches share one bound.
 * Thread-safe. */
typedef struct rplsh_memory rplsh_memory;

typedef struct rplsh_memory_stats {
  int device;
  size_t budget_bytes;
  size_t reserved_bytes;            /* charged now (blocks and accounting charges) */
  size_t peak_reserved_bytes;
  size_t cached_bytes;              /* free blocks kept for reuse */
  size_t available_without_malloc;  /* largest reservation served from the cache now */
  size_t high_water_malloc_bytes;   /* peak bytes obtained from hipMalloc */
  int64_t reservations, waits, mallocs;
} rplsh_memory_stats;

/* budget_bytes 0: a quarter of the device's free memory. */
rplsh_status rplsh_memory_create(int device, size_t budget_bytes, rplsh_memory** out);
/* Fails with RPLSH_ERR_STATE while reservations are held. */
rplsh_status rplsh_memory_destroy(rplsh_memory* m);
rplsh_status rplsh_memory_reserve(rplsh_memory* m, void* stream, size_t bytes, double timeout_s,
                                  void** ptr);
rplsh_status rplsh_memory_release(rplsh_memory* m, void* ptr);
rplsh_status rplsh_memory_get_stats(rplsh_memory* m, rplsh_memory_stats* out);

/* ----------------------------------------------------------- serving -- */
/* Continuous-batching request loop over resident indexes (no reference
 * counterpart). `slots` worker threads each own a HIP stream and a workspace;
 * an idle worker admits the oldest queued request -- plus the queued
 * requests for the same index and k behind it, up to max_batch_queries --
 * when the batch's workspace fits in what is left of the HBM budget, so
 * requests arriving while others run are admitted as slots and memory free
 * up, and their kernels overlap on the device. A request whose workspace can
 * never fit is completed with RPLSH_ERR_BUDGET. Batches up to
 * graph_max_queries run as a replay of a hipGraph captured once per (index,
 * k, power-of-two batch bucket); the live query count is read on device, so
 * one graph serves every batch size of its bucket; a graph is keyed by the
 * index, k, the bucket and the search parameters the batch resolves to (the
 * index's tuning entry nearest its size), so a replay runs exactly what an
 * eager batch would. Results equal those of rplsh_bf_search on the same
 * queries. The budget covers the workspaces of admitted batches (plus a
 * reserve for re-running up to 64 failed queries each) and those held by
 * cached graphs. Request queries and results pass through pinned staging
 * buffers owned by each slot. While a server holds an index (from
 * rplsh_server_add_index until rplsh_server_destroy), rplsh_bf_add, _reset,
 * _set_precision, _autotune, _set_tuning, _clear_tuning and _destroy on it
 * fail with RPLSH_ERR_STATE. */
typedef struct rplsh_server rplsh_server;

typedef struct rplsh_server_config {
  int device;
  int slots;               /* concurrent batches (default 4) */
  size_t hbm_budget_bytes; /* admission budget for batch workspaces (0 = 1/4 of free HBM) */
  int max_batch_queries;   /* co-batching limit (default 4096) */
  int use_graphs;          /* replay captured hipGraphs for batches up to
                              graph_max_queries (default 1) */
  int graph_max_queries;   /* largest batch bucket captured (default 1024) */
  size_t graph_cache_bytes;/* device bytes of cached graph workspaces per server
                              (default 8 GiB); least recently used are evicted */
  rplsh_memory* memory;    /* budget to admit against (not owned, must outlive the
                              server); NULL: the server's own, of hbm_budget_bytes */
  int persistent_max_queries; /* BF16 / FP8 batches of up to this many queries (with
                              k' <= 512) run in the persistent kernel, which takes up to
                              16 queued batches of one element kind and padded
                              dimension -- different indexes and k -- in one run (a
                              launch per step: prep, each filter level and its select,
                              re-rank); 0: off (default 256) */
} rplsh_server_config;

typedef struct rplsh_request_info {
  int status;           /* rplsh_status of the request itself */
  double latency_ms;    /* submit -> results in place */
  double queue_ms;      /* submit -> admitted */
  int batch_requests;   /* requests in the batch it ran in */
  int64_t batch_queries;
  float device_ms;      /* device time of that batch: query upload, search and
                           result download (HIP events on the slot's stream) */
  int kernel_batches;   /* batches that shared its persistent-kernel run (0: it ran
                           on the launched pipeline) */
} rplsh_request_info;

typedef struct rplsh_server_stats {
  int64_t submitted, completed, rejected, queries, batches;
  double p50_ms, p99_ms, mean_ms, max_ms; /* latency of completed requests */
  double wall_s;                          /* first submit -> last completion */
  double qps;                             /* completed queries / wall_s */
  size_t peak_reserved_bytes;             /* peak sum of admitted batch workspaces */
  int peak_inflight;                      /* peak concurrent batches */
  int64_t graph_launches;                 /* batches run by replaying a captured graph */
  int64_t graph_captures;                 /* graphs captured */
  size_t graph_cache_bytes;               /* device bytes held by cached graphs */
  int64_t persistent_launches;            /* persistent-kernel runs (groups of batches) */
  int64_t persistent_batches;             /* batches they ran */
} rplsh_server_stats;

void rplsh_server_config_default(rplsh_server_config* cfg);
rplsh_status rplsh_server_create(const rplsh_server_config* cfg, rplsh_server** out);
/* Drains the queue, stops the workers. */
rplsh_status rplsh_server_destroy(rplsh_server* s);
/* Registers a resident index (not owned); its id is returned. */
rplsh_status rplsh_server_add_index(rplsh_server* s, rplsh_bf* index, int* index_id);
/* Enqueues a request; host pointers, kept alive by the caller until it completes. */
rplsh_status rplsh_server_submit(rplsh_server* s, int index_id, int64_t nq, const float* queries,
                                 int k, int32_t* indices, float* distances, uint64_t* request);
rplsh_status rplsh_server_poll(rplsh_server* s, uint64_t request, int* done);
/* Waits up to timeout_s (< 0: forever) for a request; RPLSH_ERR_STATE on timeout. */
rplsh_status rplsh_server_wait(rplsh_server* s, uint64_t request, double timeout_s,
                               rplsh_request_info* info);
rplsh_status rplsh_server_drain(rplsh_server* s);
rplsh_status rplsh_server_get_stats(rplsh_server* s, rplsh_server_stats* out);
rplsh_status rplsh_server_reset_stats(rplsh_server* s);
/* The budget the server admits against (its own or the one it was given). */
rplsh_status rplsh_server_memory(rplsh_server* s, rplsh_memory** out);
/* Device bytes one batch of nq queries with this k needs on an index. */
rplsh_status rplsh_bf_batch_bytes(rplsh_bf* index, int64_t nq, int k, size_t* bytes);

/* --------------------------------------------------- synthetic data -- */
/* Counter-based, thread-count independent generators (splitmix64 hashing +
 * Box-Muller in double); identical output on every x86-64 host running the
 * same libm. Host pointers, row-major [n][d]. */
rplsh_status rplsh_gen_uniform(float* out, int64_t n, int d, uint64_t seed, float lo, float hi);
rplsh_status rplsh_gen_gaussian(float* out, int64_t n, int d, uint64_t seed, int unit_normalize);
/* Gaussian mixture: centres U[-1,1)^d from center_seed; point i = centre[c_i]
 * + sigma * N(0, I), c_i and noise from point_seed. dup_fraction of the rows
 * (Bernoulli per row) are overwritten with exact copies of non-duplicate
 * rows. */
rplsh_status rplsh_gen_mixture(float* out, int64_t n, int d, int centers, float sigma,
                               uint64_t center_seed, uint64_t point_seed, double dup_fraction);

/* ---- datasets (include/dataset.h:14-204) -----------------------------------
 * TEXMEX vector files: records [int32 dim][dim elements], elem_bytes 4 for
 * .fvecs (float) / .ivecs (int32), 1 for .bvecs (uint8).
 * info: record count and dimension (0, 0 for an empty file); fails unless the
 * file is a whole number of records. read: records [first, first+count) packed
 * as [count][dim] into host memory (on_device = 0) or device memory of
 * `device` (streamed through pinned staging), checking every record's header
 * (loadXvecsData, dataset.h:14-44). write: the inverse (saveXvecsData, a TODO
 * in the reference, dataset.h:46); append != 0 appends. */
rplsh_status rplsh_xvecs_info(const char* path, int elem_bytes, int64_t* num, int* dim);
rplsh_status rplsh_xvecs_read(const char* path, int elem_bytes, int64_t first, int64_t count,
                              void* out, int on_device, int device);
rplsh_status rplsh_xvecs_write(const char* path, int elem_bytes, const void* data, int64_t num,
                               int dim, int append);
/* Dataset::evaluate (dataset.h:135-152): mean over nq queries of
 * |first k of result row ∩ first k of truth row| / k; host pointers, inputs
 * left untouched; -1 entries of result never count; missed[q] = k - hits
 * (optional, the reference's `diff`). */
rplsh_status rplsh_recall_at(const int32_t* result, int64_t nq, int ld_result, const int32_t* truth,
                             int ld_truth, int k, int32_t* missed, double* recall);


#ifdef __cplusplus
}
#endif
#endif /* RPLSH_H */
