/*
 * CPU oracle for the brute-force L2 kNN path -- TEST INFRASTRUCTURE ONLY.
 *
 * Used by tests/, the smoke check and the benchmark's CPU-baseline leg, never
 * by the product. See oracle.h for the canonical FP32 operation order.
 *
 * Build: oracle/Makefile (gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off).
 * -ffp-contract=off keeps gcc from fusing a*b+c outside the explicit fmaf
 * calls, so the canonical order is exactly what the source says.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Rows per block of the transposed database tile: the inner loop runs over
 * RB rows with one independent fmaf chain per row, which vectorises. */
#define RB 64

/* ---------------------------------------------------------------- norms -- */

/* Follows computeL2Norm (src/internal/norm.cu:191-202, squared=true): the
 * squared L2 norm of every row, here in the canonical fmaf order. */
void oracle_norms_canonical(const float* x, int64_t n, int d, float* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const float* r = x + i * (int64_t)d;
    float s = 0.0f;
    for (int j = 0; j < d; ++j) s = fmaf(r[j], r[j], s);
    out[i] = s;
  }
}

float oracle_dot_canonical(const float* a, const float* b, int d) {
  float s = 0.0f;
  for (int j = 0; j < d; ++j) s = fmaf(a[j], b[j], s);
  return s;
}

/* ------------------------------------------------- bounded (key,idx) heap -- */

typedef struct {
  float key;
  int32_t idx;
} kv32;

/* (a < b) in (key, index) lexicographic order. */
static inline int kv32_less(float ak, int32_t ai, float bk, int32_t bi) {
  return ak < bk || (ak == bk && ai < bi);
}

/* Max-heap of size k on (key, idx); heap[0] is the current worst kept entry. */
static void heap32_push(kv32* h, int k, float key, int32_t idx) {
  if (!kv32_less(key, idx, h[0].key, h[0].idx)) return;
  int i = 0;
  for (;;) {
    int l = 2 * i + 1, r = l + 1, m = i;
    float mk = key;
    int32_t mi = idx;
    if (l < k && kv32_less(mk, mi, h[l].key, h[l].idx)) { m = l; mk = h[l].key; mi = h[l].idx; }
    if (r < k && kv32_less(mk, mi, h[r].key, h[r].idx)) { m = r; }
    if (m == i) break;
    h[i] = h[m];
    i = m;
  }
  h[i].key = key;
  h[i].idx = idx;
}

static int kv32_cmp(const void* a, const void* b) {
  const kv32* x = (const kv32*)a;
  const kv32* y = (const kv32*)b;
  if (kv32_less(x->key, x->idx, y->key, y->idx)) return -1;
  if (kv32_less(y->key, y->idx, x->key, x->idx)) return 1;
  return 0;
}

/* ------------------------------------------ canonical FP32 restatement -- */

void oracle_search_canonical(const float* x, int64_t n, int d, const float* xnorms,
                             const float* q, int64_t nq, int k, int exact,
                             int32_t* out_idx, float* out_dist, int nthreads) {
  if (nq <= 0 || k <= 0) return;
  /* Empty index: FLT_MAX / -1 everywhere (search.cu:93-100). */
  if (n <= 0) {
    for (int64_t i = 0; i < nq * (int64_t)k; ++i) {
      out_idx[i] = -1;
      out_dist[i] = FLT_MAX;
    }
    return;
  }
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  int nt = 1;
  (void)nthreads;
#endif
  float* xn_own = NULL;
  if (!xnorms) {
    xn_own = (float*)malloc(sizeof(float) * (size_t)n);
    oracle_norms_canonical(x, n, d, xn_own);
    xnorms = xn_own;
  }
  float* qn = (float*)malloc(sizeof(float) * (size_t)nq);
  oracle_norms_canonical(q, nq, d, qn);

  /* Heap slots: k, but never more than n real candidates (k > n pads). */
  const int kk = k;
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
    int tn = omp_get_num_threads();
#else
    int tid = 0, tn = 1;
#endif
    int64_t q0 = nq * tid / tn, q1 = nq * (tid + 1) / tn;
    int64_t mq = q1 - q0;
    if (mq > 0) {
      kv32* heaps = (kv32*)malloc(sizeof(kv32) * (size_t)kk * (size_t)mq);
      for (int64_t i = 0; i < (int64_t)kk * mq; ++i) {
        heaps[i].key = INFINITY;
        heaps[i].idx = INT32_MAX;
      }
      float* xt = (float*)aligned_alloc(64, sizeof(float) * (size_t)d * RB);
      float acc[RB] __attribute__((aligned(64)));
      for (int64_t r0 = 0; r0 < n; r0 += RB) {
        int rb = (int)((n - r0) < RB ? (n - r0) : RB);
        /* Transpose the row block: xt[j][r] = x[r0 + r][j]; pad rows with 0. */
        for (int r = 0; r < RB; ++r) {
          const float* src = x + (r0 + (r < rb ? r : 0)) * (int64_t)d;
          for (int j = 0; j < d; ++j) xt[(int64_t)j * RB + r] = r < rb ? src[j] : 0.0f;
        }
        for (int64_t qi = 0; qi < mq; ++qi) {
          const float* qv = q + (q0 + qi) * (int64_t)d;
          for (int r = 0; r < RB; ++r) acc[r] = 0.0f;
          for (int j = 0; j < d; ++j) {
            const float qj = qv[j];
            const float* xr = xt + (int64_t)j * RB;
            for (int r = 0; r < RB; ++r) acc[r] = fmaf(qj, xr[r], acc[r]);
          }
          kv32* h = heaps + qi * kk;
          float top = h[0].key;
          for (int r = 0; r < rb; ++r) {
            float rel = fmaf(-2.0f, acc[r], xnorms[r0 + r]);
            if (rel <= top) {
              heap32_push(h, kk, rel, (int32_t)(r0 + r));
              top = h[0].key;
            }
          }
        }
      }
      for (int64_t qi = 0; qi < mq; ++qi) {
        kv32* h = heaps + qi * kk;
        qsort(h, (size_t)kk, sizeof(kv32), kv32_cmp);
        int64_t row = q0 + qi;
        for (int j = 0; j < kk; ++j) {
          if (h[j].idx == INT32_MAX) {
            out_idx[row * kk + j] = -1;
            out_dist[row * kk + j] = FLT_MAX;
          } else {
            out_idx[row * kk + j] = h[j].idx;
            out_dist[row * kk + j] = exact ? h[j].key + qn[row] : h[j].key;
          }
        }
      }
      free(xt);
      free(heaps);
    }
  }
  free(qn);
  free(xn_own);
}

/* ---------------------------------------------------- FP64 brute force -- */

typedef struct {
  double key;
  int32_t idx;
} kv64;

static inline int kv64_less(double ak, int32_t ai, double bk, int32_t bi) {
  return ak < bk || (ak == bk && ai < bi);
}

static void heap64_push(kv64* h, int k, double key, int32_t idx) {
  if (!kv64_less(key, idx, h[0].key, h[0].idx)) return;
  int i = 0;
  for (;;) {
    int l = 2 * i + 1, r = l + 1, m = i;
    double mk = key;
    int32_t mi = idx;
    if (l < k && kv64_less(mk, mi, h[l].key, h[l].idx)) { m = l; mk = h[l].key; mi = h[l].idx; }
    if (r < k && kv64_less(mk, mi, h[r].key, h[r].idx)) { m = r; }
    if (m == i) break;
    h[i] = h[m];
    i = m;
  }
  h[i].key = key;
  h[i].idx = idx;
}

static int kv64_cmp(const void* a, const void* b) {
  const kv64* x = (const kv64*)a;
  const kv64* y = (const kv64*)b;
  if (kv64_less(x->key, x->idx, y->key, y->idx)) return -1;
  if (kv64_less(y->key, y->idx, x->key, x->idx)) return 1;
  return 0;
}

void oracle_search_fp64(const float* x, int64_t n, int d, const float* q, int64_t nq,
                        int k, int32_t* out_idx, double* out_dist, int nthreads) {
  if (nq <= 0 || k <= 0) return;
#ifdef _OPENMP
  int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  int nt = 1;
  (void)nthreads;
#endif
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    int tid = omp_get_thread_num();
    int tn = omp_get_num_threads();
#else
    int tid = 0, tn = 1;
#endif
    int64_t q0 = nq * tid / tn, q1 = nq * (tid + 1) / tn;
    int64_t mq = q1 - q0;
    if (mq > 0) {
      kv64* heaps = (kv64*)malloc(sizeof(kv64) * (size_t)k * (size_t)mq);
      for (int64_t i = 0; i < (int64_t)k * mq; ++i) {
        heaps[i].key = INFINITY;
        heaps[i].idx = INT32_MAX;
      }
      double* xt = (double*)aligned_alloc(64, sizeof(double) * (size_t)d * RB);
      double acc[RB] __attribute__((aligned(64)));
      for (int64_t r0 = 0; r0 < n; r0 += RB) {
        int rb = (int)((n - r0) < RB ? (n - r0) : RB);
        for (int r = 0; r < RB; ++r) {
          const float* src = x + (r0 + (r < rb ? r : 0)) * (int64_t)d;
          for (int j = 0; j < d; ++j) xt[(int64_t)j * RB + r] = r < rb ? (double)src[j] : 0.0;
        }
        for (int64_t qi = 0; qi < mq; ++qi) {
          const float* qv = q + (q0 + qi) * (int64_t)d;
          for (int r = 0; r < RB; ++r) acc[r] = 0.0;
          for (int j = 0; j < d; ++j) {
            const double qj = (double)qv[j];
            const double* xr = xt + (int64_t)j * RB;
            for (int r = 0; r < RB; ++r) {
              double t = qj - xr[r];
              acc[r] = acc[r] + t * t;
            }
          }
          kv64* h = heaps + qi * k;
          for (int r = 0; r < rb; ++r) {
            if (acc[r] <= h[0].key) heap64_push(h, k, acc[r], (int32_t)(r0 + r));
          }
        }
      }
      for (int64_t qi = 0; qi < mq; ++qi) {
        kv64* h = heaps + qi * k;
        qsort(h, (size_t)k, sizeof(kv64), kv64_cmp);
        int64_t row = q0 + qi;
        for (int j = 0; j < k; ++j) {
          out_idx[row * k + j] = h[j].idx == INT32_MAX ? -1 : h[j].idx;
          out_dist[row * k + j] = h[j].idx == INT32_MAX ? INFINITY : h[j].key;
        }
      }
      free(xt);
      free(heaps);
    }
  }
}

void oracle_pair_dist_fp64(const float* x, int d, const float* q, int64_t nq, int k,
                           const int32_t* idx, double* out) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < nq; ++i) {
    const float* qv = q + i * (int64_t)d;
    for (int j = 0; j < k; ++j) {
      int32_t id = idx[i * k + j];
      if (id < 0) {
        out[i * k + j] = INFINITY;
        continue;
      }
      const float* xv = x + (int64_t)id * d;
      double s = 0.0;
      for (int t = 0; t < d; ++t) {
        double v = (double)qv[t] - (double)xv[t];
        s = s + v * v;
      }
      out[i * k + j] = s;
    }
  }
}

/* --------------------------------------------------------------- recall -- */

static int i32_cmp(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* Mean of |R ∩ G| / k over queries (include/dataset.h:135-152, with the
 * set intersection of src/internal/algorithms.cu:10-28), on copies. */
double oracle_recall(const int32_t* result, const int32_t* truth, int64_t nq, int k) {
  if (nq <= 0 || k <= 0) return 1.0;
  int32_t* a = (int32_t*)malloc(sizeof(int32_t) * (size_t)k);
  int32_t* b = (int32_t*)malloc(sizeof(int32_t) * (size_t)k);
  double total = 0.0;
  for (int64_t i = 0; i < nq; ++i) {
    memcpy(a, result + i * k, sizeof(int32_t) * (size_t)k);
    memcpy(b, truth + i * k, sizeof(int32_t) * (size_t)k);
    qsort(a, (size_t)k, sizeof(int32_t), i32_cmp);
    qsort(b, (size_t)k, sizeof(int32_t), i32_cmp);
    int ia = 0, ib = 0, hit = 0;
    while (ia < k && ib < k) {
      if (a[ia] < b[ib]) {
        ++ia;
      } else if (b[ib] < a[ia]) {
        ++ib;
      } else {
        if (a[ia] >= 0) ++hit;
        ++ia;
        ++ib;
      }
    }
    total += (double)hit / (double)k;
  }
  free(a);
  free(b);
  return total / (double)nq;
}

/* ------------------------------------------------------------- encoders -- */

static inline uint32_t f2u(float v) {
  uint32_t u;
  memcpy(&u, &v, 4);
  return u;
}

static inline float u2f(uint32_t u) {
  float v;
  memcpy(&v, &u, 4);
  return v;
}

uint16_t oracle_bf16_rne(float v) {
  uint32_t u = f2u(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x0040u);
  uint32_t lsb = (u >> 16) & 1u;
  u += 0x7fffu + lsb;
  return (uint16_t)(u >> 16);
}

float oracle_bf16_to_float(uint16_t h) { return u2f((uint32_t)h << 16); }

/* e4m3 (FN): bias 7, 3 mantissa bits, subnormal step 2^-9, max 448,
 * S.1111.111 = NaN, no infinities. Rounding is done on the exact value in
 * double: the quantum of the target binade is a power of two, so a / quantum
 * is exact and rint (round-half-even) gives the RNE result. */
uint8_t oracle_e4m3_rne_sat(float v) {
  if (isnan(v)) return 0x7f;
  uint8_t sign = signbit(v) ? 0x80 : 0x00;
  double a = fabs((double)v);
  if (a >= 448.0) return sign | 0x7e;
  double quantum;
  if (a < ldexp(1.0, -6)) {
    quantum = ldexp(1.0, -9);
  } else {
    int e;
    frexp(a, &e); /* a = f * 2^e, f in [0.5, 1) -> binade 2^(e-1) */
    quantum = ldexp(1.0, e - 1 - 3);
  }
  double r = rint(a / quantum) * quantum;
  if (r >= 448.0) return sign | 0x7e;
  if (r == 0.0) return sign;
  if (r < ldexp(1.0, -6)) {
    int m = (int)(r / ldexp(1.0, -9));
    return sign | (uint8_t)m;
  }
  int e;
  double f = frexp(r, &e); /* r = f * 2^e, f in [0.5,1) */
  int ebiased = (e - 1) + 7;
  int m = (int)((f * 2.0 - 1.0) * 8.0);
  return sign | (uint8_t)(ebiased << 3) | (uint8_t)m;
}

float oracle_e4m3_to_float(uint8_t b) {
  int s = b >> 7, e = (b >> 3) & 0xf, m = b & 7;
  if (e == 0xf && m == 7) return NAN;
  double v = e == 0 ? ldexp((double)m, -9) : ldexp(1.0 + m / 8.0, e - 7);
  return (float)(s ? -v : v);
}

/* ------------------------------------------------ random-projection LSH -- */
/* Restatement of the library's LSH front end (IndexRPLSH). The reference
 * has no LSH step to follow (its only piece is the Gaussian projection
 * initialiser, src/internal/matrix_ops.cu:71-76); this is the specification
 * the GPU path implements, written as plainly as possible. */

void oracle_lsh_hash(const float* x, int64_t n, int d, const float* w, const float* off, int L,
                     int K, uint32_t* codes, float* margins) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    const float* xr = x + i * (int64_t)d;
    for (int t = 0; t < L; ++t) {
      uint32_t c = 0;
      for (int b = 0; b < K; ++b) {
        const float s = oracle_dot_canonical(w + ((int64_t)t * K + b) * d, xr, d);
        const float o = off ? off[(int64_t)t * K + b] : 0.0f;
        if (s >= o) c |= 1u << b;
        if (margins) margins[i * (int64_t)L * K + (int64_t)t * K + b] = s - o;
      }
      codes[(int64_t)t * n + i] = c;
    }
  }
}

typedef struct {
  uint32_t code;
  int32_t row;
} code_row;

static int code_row_cmp(const void* a, const void* b) {
  const code_row* x = (const code_row*)a;
  const code_row* y = (const code_row*)b;
  if (x->code != y->code) return x->code < y->code ? -1 : 1;
  return (x->row > y->row) - (x->row < y->row);
}

/* Order of the probe bits: |dot| as its IEEE bit pattern (NaN after +inf),
 * then the bit index. */
static uint64_t probe_key(float m, int b) {
  return ((uint64_t)f2u(fabsf(m)) << 5) | (uint64_t)b;
}

/* BF16 dot product in piece/tree order (the LSH pre-rank's score; the order
 * is this library's own, chosen so that G lanes can share a row):
 * both vectors are rounded to BF16 (RNE) and zero padded to P = ceil(d / 8)
 * pieces of 8 elements; G = min(64, P rounded up to a power of two) partial
 * sums, sum p chaining fmaf over the elements of pieces p, p + G, ... in
 * element order from +0 (sums past P stay +0); then s[p] += s[p + m] for
 * m = G/2, ..., 1 (p < m), and the result is s[0]. */
float oracle_bf16_tree_dot(const float* a, const float* b, int d) {
  const int P = (d + 7) / 8;
  int G = 1;
  while (G < P && G < 64) G <<= 1;
  float s[64];
  for (int p = 0; p < G; ++p) {
    float acc = 0.0f;
    for (int c = p; c < P; c += G) {
      for (int t = 0; t < 8; ++t) {
        const int j = 8 * c + t;
        const float u = j < d ? oracle_bf16_to_float(oracle_bf16_rne(a[j])) : 0.0f;
        const float v = j < d ? oracle_bf16_to_float(oracle_bf16_rne(b[j])) : 0.0f;
        acc = fmaf(u, v, acc);
      }
    }
    s[p] = acc;
  }
  for (int m = G / 2; m >= 1; m >>= 1)
    for (int p = 0; p < m; ++p) s[p] = s[p] + s[p + m];
  return s[0];
}

void oracle_lsh_search(const float* x, int64_t n, int d, const float* w, const float* off, int L,
                       int K, int probes, int budget, int prerank, const float* q, int64_t nq, int k,
                       int32_t* out_idx, float* out_dist, int64_t* raw_total,
                       int64_t* distinct_total) {
  if (raw_total) *raw_total = 0;
  if (distinct_total) *distinct_total = 0;
  if (nq <= 0 || k <= 0) return;
  if (n <= 0) {
    for (int64_t i = 0; i < nq * (int64_t)k; ++i) {
      out_idx[i] = -1;
      out_dist[i] = FLT_MAX;
    }
    return;
  }
  uint32_t* codes = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n * L);
  oracle_lsh_hash(x, n, d, w, off, L, K, codes, NULL);
  code_row* tab = (code_row*)malloc(sizeof(code_row) * (size_t)n * L);
  for (int t = 0; t < L; ++t) {
    for (int64_t i = 0; i < n; ++i) {
      tab[(int64_t)t * n + i].code = codes[(int64_t)t * n + i];
      tab[(int64_t)t * n + i].row = (int32_t)i;
    }
    qsort(tab + (int64_t)t * n, (size_t)n, sizeof(code_row), code_row_cmp);
  }
  float* xn = (float*)malloc(sizeof(float) * (size_t)n);
  oracle_norms_canonical(x, n, d, xn);
  float* qn = (float*)malloc(sizeof(float) * (size_t)nq);
  oracle_norms_canonical(q, nq, d, qn);
  uint32_t* qc = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)nq * L);
  float* qm = (float*)malloc(sizeof(float) * (size_t)nq * L * K);
  oracle_lsh_hash(q, nq, d, w, off, L, K, qc, qm);
  int64_t raw = 0, distinct = 0;

#pragma omp parallel for schedule(dynamic, 4) reduction(+ : raw, distinct)
  for (int64_t qi = 0; qi < nq; ++qi) {
    int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)(budget > 0 ? budget : 1));
    int R = 0;
    for (int p = 0; p <= probes && R < budget; ++p) {
      for (int t = 0; t < L && R < budget; ++t) {
        uint32_t code = qc[(int64_t)t * nq + qi];
        if (p > 0) {
          /* the bit of the p-th smallest key among the K bits */
          const float* m = qm + qi * (int64_t)L * K + (int64_t)t * K;
          uint64_t keys[32];
          for (int b = 0; b < K; ++b) keys[b] = probe_key(m[b], b);
          for (int a = 1; a < K; ++a) { /* insertion sort, ascending */
            uint64_t v = keys[a];
            int j = a - 1;
            while (j >= 0 && keys[j] > v) {
              keys[j + 1] = keys[j];
              --j;
            }
            keys[j + 1] = v;
          }
          code ^= 1u << (int)(keys[p - 1] & 31u);
        }
        const code_row* s = tab + (int64_t)t * n;
        int64_t lo = 0, hi = n;
        while (lo < hi) { /* first entry with code >= probe */
          int64_t mid = (lo + hi) / 2;
          if (s[mid].code < code) lo = mid + 1; else hi = mid;
        }
        for (int64_t r = lo; r < n && s[r].code == code && R < budget; ++r) buf[R++] = s[r].row;
      }
    }
    raw += R;
    qsort(buf, (size_t)R, sizeof(int32_t), i32_cmp);
    int nd = 0;  /* distinct rows, compacted to the front of buf */
    for (int j = 0; j < R; ++j)
      if (j == 0 || buf[j] != buf[j - 1]) buf[nd++] = buf[j];
    distinct += nd;
    const float* qv = q + qi * (int64_t)d;
    int nc = nd;
    const int kpre = prerank > 0 ? (prerank * k < budget ? prerank * k : budget) : budget;
    if (prerank > 0 && kpre < budget && nd > kpre) {
      /* BF16 pre-rank: keep the kpre smallest (rel_b, index) */
      kv32* pre = (kv32*)malloc(sizeof(kv32) * (size_t)nd);
      for (int j = 0; j < nd; ++j) {
        pre[j].key = fmaf(-2.0f, oracle_bf16_tree_dot(qv, x + (int64_t)buf[j] * d, d), xn[buf[j]]);
        pre[j].idx = buf[j];
      }
      qsort(pre, (size_t)nd, sizeof(kv32), kv32_cmp);
      for (int j = 0; j < kpre; ++j) buf[j] = pre[j].idx;
      nc = kpre;
      free(pre);
    }
    kv32* h = (kv32*)malloc(sizeof(kv32) * (size_t)k);
    for (int j = 0; j < k; ++j) {
      h[j].key = INFINITY;
      h[j].idx = INT32_MAX;
    }
    for (int j = 0; j < nc; ++j) {
      const int32_t r = buf[j];
      const float rel = fmaf(-2.0f, oracle_dot_canonical(qv, x + (int64_t)r * d, d), xn[r]);
      heap32_push(h, k, rel, r);
    }
    nd = nc;
    qsort(h, (size_t)k, sizeof(kv32), kv32_cmp);
    for (int j = 0; j < k; ++j) {
      const int64_t o = qi * (int64_t)k + j;
      if (j < nd) {
        out_idx[o] = h[j].idx;
        out_dist[o] = h[j].key + qn[qi];
      } else {
        out_idx[o] = -1;
        out_dist[o] = FLT_MAX;
      }
    }
    free(h);
    free(buf);
  }
  if (raw_total) *raw_total = raw;
  if (distinct_total) *distinct_total = distinct;
  free(codes);
  free(tab);
  free(xn);
  free(qn);
  free(qc);
  free(qm);
}
