/*
 * oracle.c -- CPU checker for the jBLAS GEMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library or the
 * Python/Julia host layer) links, loads or calls this file.  It is used by the
 * tests, by smoke() and by the CPU-baseline leg of bench.py, and there only as
 * the checker (or as the timed CPU baseline), never as the thing measured on
 * the GPU.
 *
 * Contents (each function cites the reference it restates; paths are relative
 * to the jBLAS.jl repository):
 *   1. SplitMix64 counter-based input generator (replaces src/randmat.jl:5-14,
 *      which draws unseeded normals; the seeded uniform stream is this build's).
 *   2. orc_jmul_*  : restatement of jmul!'s per-element recurrence and tile loop
 *      order (src/gemm.jl:71-91 init product, :149-170 fma k-step, :3-11 store,
 *      :244-348 driver, cc outer / rc inner at :313), extended to the fringe
 *      that the reference drops (src/gemm.jl:266-267, :340-345).
 *   3. orc_gemm_recurrence : the same recurrence under BLAS alpha/beta/op
 *      semantics (the reference has none; Julia's LinearAlgebra.BLAS.gemm!
 *      argument meaning is used).
 *   4. orc_gemm_dot2 : FP64 oracle, compensated dot product "Dot2" of
 *      Ogita, Rump & Oishi, SIAM J. Sci. Comput. 26(6), 2005 (TwoProduct via
 *      fma, TwoSum of Knuth), with a double-double alpha/beta epilogue.
 *   5. orc_dgemm_ozaki : exact restatement of the split-BF16 emulated DGEMM
 *      (Ozaki, Ogita, Oishi & Rump, Numer. Algorithms 59, 2012, with
 *      exponent-scaled balanced 9-bit digits, |d| <= 256, each exact in BF16);
 *      integer level sums are formed exactly in int64, so the result bits are
 *      fully determined.  orc_dgemm_ozaki_i8: the same with signed 8-bit
 *      digits (the INT8-slice method).
 *   6. orc_pick_kernel_size / orc_blocking_structure : integer tile policy
 *      (src/kernel_structure.jl:76-99, src/memory_management.jl:78-177,
 *      src/miscellaneous.jl:8-14).
 *
 * Parity pinning (what checks each function; see tests/test_oracle.py):
 *   SplitMix64      -- published SplitMix64 first outputs for seed 0 and a
 *                      numpy restatement; uniform/wide are exact arithmetic.
 *   orc_jmul_*      -- bit-identical to a naive triple loop in the same order;
 *                      integer-valued KATs (exact for any order), identity,
 *                      permutation, all-ones, rank-1, power-of-two scaling.
 *   orc_gemm_dot2   -- sums with known exact values (cancellation cases that
 *                      plain summation gets wrong), fractions.Fraction brute
 *                      force on tiny inputs, numpy dgemm within gamma_K.
 *   orc_dgemm_ozaki -- inputs exactly representable in s digits give the
 *                      exact product; digits reconstruct the scaled input;
 *                      error vs Dot2 shrinks by at least 2^-7 per added slice
 *                      (9 bits per slice in the BF16 split, 8 in INT8).
 *   tile policy     -- hand-derived table for AVX2/AVX-512 register files.
 *
 * Build: oracle/Makefile (gcc, -ffp-contract=off so no fused operation is
 * introduced beyond the explicit fma() calls).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* 1. SplitMix64 counter-based generator                                     */
/* ------------------------------------------------------------------------ */
#define SM_GAMMA 0x9e3779b97f4a7c15ULL

static inline uint64_t sm_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
/* i-th output (0-based) of the SplitMix64 sequence started at state `seed`. */
ORC_API uint64_t orc_splitmix64(uint64_t seed, uint64_t i) { return sm_mix(seed + (i + 1) * SM_GAMMA); }

/* u = 2*((x>>11)*2^-53) - 1 in [-1, 1); exact arithmetic. */
ORC_API void orc_fill_uniform(uint64_t seed, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = 2.0 * ((double)(orc_splitmix64(seed, (uint64_t)i) >> 11) * 0x1.0p-53) - 1.0;
}
/* a = u * 2^e, u from stream index 2i, e = emin + floor((emax-emin+1)*v) from 2i+1. */
ORC_API void orc_fill_wide(uint64_t seed, int64_t n, int emin, int emax, double* out) {
  for (int64_t i = 0; i < n; ++i) {
    double u = 2.0 * ((double)(orc_splitmix64(seed, 2 * (uint64_t)i) >> 11) * 0x1.0p-53) - 1.0;
    double v = (double)(orc_splitmix64(seed, 2 * (uint64_t)i + 1) >> 11) * 0x1.0p-53;
    int e = emin + (int)floor((double)(emax - emin + 1) * v);
    out[i] = ldexp(u, e);
  }
}

/* ------------------------------------------------------------------------ */
/* 2. jmul! restatement                                                      */
/* ------------------------------------------------------------------------ */
/*
 * D (M x P) = A (M x N) * X (N x P), column-major with leading dimensions.
 * Per element (src/gemm.jl:86 then :165):
 *   acc_1 = fl(A[i,1] * X[1,j]);  acc_n = fma(A[i,n], X[n,j], acc_{n-1}), n = 2..N.
 * Tiles of rows x cols (pick_kernel_size), cc outer, rc inner (src/gemm.jl:313).
 * Unlike the reference, the fringe rows/columns are computed with the same
 * recurrence (the reference leaves them unwritten, src/gemm.jl:340-345).
 */
#define JMUL_TILE(NAME, T, FMA, R, C)                                                              \
  static void NAME(int64_t r0, int64_t c0, int64_t N, const T* A, int64_t lda, const T* X,        \
                   int64_t ldx, T* D, int64_t ldd) {                                              \
    T acc[C][R];                                                                                   \
    for (int c = 0; c < C; ++c) {                                                                  \
      const T x = X[(c0 + c) * ldx];                                                               \
      for (int r = 0; r < R; ++r) acc[c][r] = A[r0 + r] * x;                                       \
    }                                                                                              \
    for (int64_t n = 1; n < N; ++n) {                                                              \
      const T* a = A + n * lda + r0;                                                               \
      for (int c = 0; c < C; ++c) {                                                                \
        const T x = X[n + (c0 + c) * ldx];                                                         \
        for (int r = 0; r < R; ++r) acc[c][r] = FMA(a[r], x, acc[c][r]);                           \
      }                                                                                            \
    }                                                                                              \
    for (int c = 0; c < C; ++c)                                                                    \
      for (int r = 0; r < R; ++r) D[r0 + r + (c0 + c) * ldd] = acc[c][r];                          \
  }

JMUL_TILE(tile_d_40x5, double, fma, 40, 5)
JMUL_TILE(tile_d_12x4, double, fma, 12, 4)
JMUL_TILE(tile_s_80x5, float, fmaf, 80, 5)
JMUL_TILE(tile_s_24x4, float, fmaf, 24, 4)

/* Fringe: same recurrence, any tile extent. */
#define JMUL_EDGE(NAME, T, FMA)                                                                    \
  static void NAME(int64_t r0, int64_t nr, int64_t c0, int64_t nc, int64_t N, const T* A,         \
                   int64_t lda, const T* X, int64_t ldx, T* D, int64_t ldd) {                     \
    for (int64_t c = 0; c < nc; ++c)                                                               \
      for (int64_t r = 0; r < nr; ++r) {                                                           \
        T acc = A[r0 + r] * X[(c0 + c) * ldx];                                                     \
        for (int64_t n = 1; n < N; ++n) acc = FMA(A[r0 + r + n * lda], X[n + (c0 + c) * ldx], acc); \
        D[r0 + r + (c0 + c) * ldd] = acc;                                                          \
      }                                                                                            \
  }
JMUL_EDGE(edge_d, double, fma)
JMUL_EDGE(edge_s, float, fmaf)

#define JMUL_DRIVER(NAME, T, TILE_BIG, RB, CB, TILE_SMALL, RS, CS, EDGE)                          \
  ORC_API int NAME(int64_t M, int64_t N, int64_t P, const T* A, int64_t lda, const T* X,          \
                   int64_t ldx, T* D, int64_t ldd, int avx512, int nthreads) {                    \
    if (M < 0 || N < 1 || P < 0) return -1;                                                        \
    const int64_t rows = avx512 ? RB : RS, cols = avx512 ? CB : CS;                                \
    const int64_t rch = M / rows, cch = P / cols;                                                  \
    (void)nthreads;                                                                                \
    _Pragma("omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)")         \
    for (int64_t cc = 0; cc <= cch; ++cc) {                                                        \
      const int64_t c0 = cc * cols, nc = (cc < cch) ? cols : P - cch * cols;                       \
      if (nc == 0) continue;                                                                       \
      for (int64_t rc = 0; rc <= rch; ++rc) {                                                      \
        const int64_t r0 = rc * rows, nr = (rc < rch) ? rows : M - rch * rows;                     \
        if (nr == 0) continue;                                                                     \
        if (nr == rows && nc == cols) {                                                            \
          if (avx512) TILE_BIG(r0, c0, N, A, lda, X, ldx, D, ldd);                                 \
          else TILE_SMALL(r0, c0, N, A, lda, X, ldx, D, ldd);                                      \
        } else {                                                                                   \
          EDGE(r0, nr, c0, nc, N, A, lda, X, ldx, D, ldd);                                         \
        }                                                                                          \
      }                                                                                            \
    }                                                                                              \
    return 0;                                                                                      \
  }
JMUL_DRIVER(orc_jmul_d, double, tile_d_40x5, 40, 5, tile_d_12x4, 12, 4, edge_d)
JMUL_DRIVER(orc_jmul_s, float, tile_s_80x5, 80, 5, tile_s_24x4, 24, 4, edge_s)

/* Naive triple loop, same per-element order; pins the tiled restatement. */
ORC_API void orc_jmul_naive_d(int64_t M, int64_t N, int64_t P, const double* A, int64_t lda,
                              const double* X, int64_t ldx, double* D, int64_t ldd) {
  for (int64_t j = 0; j < P; ++j)
    for (int64_t i = 0; i < M; ++i) {
      double acc = A[i] * X[j * ldx];
      for (int64_t n = 1; n < N; ++n) acc = fma(A[i + n * lda], X[n + j * ldx], acc);
      D[i + j * ldd] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* 3. BLAS-semantics wrapper of the recurrence                               */
/* ------------------------------------------------------------------------ */
/* Element access of op(A) (m x k) and op(B) (k x n), column-major storage. */
static inline double opA(char t, const double* A, int64_t lda, int64_t i, int64_t p) {
  return (t == 'N' || t == 'n') ? A[i + p * lda] : A[p + i * lda];
}
static inline double opB(char t, const double* B, int64_t ldb, int64_t p, int64_t j) {
  return (t == 'N' || t == 'n') ? B[p + j * ldb] : B[j + p * ldb];
}

/*
 * C = alpha*op(A)*op(B) + beta*C with the jmul! recurrence for the dot
 * product and the epilogue c = alpha*acc, then c = fma(beta, c_old, c) when
 * beta != 0 (beta == 0 never reads C).  alpha == 0 or k == 0 gives beta*C.
 */
/*
 * Packed copies of op(A) (m x k) and op(B) (k x n), both column-major and
 * contiguous, so that the chains of 8 consecutive rows run side by side in
 * one vector (the per-element operation order is unchanged).
 */
static double* pack_op(char t, const double* X, int64_t ldx, int64_t rows, int64_t cols, int nthreads) {
  double* P = (double*)malloc(sizeof(double) * (size_t)(rows * cols + 1));
  if (!P) return NULL;
  const int tr = !(t == 'N' || t == 'n');
  (void)nthreads;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t c = 0; c < cols; ++c)
    for (int64_t r = 0; r < rows; ++r) P[r + c * rows] = tr ? X[c + r * ldx] : X[r + c * ldx];
  return P;
}

/* Chains for rows [i0, i0+nr) of column j over k in [c0, c1): acc[r] for each row. */
static inline void chain_block(const double* Ap, const double* Bp, int64_t m, int64_t i0, int nr, int64_t j,
                               int64_t k, int64_t c0, int64_t c1, double* acc) {
  const double* b = Bp + j * k;
  if (nr == 8) {
    double a8[8];
    for (int r = 0; r < 8; ++r) a8[r] = Ap[i0 + r + c0 * m] * b[c0];
    for (int64_t p = c0 + 1; p < c1; ++p) {
      const double* a = Ap + i0 + p * m;
      const double x = b[p];
      for (int r = 0; r < 8; ++r) a8[r] = fma(a[r], x, a8[r]);
    }
    for (int r = 0; r < 8; ++r) acc[r] = a8[r];
  } else {
    for (int r = 0; r < nr; ++r) {
      double s = Ap[i0 + r + c0 * m] * b[c0];
      for (int64_t p = c0 + 1; p < c1; ++p) s = fma(Ap[i0 + r + p * m], b[p], s);
      acc[r] = s;
    }
  }
}

/*
 * Shared body of the BLAS-form recurrence: k cut into chunks of kc (kc >= k:
 * one chain), chunk results summed in chunk order, then the epilogue.
 */
static void gemm_recurrence_impl(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                                 double* C, int64_t ldc, int64_t kc, int nthreads) {
  if (alpha == 0.0 || k == 0) {
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < m; ++i) C[i + j * ldc] = (beta == 0.0) ? 0.0 : beta * C[i + j * ldc];
    return;
  }
  if (kc < 1 || kc > k) kc = k;
  double* Ap = pack_op(ta, A, lda, m, k, nthreads);
  double* Bp = pack_op((tb == 'N' || tb == 'n') ? 'N' : 'T', B, ldb, k, n, nthreads);
  if (!Ap || !Bp) {
    free(Ap);
    free(Bp);
    return;
  }
  (void)nthreads;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i0 = 0; i0 < m; i0 += 8) {
      const int nr = (int)(m - i0 < 8 ? m - i0 : 8);
      double s[8] = {0, 0, 0, 0, 0, 0, 0, 0}, acc[8];
      for (int64_t c0 = 0; c0 < k; c0 += kc) {
        const int64_t c1 = c0 + kc < k ? c0 + kc : k;
        chain_block(Ap, Bp, m, i0, nr, j, k, c0, c1, acc);
        for (int r = 0; r < nr; ++r) s[r] = (c0 == 0) ? acc[r] : s[r] + acc[r];
      }
      for (int r = 0; r < nr; ++r) {
        double c = alpha * s[r];
        if (beta != 0.0) c = fma(beta, C[i0 + r + j * ldc], c);
        C[i0 + r + j * ldc] = c;
      }
    }
  free(Ap);
  free(Bp);
}

ORC_API void orc_gemm_recurrence(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                                 const double* A, int64_t lda, const double* B, int64_t ldb,
                                 double beta, double* C, int64_t ldc, int nthreads) {
  gemm_recurrence_impl(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, k, nthreads);
}

/* The same recurrence on sampled entries (ii[s], jj[s]); c0[s] is the old C entry. */
ORC_API void orc_gemm_recurrence_sample(char ta, char tb, int64_t k, double alpha, const double* A,
                                        int64_t lda, const double* B, int64_t ldb, double beta,
                                        const double* c0, int64_t ns, const int64_t* ii,
                                        const int64_t* jj, double* out, int nthreads) {
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t s = 0; s < ns; ++s) {
    const int64_t i = ii[s], j = jj[s];
    double c;
    if (alpha == 0.0 || k == 0) {
      c = (beta == 0.0) ? 0.0 : beta * c0[s];
    } else {
      double acc = opA(ta, A, lda, i, 0) * opB(tb, B, ldb, 0, j);
      for (int64_t p = 1; p < k; ++p) acc = fma(opA(ta, A, lda, i, p), opB(tb, B, ldb, p, j), acc);
      c = alpha * acc;
      if (beta != 0.0) c = fma(beta, c0[s], c);
    }
    out[s] = c;
  }
}

/*
 * Split-K form of the recurrence (the grouped launch's deterministic split-K):
 * k is cut into chunks [q*kc, min(k, (q+1)*kc)); each chunk is its own chain
 * (first product, then fma in increasing k), and the chunk results are summed
 * in chunk order, s = p_0; s = s + p_q.  kc >= k gives the plain recurrence.
 */
ORC_API void orc_gemm_recurrence_chunked(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                                         const double* A, int64_t lda, const double* B, int64_t ldb,
                                         double beta, double* C, int64_t ldc, int64_t kc, int nthreads) {
  gemm_recurrence_impl(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kc, nthreads);
}

/* ------------------------------------------------------------------------ */
/* 4. Dot2 FP64 oracle                                                       */
/* ------------------------------------------------------------------------ */
static inline void two_sum(double a, double b, double* s, double* e) {
  double x = a + b, z = x - a;
  *e = (a - (x - z)) + (b - z);
  *s = x;
}
static inline void two_prod(double a, double b, double* p, double* e) {
  double x = a * b;
  *e = fma(a, b, -x);
  *p = x;
}

/* Compensated alpha*(hi+lo) + beta*c0, one final rounding of a double-double. */
static inline double dd_epilogue(double hi, double lo, double alpha, double beta, double c0) {
  double p1, e1, p2 = 0.0, e2 = 0.0, s, es;
  two_prod(alpha, hi, &p1, &e1);
  e1 += alpha * lo;
  if (beta != 0.0) two_prod(beta, c0, &p2, &e2);
  two_sum(p1, p2, &s, &es);
  return s + (es + e1 + e2);
}

static inline void dot2_entry(char ta, char tb, int64_t k, const double* A, int64_t lda,
                              const double* B, int64_t ldb, int64_t i, int64_t j, double* hi,
                              double* lo) {
  double s = 0.0, c = 0.0;
  for (int64_t p = 0; p < k; ++p) {
    double pr, ep, es;
    two_prod(opA(ta, A, lda, i, p), opB(tb, B, ldb, p, j), &pr, &ep);
    two_sum(s, pr, &s, &es);
    c += ep + es;
  }
  two_sum(s, c, hi, lo);
}

ORC_API void orc_gemm_dot2(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                           const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                           double* C, int64_t ldc, int nthreads) {
  (void)nthreads;
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      double hi = 0.0, lo = 0.0;
      if (alpha != 0.0) dot2_entry(ta, tb, k, A, lda, B, ldb, i, j, &hi, &lo);
      C[i + j * ldc] = dd_epilogue(hi, lo, alpha, beta, beta != 0.0 ? C[i + j * ldc] : 0.0);
    }
}

/* Dot2 on sampled entries (ii[s], jj[s]); c0[s] is the old C entry. */
ORC_API void orc_gemm_dot2_sample(char ta, char tb, int64_t k, double alpha, const double* A,
                                  int64_t lda, const double* B, int64_t ldb, double beta,
                                  const double* c0, int64_t ns, const int64_t* ii,
                                  const int64_t* jj, double* out, int nthreads) {
  (void)nthreads;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t s = 0; s < ns; ++s) {
    double hi = 0.0, lo = 0.0;
    if (alpha != 0.0) dot2_entry(ta, tb, k, A, lda, B, ldb, ii[s], jj[s], &hi, &lo);
    out[s] = dd_epilogue(hi, lo, alpha, beta, beta != 0.0 ? c0[s] : 0.0);
  }
}

/* Float inputs promoted to double (FP64 oracle for the SGEMM paths). */
ORC_API void orc_sgemm_dot2(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                            const float* A, int64_t lda, const float* B, int64_t ldb, double beta,
                            const float* C0, int64_t ldc, double* out, int64_t ldo, int nthreads) {
  (void)nthreads;
  const int an = (ta == 'N' || ta == 'n'), bn = (tb == 'N' || tb == 'n');
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      double s = 0.0, c = 0.0, hi = 0.0, lo = 0.0;
      if (alpha != 0.0) {
        for (int64_t p = 0; p < k; ++p) {
          double a = an ? A[i + p * lda] : A[p + i * lda];
          double b = bn ? B[p + j * ldb] : B[j + p * ldb];
          double pr, ep, es;
          two_prod(a, b, &pr, &ep);
          two_sum(s, pr, &s, &es);
          c += ep + es;
        }
        two_sum(s, c, &hi, &lo);
      }
      out[i + j * ldo] = dd_epilogue(hi, lo, alpha, beta, beta != 0.0 ? (double)C0[i + j * ldc] : 0.0);
    }
}

/* kernel! (src/kernels.jl:212-241, kernel_quote :225-233): D += A*X with each
 * element's chain starting from D -- acc = D[i,j]; acc = fma(A[i,n], X[n,j], acc)
 * for n = 0..N-1 in increasing n; leading dimensions lda, ldx, ldd. */
ORC_API void orc_kernel_acc_d(int64_t M, int64_t N, int64_t P, const double* A, int64_t lda, const double* X,
                              int64_t ldx, double* D, int64_t ldd) {
  for (int64_t j = 0; j < P; ++j)
    for (int64_t i = 0; i < M; ++i) {
      double acc = D[i + j * ldd];
      for (int64_t n = 0; n < N; ++n) acc = fma(A[i + n * lda], X[n + j * ldx], acc);
      D[i + j * ldd] = acc;
    }
}

/* ------------------------------------------------------------------------ */
/* 5. Split-BF16 emulated DGEMM, exact restatement                           */
/* ------------------------------------------------------------------------ */
/*
 * Scheme (this build's definition; the reference has no emulated path):
 *   E_i = exponent of max_p |op(A)[i,p]| (frexp: max = f*2^E, f in [0.5,1)),
 *         0 for an all-zero row; F_j likewise for the columns of op(B).
 *   x = ldexp(a, -E_i) in (-1, 1);  r = x*2^8;  d_1 = rint(r); r -= d_1;
 *   for t >= 2: r *= 2^9; d_t = rint(r); r -= d_t.      (|d_t| <= 256)
 *   so  a ~= 2^E_i * sum_t d_t * 2^(-8-9(t-1)).  Every integer of magnitude
 *   <= 256 has at most 8 significant bits and is exact in BF16, so a digit
 *   carries 9 bits (balanced, with sign).
 *   Level sums over k-chunks of their own length kc_l:
 *     S_l = sum_{t+u=l, t,u<=s} sum_{p in chunk} dA_t[i,p]*dB_u[p,j]   (exact integers),
 *   for l = 2..s+1 (triangular truncation).  Level l has l-1 pairs of products
 *   |d_t d_u| <= 2^16 and kc_l = largest power of two with kc_l * (l-1) <= 256
 *   (256, 128, 64, 64, 32, ...), so that every level sum of a chunk stays
 *   within 2^24 and is exact in FP32 whatever the summation order.
 *   At every multiple of the shortest chunk and at k, in increasing l, each
 *   level whose chunk ends there is flushed:  g = (l-2)/3, j = (l-2)%3,
 *     G_g = fma((double)S_l, 2^(-9j), G_g), S_l = 0     (groups of three levels:
 *   G_g is a multiple of 2^-18 and, when k*s <= 2^18, below 2^35, so that every
 *   flush is then exact; for s > 6 the groups hold four levels, g = (l-2)/4,
 *   j = (l-2)%4, and w_g below is ldexp(G_g, -36g)).
 *   v = compensated sum over g from the smallest: w_g = ldexp(G_g, -27g);
 *       (hi, e) = TwoSum(hi, w_g), lo += e;  v = hi + lo;
 *   acc = ldexp(v, E_i + F_j - 16);  c = alpha*acc;  c = fma(beta, c_old, c) if beta != 0.
 * Rows/columns containing a non-finite value make the routine return 1.
 */
/* kc_l of the 0-based level L = l - 2 (L + 1 pairs). */
ORC_API int orc_ozaki_level_chunk(int L) {
  int kc = 256;
  while (kc * (L + 1) > 256) kc >>= 1;
  return kc;
}
static int row_exponent(const double* v, int64_t n, int64_t stride, int* e) {
  double mx = 0.0;
  for (int64_t p = 0; p < n; ++p) {
    double a = fabs(v[p * stride]);
    if (!isfinite(a)) return 1;
    if (a > mx) mx = a;
  }
  int ex = 0;
  if (mx > 0.0) frexp(mx, &ex);
  *e = ex;
  return 0;
}

static void split_digits(double a, int e, int s, int16_t* d, int64_t dstride) {
  double r = ldexp(a, -e) * 256.0;
  for (int t = 0; t < s; ++t) {
    if (t > 0) r *= 512.0;
    double q = rint(r);
    r -= q;
    d[t * dstride] = (int16_t)q;
  }
}

/* Digits of op(A) rows: dA[t][i][p] (t-major, then row i, then p), exponents eA[i]. */
ORC_API int orc_ozaki_split_a(char ta, int64_t m, int64_t k, const double* A, int64_t lda, int s,
                              int* eA, int16_t* dA) {
  const int an = (ta == 'N' || ta == 'n');
  for (int64_t i = 0; i < m; ++i) {
    int e;
    if (row_exponent(an ? A + i : A + i * lda, k, an ? lda : 1, &e)) return 1;
    eA[i] = e;
    for (int64_t p = 0; p < k; ++p)
      split_digits(an ? A[i + p * lda] : A[p + i * lda], e, s, dA + i * k + p, m * k);
  }
  return 0;
}
/* Digits of op(B) columns: dB[t][j][p], exponents eB[j]. */
ORC_API int orc_ozaki_split_b(char tb, int64_t k, int64_t n, const double* B, int64_t ldb, int s,
                              int* eB, int16_t* dB) {
  const int bn = (tb == 'N' || tb == 'n');
  for (int64_t j = 0; j < n; ++j) {
    int e;
    if (row_exponent(bn ? B + j * ldb : B + j, k, bn ? 1 : ldb, &e)) return 1;
    eB[j] = e;
    for (int64_t p = 0; p < k; ++p)
      split_digits(bn ? B[p + j * ldb] : B[j + p * ldb], e, s, dB + j * k + p, n * k);
  }
  return 0;
}

/* kcl[L]: chunk length of 0-based level L, each a multiple of the shortest.
 * db: bits per digit (level weight 2^-db), gl: levels per FP64 group,
 * sh: exponent of the top digit pair (2 * bits of the first digit). */
static double ozaki_entry(int64_t m, int64_t n, int64_t k, int s, const int64_t* kcl, const int* eA, const int16_t* dA,
                          const int* eB, const int16_t* dB, int64_t i, int64_t j, int db, int gl, int sh) {
  const int ngroups = (s + gl - 1) / gl; /* levels 2..s+1 -> groups of gl */
  double G[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int64_t S[20] = {0};
  int64_t kmin = kcl[0];
  for (int L = 1; L < s; ++L)
    if (kcl[L] < kmin) kmin = kcl[L];
  for (int64_t c0 = 0; c0 < k; c0 += kmin) {
    const int64_t c1 = c0 + kmin < k ? c0 + kmin : k;
    for (int t = 1; t <= s; ++t)
      for (int u = 1; u + t <= s + 1; ++u) {
        const int16_t* a = dA + (int64_t)(t - 1) * m * k + i * k;
        const int16_t* b = dB + (int64_t)(u - 1) * n * k + j * k;
        int64_t acc = 0;
        for (int64_t p = c0; p < c1; ++p) acc += (int64_t)a[p] * (int64_t)b[p];
        S[t + u] += acc;
      }
    for (int l = 2; l <= s + 1; ++l) {
      if (c1 % kcl[l - 2] != 0 && c1 != k) continue;
      const int g = (l - 2) / gl, jj = (l - 2) % gl;
      G[g] = fma((double)S[l], ldexp(1.0, -db * jj), G[g]);
      S[l] = 0;
    }
  }
  double hi = 0.0, lo = 0.0;
  for (int g = ngroups - 1; g >= 0; --g) {
    double w = ldexp(G[g], -db * gl * g), e;
    two_sum(hi, w, &hi, &e);
    lo += e;
  }
  return ldexp(hi + lo, eA[i] + eB[j] - sh);
}

ORC_API int orc_dgemm_ozaki(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                            const double* A, int64_t lda, const double* B, int64_t ldb,
                            double beta, double* C, int64_t ldc, int s, int nthreads) {
  if (s < 1 || s > 12) return -16;
  int64_t kcl[12];
  for (int L = 0; L < s; ++L) kcl[L] = orc_ozaki_level_chunk(L);
  (void)nthreads;
  if (alpha == 0.0 || k == 0) {
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < m; ++i) C[i + j * ldc] = beta == 0.0 ? 0.0 : beta * C[i + j * ldc];
    return 0;
  }
  int* eA = (int*)malloc(sizeof(int) * (m + 1));
  int* eB = (int*)malloc(sizeof(int) * (n + 1));
  int16_t* dA = (int16_t*)malloc(sizeof(int16_t) * (size_t)(s * m * k + 1));
  int16_t* dB = (int16_t*)malloc(sizeof(int16_t) * (size_t)(s * n * k + 1));
  int rc = 0;
  if (!eA || !eB || !dA || !dB) { rc = 2; goto done; }
  if (orc_ozaki_split_a(ta, m, k, A, lda, s, eA, dA) || orc_ozaki_split_b(tb, k, n, B, ldb, s, eB, dB)) {
    rc = 1;
    goto done;
  }
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      double acc = ozaki_entry(m, n, k, s, kcl, eA, dA, eB, dB, i, j, 9, s > 6 ? 4 : 3, 16);
      double c = alpha * acc;
      if (beta != 0.0) c = fma(beta, C[i + j * ldc], c);
      C[i + j * ldc] = c;
    }
done:
  free(eA); free(eB); free(dA); free(dB);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* 5b. Split-INT8 emulated DGEMM, exact restatement                          */
/* ------------------------------------------------------------------------ */
/*
 * The same scheme with every digit an int8 (v_mfma_i32_32x32x32_i8: exact
 * int32 sums), as this build defines it:
 *   E_i = frexp exponent of max_p |op(A)[i,p]|, plus one if that max scaled
 *         by 2^-E_i exceeds 126.5/128 (so that the top digit fits in int8);
 *   X = rint(ldexp(a, -E_i) * 2^(8s-1))   (an integer, |X| < 2^(8s-1));
 *   the s digits are X's signed base-256 digits, least significant first:
 *     d_s = (int8)(X & 255), X = (X - d_s)/256, ..., d_1 = X   (all in [-128, 127]);
 *   so a ~= 2^E_i * sum_t d_t * 2^(-7-8(t-1)).
 * Level sums as above, over chunks of kc = 8192 (|S_l| <= s * kc * 2^14 <= 2^30:
 * exact in int32), flushed in the same order into FP64 groups of four levels,
 * G_g = fma((double)S_l, 2^(-8j), G_g) with g = (l-2)/4, j = (l-2)%4; the
 * combine is the BF16 one with w_g = ldexp(G_g, -32g), then
 * acc = ldexp(v, E_i + F_j - 14).
 */
#define ORC_I8_CHUNK 8192
ORC_API int orc_ozaki_i8_chunk(void) { return ORC_I8_CHUNK; }

static int row_exponent_i8(const double* v, int64_t n, int64_t stride, int* e) {
  double mx = 0.0;
  for (int64_t p = 0; p < n; ++p) {
    double a = fabs(v[p * stride]);
    if (!isfinite(a)) return 1;
    if (a > mx) mx = a;
  }
  int ex = 0;
  if (mx > 0.0) {
    double f = frexp(mx, &ex);
    if (f > 126.5 / 128.0) ex += 1;
  }
  *e = ex;
  return 0;
}

static void split_digits_i8(double a, int e, int s, int16_t* d, int64_t dstride) {
  int64_t X = (int64_t)rint(ldexp(a, 8 * s - 1 - e));
  for (int t = s - 1; t >= 1; --t) {
    const int64_t lo = (int64_t)(int8_t)(uint8_t)(X & 255);
    d[t * dstride] = (int16_t)lo;
    X = (X - lo) / 256;
  }
  d[0] = (int16_t)X;
}

ORC_API int orc_ozaki_i8_split_a(char ta, int64_t m, int64_t k, const double* A, int64_t lda, int s,
                                 int* eA, int16_t* dA) {
  const int an = (ta == 'N' || ta == 'n');
  for (int64_t i = 0; i < m; ++i) {
    int e;
    if (row_exponent_i8(an ? A + i : A + i * lda, k, an ? lda : 1, &e)) return 1;
    eA[i] = e;
    for (int64_t p = 0; p < k; ++p)
      split_digits_i8(an ? A[i + p * lda] : A[p + i * lda], e, s, dA + i * k + p, m * k);
  }
  return 0;
}
ORC_API int orc_ozaki_i8_split_b(char tb, int64_t k, int64_t n, const double* B, int64_t ldb, int s,
                                 int* eB, int16_t* dB) {
  const int bn = (tb == 'N' || tb == 'n');
  for (int64_t j = 0; j < n; ++j) {
    int e;
    if (row_exponent_i8(bn ? B + j * ldb : B + j, k, bn ? 1 : ldb, &e)) return 1;
    eB[j] = e;
    for (int64_t p = 0; p < k; ++p)
      split_digits_i8(bn ? B[p + j * ldb] : B[j + p * ldb], e, s, dB + j * k + p, n * k);
  }
  return 0;
}

ORC_API int orc_dgemm_ozaki_i8(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                               const double* A, int64_t lda, const double* B, int64_t ldb,
                               double beta, double* C, int64_t ldc, int s, int nthreads) {
  if (s < 1 || s > 8) return -16;
  int64_t kcl[8];
  for (int L = 0; L < s; ++L) kcl[L] = ORC_I8_CHUNK;
  if (alpha == 0.0 || k == 0) {
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < m; ++i) C[i + j * ldc] = beta == 0.0 ? 0.0 : beta * C[i + j * ldc];
    return 0;
  }
  int* eA = (int*)malloc(sizeof(int) * (m + 1));
  int* eB = (int*)malloc(sizeof(int) * (n + 1));
  int16_t* dA = (int16_t*)malloc(sizeof(int16_t) * (size_t)(s * m * k + 1));
  int16_t* dB = (int16_t*)malloc(sizeof(int16_t) * (size_t)(s * n * k + 1));
  int rc = 0;
  if (!eA || !eB || !dA || !dB) { rc = 2; goto done; }
  if (orc_ozaki_i8_split_a(ta, m, k, A, lda, s, eA, dA) || orc_ozaki_i8_split_b(tb, k, n, B, ldb, s, eB, dB)) {
    rc = 1;
    goto done;
  }
#pragma omp parallel for schedule(dynamic) num_threads(nthreads > 0 ? nthreads : 1)
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) {
      double acc = ozaki_entry(m, n, k, s, kcl, eA, dA, eB, dB, i, j, 8, 4, 14);
      double c = alpha * acc;
      if (beta != 0.0) c = fma(beta, C[i + j * ldc], c);
      C[i + j * ldc] = c;
    }
done:
  free(eA); free(eB); free(dA); free(dB);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* 6. Tile policy restatements                                               */
/* ------------------------------------------------------------------------ */
/* src/kernel_structure.jl:76-99 -> out = (elements_per_register, rows, cols). */
ORC_API void orc_pick_kernel_size(int64_t register_size, int64_t register_count, int64_t t_size,
                                  int64_t* out) {
  const int64_t epr = register_size / t_size;
  const int64_t max_total = epr * register_count;
  const int64_t ncl = (max_total + epr - 1) / epr; /* cld */
  int64_t prow = 0, pcol = 0;
  double prev = -INFINITY;
  for (int64_t a = 1; a <= ncl; ++a) {
    int64_t rows = a * epr;
    int64_t cols = (register_count - a - 1) / a; /* Julia div truncates toward zero */
    double ratio = (double)(rows * cols) / (double)(cols + a);
    if (ratio < prev) break;
    prev = ratio;
    prow = rows;
    pcol = cols;
  }
  out[0] = epr; out[1] = prow; out[2] = pcol;
}

/* round(Int, x/y)*y with Julia's default RoundNearest (ties to even); src/miscellaneous.jl:12 */
static int64_t round_to_nearest(int64_t x, int64_t y) { return (int64_t)rint((double)x / (double)y) * y; }

/* src/memory_management.jl:157-177 */
static void divide_into_rough_square(int64_t L, int64_t M, int64_t P, int64_t n, int64_t mbase,
                                     int64_t pbase, int64_t* mo, int64_t* po) {
  int64_t Mhalf = round_to_nearest(M >> 1, mbase);
  if (Mhalf < mbase) Mhalf = mbase;
  int64_t Phalf = round_to_nearest(P >> 1, pbase);
  if (Phalf < pbase) Phalf = pbase;
  int64_t bsize = Mhalf * Phalf + Mhalf * n + n * Phalf;
  if (bsize < L) { *mo = Mhalf; *po = Phalf; return; }
  int64_t ub = (int64_t)floor(sqrt((double)(n * n + L)) - (double)n);
  int64_t m2 = round_to_nearest(ub, mbase);
  if (m2 < mbase) m2 = mbase;
  int64_t p2;
  if (m2 >= M) {
    m2 = M;
    p2 = (L - m2 * n) / (m2 + n);
    if (p2 > P) p2 = P;
  } else {
    p2 = ub / pbase * pbase;
    if (p2 >= P) {
      p2 = P;
      int64_t t = (L - p2 * n) / (p2 + n);
      m2 = t < M ? t : M;
    }
  }
  *mo = m2; *po = p2;
}

/*
 * src/memory_management.jl:78-140 with D_count = A_count = X_count = 1.
 * out[0..8] = (m1,n1,p1, m2,n2,p2, m3,n3,p3); returns the level (0..3).
 */
ORC_API int orc_blocking_structure(int64_t M, int64_t N, int64_t P, int64_t t_size,
                                   const int64_t* cache_size, int64_t register_size,
                                   int64_t register_count, int64_t* out) {
  const int64_t total = M * N + N * P + M * P;
  const int64_t L1 = cache_size[0] / t_size, L2 = cache_size[1] / t_size, L3 = cache_size[2] / t_size;
  int64_t ks[3];
  orc_pick_kernel_size(register_size, register_count, t_size, ks);
  const int64_t m1 = ks[1], p1 = ks[2];
  if (L1 > total) {
    int64_t v[9] = {m1 < M ? m1 : M, N, p1 < P ? p1 : P, M, N, P, M, N, P};
    memcpy(out, v, sizeof v);
    return 0;
  }
  int64_t n1 = (L1 - m1 * p1) / (m1 + p1), n2;
  if (n1 > N) n2 = n1 = N; else n2 = n1;
  if (L2 > total) {
    int64_t v[9] = {m1, n1, p1, M, N, P, M, N, P};
    memcpy(out, v, sizeof v);
    return 1;
  }
  int64_t m2, p2;
  divide_into_rough_square(L2, M, P, n2, m1, p1, &m2, &p2);
  if (L3 > total) {
    int64_t v[9] = {m1, n1, p1, m2, n2, p2, M, N, P};
    memcpy(out, v, sizeof v);
    return 2;
  }
  int64_t n3 = (L3 - m2 * p2) / (m2 + p2), m3, p3;
  if (n3 > N) {
    n3 = N;
    divide_into_rough_square(L3, M, P, N, m2, p2, &m3, &p3);
  } else {
    m3 = m2; p3 = p2;
  }
  int64_t v[9] = {m1, n1, p1, m2, n2, p2, m3, n3, p3};
  memcpy(out, v, sizeof v);
  return 3;
}

/* Threads an OpenMP parallel region with num_threads(nthreads) actually runs
 * (the CPU-baseline leg reports this count). */
ORC_API int orc_threads_used(int nthreads) {
  int used = 1;
  (void)nthreads;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
  {
#ifdef _OPENMP
#pragma omp single
    used = omp_get_num_threads();
#endif
  }
  return used;
}
