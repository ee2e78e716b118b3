/*
 * asan_check.c -- runs every entry point of the CPU checker (oracle.c) under
 * AddressSanitizer and UndefinedBehaviorSanitizer on small operands with
 * exact-size heap buffers (so any read or write past an operand is caught),
 * including fringe shapes, both transposes, strided leading dimensions and
 * k-chunk boundaries, and cross-checks a few results against plain loops.
 *
 * TEST INFRASTRUCTURE ONLY: built by `make -C oracle asan` and run by
 * tests/test_oracle.py; nothing in the product links it.  Exit status 0 means
 * every check passed and the sanitizers reported nothing.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

uint64_t orc_splitmix64(uint64_t seed, uint64_t i);
void orc_fill_uniform(uint64_t seed, int64_t n, double* out);
void orc_fill_wide(uint64_t seed, int64_t n, int emin, int emax, double* out);
int orc_jmul_d(int64_t M, int64_t N, int64_t P, const double* A, int64_t lda, const double* X, int64_t ldx, double* D,
               int64_t ldd, int avx512, int nthreads);
int orc_jmul_s(int64_t M, int64_t N, int64_t P, const float* A, int64_t lda, const float* X, int64_t ldx, float* D,
               int64_t ldd, int avx512, int nthreads);
void orc_jmul_naive_d(int64_t M, int64_t N, int64_t P, const double* A, int64_t lda, const double* X, int64_t ldx,
                      double* D, int64_t ldd);
void orc_gemm_recurrence(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                         int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                         int nthreads);
void orc_gemm_recurrence_sample(char ta, char tb, int64_t k, double alpha, const double* A, int64_t lda,
                                const double* B, int64_t ldb, double beta, const double* c0, int64_t ns,
                                const int64_t* ii, const int64_t* jj, double* out, int nthreads);
void orc_gemm_recurrence_chunked(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                                 int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                                 int64_t kc, int nthreads);
void orc_gemm_dot2(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                   const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int nthreads);
void orc_gemm_dot2_sample(char ta, char tb, int64_t k, double alpha, const double* A, int64_t lda,
                          const double* B, int64_t ldb, double beta, const double* c0, int64_t ns,
                          const int64_t* ii, const int64_t* jj, double* out, int nthreads);
void orc_sgemm_dot2(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const float* A, int64_t lda,
                    const float* B, int64_t ldb, double beta, const float* C0, int64_t ldc, double* out,
                    int64_t ldo, int nthreads);
void orc_kernel_acc_d(int64_t M, int64_t N, int64_t P, const double* A, int64_t lda, const double* X, int64_t ldx,
                      double* D, int64_t ldd);
int orc_ozaki_level_chunk(int L);
int orc_ozaki_split_a(char ta, int64_t m, int64_t k, const double* A, int64_t lda, int s, int* eA, int16_t* dA);
int orc_ozaki_split_b(char tb, int64_t k, int64_t n, const double* B, int64_t ldb, int s, int* eB, int16_t* dB);
int orc_dgemm_ozaki(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                    const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int s, int nthreads);
int orc_ozaki_i8_chunk(void);
int orc_ozaki_i8_split_a(char ta, int64_t m, int64_t k, const double* A, int64_t lda, int s, int* eA, int16_t* dA);
int orc_ozaki_i8_split_b(char tb, int64_t k, int64_t n, const double* B, int64_t ldb, int s, int* eB, int16_t* dB);
int orc_dgemm_ozaki_i8(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                       int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int s,
                       int nthreads);
void orc_pick_kernel_size(int64_t register_size, int64_t register_count, int64_t t_size, int64_t* out);
int orc_blocking_structure(int64_t M, int64_t N, int64_t P, int64_t t_size, const int64_t* cache_size,
                           int64_t register_size, int64_t register_count, int64_t* out);
int orc_threads_used(int nthreads);

static int failures = 0;
#define CHECK(cond, ...)                     \
  do {                                       \
    if (!(cond)) {                           \
      fprintf(stderr, "FAIL: " __VA_ARGS__); \
      fprintf(stderr, "\n");                 \
      ++failures;                            \
    }                                        \
  } while (0)

/* exact-size heap buffers: one element past the end is outside the allocation */
static double* dmat(int64_t rows, int64_t cols, int64_t ld, uint64_t seed) {
  const int64_t n = rows == 0 || cols == 0 ? 1 : ld * (cols - 1) + rows;
  double* p = malloc((size_t)n * sizeof(double));
  orc_fill_uniform(seed, n, p);
  return p;
}
static float* smat(int64_t rows, int64_t cols, int64_t ld, uint64_t seed) {
  const int64_t n = rows == 0 || cols == 0 ? 1 : ld * (cols - 1) + rows;
  double* t = malloc((size_t)n * sizeof(double));
  orc_fill_uniform(seed, n, t);
  float* p = malloc((size_t)n * sizeof(float));
  for (int64_t i = 0; i < n; ++i) p[i] = (float)t[i];
  free(t);
  return p;
}

static void check_jmul(int64_t M, int64_t N, int64_t P) {
  double* A = dmat(M, N, M, 1);
  double* X = dmat(N, P, N, 2);
  double* D = dmat(M, P, M, 3);
  double* R = dmat(M, P, M, 4);
  for (int avx = 0; avx < 2; ++avx) {
    CHECK(orc_jmul_d(M, N, P, A, M, X, N, D, M, avx, 2) == 0, "jmul_d rc");
    orc_jmul_naive_d(M, N, P, A, M, X, N, R, M);
    CHECK(memcmp(D, R, (size_t)(M * P) * sizeof(double)) == 0, "jmul_d != naive at %ldx%ldx%ld avx512=%d",
          (long)M, (long)N, (long)P, avx);
  }
  float* As = smat(M, N, M, 5);
  float* Xs = smat(N, P, N, 6);
  float* Ds = smat(M, P, M, 7);
  for (int avx = 0; avx < 2; ++avx) CHECK(orc_jmul_s(M, N, P, As, M, Xs, N, Ds, M, avx, 2) == 0, "jmul_s rc");
  free(A), free(X), free(D), free(R), free(As), free(Xs), free(Ds);
}

static void check_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k) {
  const int64_t ar = ta == 'N' ? m : k, ac = ta == 'N' ? k : m, br = tb == 'N' ? k : n, bc = tb == 'N' ? n : k;
  const int64_t lda = ar + 3, ldb = br + 1, ldc = m + 2;
  double* A = dmat(ar, ac, lda, 11);
  double* B = dmat(br, bc, ldb, 12);
  double* C = dmat(m, n, ldc, 13);
  double* C2 = dmat(m, n, ldc, 13);
  double* C3 = dmat(m, n, ldc, 13);
  orc_gemm_recurrence(ta, tb, m, n, k, 1.5, A, lda, B, ldb, 0.5, C, ldc, 2);
  orc_gemm_recurrence_chunked(ta, tb, m, n, k, 1.5, A, lda, B, ldb, 0.5, C2, ldc, k + 1, 2);
  int same = 1;
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) same &= C[i + j * ldc] == C2[i + j * ldc];
  CHECK(same, "chunked(kc > k) != recurrence %c%c", ta, tb);
  orc_gemm_recurrence_chunked(ta, tb, m, n, k, 1.5, A, lda, B, ldb, 0.5, C3, ldc, 7, 2);
  double* D2 = dmat(m, n, ldc, 13);
  orc_gemm_dot2(ta, tb, m, n, k, 1.5, A, lda, B, ldb, 0.5, D2, ldc, 2);
  double err = 0;
  for (int64_t j = 0; j < n; ++j)
    for (int64_t i = 0; i < m; ++i) err = fmax(err, fabs(C[i + j * ldc] - D2[i + j * ldc]));
  CHECK(err <= 2.0 * k * ldexp(1.0, -53) * (1.5 * k + 1.0), "recurrence vs Dot2 %g", err);
  int64_t ii[5], jj[5];
  double c0[5], out[5], out2[5];
  for (int s = 0; s < 5; ++s) {
    ii[s] = (s * 7) % m;
    jj[s] = (s * 5) % n;
    c0[s] = 0.25 * s;
  }
  orc_gemm_recurrence_sample(ta, tb, k, 1.5, A, lda, B, ldb, 0.5, c0, 5, ii, jj, out, 2);
  orc_gemm_dot2_sample(ta, tb, k, 1.5, A, lda, B, ldb, 0.5, c0, 5, ii, jj, out2, 2);
  for (int s = 0; s < 5; ++s) CHECK(fabs(out[s] - out2[s]) <= 1e-12 * k, "sample mismatch");
  for (int sl = 1; sl <= 8; ++sl) {
    double* E = dmat(m, n, ldc, 13);
    CHECK(orc_dgemm_ozaki(ta, tb, m, n, k, 1.5, A, lda, B, ldb, 0.5, E, ldc, sl, 2) == 0, "ozaki rc");
    CHECK(orc_dgemm_ozaki_i8(ta, tb, m, n, k, 1.5, A, lda, B, ldb, 0.5, E, ldc, sl, 2) == 0, "ozaki_i8 rc");
    free(E);
  }
  int* eA = malloc((size_t)m * sizeof(int));
  int* eB = malloc((size_t)n * sizeof(int));
  int16_t* dA = malloc((size_t)(8 * m * k) * sizeof(int16_t));
  int16_t* dB = malloc((size_t)(8 * n * k) * sizeof(int16_t));
  CHECK(orc_ozaki_split_a(ta, m, k, A, lda, 8, eA, dA) == 0, "split_a");
  CHECK(orc_ozaki_split_b(tb, k, n, B, ldb, 8, eB, dB) == 0, "split_b");
  CHECK(orc_ozaki_i8_split_a(ta, m, k, A, lda, 8, eA, dA) == 0, "i8 split_a");
  CHECK(orc_ozaki_i8_split_b(tb, k, n, B, ldb, 8, eB, dB) == 0, "i8 split_b");
  float* As = smat(ar, ac, lda, 14);
  float* Bs = smat(br, bc, ldb, 15);
  float* Cs = smat(m, n, ldc, 16);
  double* O = malloc((size_t)(m * n) * sizeof(double));
  orc_sgemm_dot2(ta, tb, m, n, k, 1.5, As, lda, Bs, ldb, 0.5, Cs, ldc, O, m, 2);
  free(A), free(B), free(C), free(C2), free(C3), free(D2), free(eA), free(eB), free(dA), free(dB);
  free(As), free(Bs), free(Cs), free(O);
}

int main(void) {
  CHECK(orc_splitmix64(0, 0) == 0xe220a8397b1dcdafull, "splitmix64 first output for seed 0");
  double w[37];
  orc_fill_wide(9, 37, -12, 12, w);
  for (int i = 0; i < 37; ++i) CHECK(fabs(w[i]) < ldexp(1.0, 12), "wide range");
  const int64_t shapes[][3] = {{1, 1, 1}, {16, 32, 14}, {13, 12, 12}, {41, 7, 6}, {80, 3, 11}, {24, 18, 18}, {3, 100, 5}};
  for (size_t s = 0; s < sizeof shapes / sizeof shapes[0]; ++s) check_jmul(shapes[s][0], shapes[s][1], shapes[s][2]);
  const char ops[] = "NT";
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      check_gemm(ops[a], ops[b], 17, 9, 300);
      check_gemm(ops[a], ops[b], 5, 23, 1);
    }
  /* kernel!: chains from D, strided sub-blocks */
  {
    double* A = dmat(9, 6, 11, 21);
    double* X = dmat(6, 4, 8, 22);
    double* D = dmat(9, 4, 10, 23);
    orc_kernel_acc_d(9, 6, 4, A, 11, X, 8, D, 10);
    free(A), free(X), free(D);
  }
  for (int L = 0; L < 12; ++L) CHECK(orc_ozaki_level_chunk(L) * (L + 1) <= 256, "level chunk %d", L);
  CHECK(orc_ozaki_i8_chunk() > 0, "i8 chunk");
  int64_t ks[3];
  orc_pick_kernel_size(64, 32, 8, ks);
  CHECK(ks[0] == 8 && ks[1] == 40 && ks[2] == 5, "pick_kernel_size AVX-512 F64");
  int64_t out[9];
  const int64_t cache[3] = {48 << 10, 2 << 20, 105 << 20};
  CHECK(orc_blocking_structure(256, 256, 256, 8, cache, 64, 32, out) >= 0, "blocking_structure");
  CHECK(orc_threads_used(2) >= 1, "threads");
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("asan_check: all oracle entry points ran clean\n");
  return 0;
}
