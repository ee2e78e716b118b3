/*
 * jblas_hip.h -- C ABI of libjblas_hip.so, the MI355X (gfx950) GEMM engine that
 * sits behind jBLAS.jl's GEMM entry point.
 *
 * Reference interface replaced (paths relative to the jBLAS.jl repository):
 *   jmul!(D::MMatrix{M,P,T}, A::MMatrix{M,N,T}, X::MMatrix{N,P,T}, ...)  src/gemm.jl:244-348
 *       "calculates D = A * X"                                             src/gemm.jl:225-243
 *   kernel!(pD, pA, pX, ::Kernel{Mk,Pk,stride_AD,stride_X,N})  D += A*X      src/kernels.jl:239-241
 *   initkernel!(pD, pA, pX, ::Kernel{...})                     D  = A*X      src/kernels.jl:273-275
 *   fastmul!(D, A, X) whole-matrix small GEMM                               src/kernels.jl:202-208
 *   pick_kernel_size / blocking_structure (static tile policy)              src/kernel_structure.jl:76-99,
 *                                                                           src/memory_management.jl:78-177
 *   mrandn (random fill)                                                    src/randmat.jl:11-14
 * The only FFI precedent in the reference is a ccall of a C-ABI void function
 * taking three host pointers (test/runtests.jl:97-101); this header keeps that
 * shape (plain pointers, sizes, no runtime types) and adds BLAS arguments.
 *
 * Conventions (all entry points):
 *   - Column-major storage, BLAS argument order and meaning:
 *       C = alpha * op(A) * op(B) + beta * C,  op(X) = X ('N') or X^T ('T'/'C'),
 *       op(A) is m x k, op(B) is k x n, C is m x n.
 *     beta == 0 never reads C; alpha == 0 or k == 0 returns beta*C without
 *     reading A or B.
 *   - Return value: 0 on success; -i when the i-th argument is invalid (BLAS
 *     xerbla numbering, counting from 1 in the order written here); a positive
 *     JB_ERR_* code for a runtime failure.  jb_last_error() gives a message
 *     (thread-local).  Validation happens before any HIP call.
 *   - Functions without the _dev suffix take HOST pointers and are synchronous:
 *     they stage operands through the library's device workspace (the drop-in
 *     for jmul!/gemm! on Julia Arrays).  *_dev functions take DEVICE pointers
 *     and are asynchronous on `stream` (a hipStream_t; NULL = default stream),
 *     except the one-shot grouped forms named below, which synchronise.
 *   - Calls are serialised inside the library (one mutex); the device workspace
 *     is owned by the library and released by jb_finalize().
 */
#ifndef JBLAS_HIP_H
#define JBLAS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------- */
#define JB_OK 0
#define JB_ERR_HIP 1         /* a HIP runtime call failed (message has details)   */
#define JB_ERR_ALLOC 2       /* device workspace could not be allocated           */
#define JB_ERR_NONFINITE 3   /* emulated path: an operand holds Inf or NaN         */
#define JB_ERR_GRAPH 4       /* hipGraph capture / instantiate / launch failed     */
#define JB_ERR_STATE 5       /* library not initialised, or a bad handle           */
#define JB_ERR_UNSUPPORTED 6 /* valid request this build does not implement        */

/* ---- methods (grouped launches and the pullback choose one) ---------------- */
#define JB_F64_NATIVE 0      /* FP64 MFMA (v_mfma_f64_16x16x4f64)                  */
#define JB_F64_RECURRENCE 1  /* FP64 VALU, jmul!'s sequential fma order, bit-exact */
#define JB_F32_NATIVE 2      /* FP32 MFMA (v_mfma_f32_16x16x4f32)                  */
#define JB_F32_BF16X3 3      /* FP32 via 3 BF16 slices on BF16 MFMA                */
#define JB_F64_OZAKI_BASE 16 /* JB_F64_OZAKI_BASE + s: split-BF16 DGEMM, s slices  */

/* ---- library state ------------------------------------------------------------ */
/* Selects the device, queries its properties (replaces the build-time CPU probe
 * deps/build.jl:1-38 -> src/cpu_info.jl) and loads the autotuning table named by
 * JB_TUNE_FILE if set.  Called implicitly by the first GEMM call (device 0). */
int jb_init(int device);
/* Frees the device workspace and cached graphs. */
int jb_finalize(void);
/* Last error message of the calling thread ("" if none). */
const char* jb_last_error(void);
/* Library version as MAJOR*10000 + MINOR*100 + PATCH. */
int jb_version(void);

/* Device properties gathered by jb_init (the GPU counterpart of cpu_info.jl's
 * REGISTER_SIZE / REGISTER_COUNT / CACHELINE_SIZE / CACHE_SIZE constants). */
typedef struct jb_device_info {
  char name[64];
  char arch[32];
  int32_t compute_units;
  int32_t clock_khz;         /* peak engine clock */
  int32_t wavefront_size;
  int32_t lds_bytes_per_cu;
  int64_t l2_bytes;
  int64_t hbm_bytes;
  double peak_fp64_tflops;   /* CU * clock * 128 flop/clk */
  double peak_bf16_tflops;   /* CU * clock * 4096 flop/clk (dense) */
  double peak_fp32_tflops;   /* CU * clock * 256 flop/clk */
} jb_device_info;
int jb_get_device_info(jb_device_info* out);

/* ---- per-kernel timing (metrics; no counterpart in the reference, whose only
 * timing is BenchmarkTools in test/runtests.jl:107) ------------------------------ */
/* When on, every kernel launch is bracketed by HIP events on its stream (not
 * inside graph capture).  Kernel names: "dgemm_mfma", "dgemm_recurrence",
 * "sgemm_mfma", "ozaki_split", "slice_pair_gemm", "bf16x3_split", "bf16x3_gemm",
 * "grouped_gemm". */
int jb_kernel_timing_enable(int on);
/* Total milliseconds and launch count of `kernel` since the last reset. */
int jb_kernel_timing_read(const char* kernel, double* total_ms, int64_t* launches);
int jb_kernel_timing_reset(void);

/* ---- jmul! drop-in ------------------------------------------------------------ */
/* D (M x P) = A (M x N) * X (N x P), column-major, leading dimensions M, N, M.
 * Per element: acc_1 = A[i,1]*X[1,j]; acc_n = fma(A[i,n], X[n,j], acc_{n-1}),
 * i.e. src/gemm.jl:86 then :165, bit-identical to the reference recurrence,
 * and (unlike src/gemm.jl:340-345) the fringe rows/columns are written too.
 * Replaces jmul!(D, A, X) src/gemm.jl:244-348.  Host pointers, synchronous. */
int jb_jmul(int64_t M, int64_t N, int64_t P, double* D, const double* A, const double* X);
int jb_jmul_dev(int64_t M, int64_t N, int64_t P, double* D, const double* A, const double* X,
                void* stream);

/* jmul! in Float32 (the reference's jmul! is generic in T, src/gemm.jl:244-246):
 * D = A * X with the same recurrence in binary32 -- acc_1 = fl(A[i,1]*X[1,j]),
 * acc_n = fmaf(A[i,n], X[n,j], acc_{n-1}) in increasing n -- on a sequential-fma
 * kernel (vector ALUs), bit-identical, fringe included.  Leading dimensions M,
 * N, M.  Host pointers (synchronous) / device pointers on `stream`. */
int jb_jmul_f32(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X);
int jb_jmul_f32_dev(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X, void* stream);

/* ---- kernel! / initkernel! drop-in ---------------------------------------------- */
/* kernel!(pD, pA, pX, Kernel{M,P,stride_AD,stride_X,N}) src/kernels.jl:212-241:
 * D (M x P) += A (M x N) * X (N x P) with each element's chain starting from D:
 * acc = D[i,j]; acc = fma(A[i,n], X[n,j], acc) for n = 1..N (the vload of D
 * then the vmuladd k-loop, src/kernels.jl:225-233).  Column-major with leading
 * dimensions ldd, lda, ldx (the reference shares one stride between A and D);
 * N == 0 leaves D unchanged.  Host pointers, synchronous. */
int jb_kernel(int64_t M, int64_t N, int64_t P, double* D, int64_t ldd, const double* A, int64_t lda,
              const double* X, int64_t ldx);
int jb_kernel_dev(int64_t M, int64_t N, int64_t P, double* D, int64_t ldd, const double* A,
                  int64_t lda, const double* X, int64_t ldx, void* stream);
/* initkernel!(pD, pA, pX, K) src/kernels.jl:242-275: D = A * X, the first
 * k-step a plain product (evmul) then fma (vmuladd), as jmul!.  N >= 1. */
int jb_initkernel(int64_t M, int64_t N, int64_t P, double* D, int64_t ldd, const double* A,
                  int64_t lda, const double* X, int64_t ldx);
int jb_initkernel_dev(int64_t M, int64_t N, int64_t P, double* D, int64_t ldd, const double* A,
                      int64_t lda, const double* X, int64_t ldx, void* stream);

/* ---- fastmul! drop-in ---------------------------------------------------------- */
/* D (M x P) = A (M x N) * X (N x P), column-major, leading dimensions M, N, M, as
 * one small-GEMM launch: one wave per 64 rows x 16 columns of D, X read through
 * wave-uniform (scalar) loads.  Same per-element recurrence as jb_jmul
 * (mulinit's plain product src/kernels.jl:3-21, then the vmuladd k-loop :93-99),
 * so the bits equal jb_jmul's.  Meant for small matrices (latency, not
 * throughput); any size is accepted.  Replaces fastmul!(D, A, X)
 * src/kernels.jl:202-208 (which, for M not a multiple of its vector width,
 * stores past D at :110 and, with N == 1, adds the single product twice; here
 * rows >= M are not stored and N == 1 gives D = A*X). */
int jb_fastmul(int64_t M, int64_t N, int64_t P, double* D, const double* A, const double* X);
int jb_fastmul_dev(int64_t M, int64_t N, int64_t P, double* D, const double* A, const double* X,
                   void* stream);
int jb_fastmul_f32(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X);
int jb_fastmul_f32_dev(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X,
                       void* stream);

/* ---- BLAS GEMM (gemm! in Julia: LinearAlgebra.BLAS.gemm!(tA,tB,alpha,A,B,beta,C)) */
/* FP64 on FP64 MFMA.  jmul!(D,A,X) == jb_dgemm('N','N',M,P,N,1,A,M,X,N,0,D,M)
 * up to the summation order of the matrix core. */
int jb_dgemm(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
             int64_t ldc);
int jb_dgemm_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                 const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                 double* C, int64_t ldc, void* stream);

/* FP64 with the reference's sequential-fma order per element (bit-exact vs the
 * jmul! recurrence under BLAS alpha/beta/op semantics; kernel!/initkernel!
 * src/kernels.jl:212-275 with arbitrary leading dimensions). */
int jb_dgemm_recurrence(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                        double* C, int64_t ldc);
int jb_dgemm_recurrence_dev(char transA, char transB, int64_t m, int64_t n, int64_t k,
                            double alpha, const double* A, int64_t lda, const double* B,
                            int64_t ldb, double beta, double* C, int64_t ldc, void* stream);

/* Emulated FP64: operands split into `nslices` (1..8) exponent-scaled BF16
 * slices of balanced 9-bit digits (|d| <= 256, weights 2^(-8-9t) times the
 * row/column exponent: 9*nslices-1 bits below the row/column scale), slice
 * pairs (t+u <= nslices+1) multiplied on BF16 MFMA with exact FP32 level sums,
 * level sums combined in compensated FP64.  Six slices (21 pair products) are
 * as accurate as native FP64 componentwise on wide-range 8192^3 inputs.  The 15th
 * argument (nslices) is validated as argument 14 in the _dev form too.
 * Non-finite operands: every element of C fed by a row of op(A) or a column of
 * op(B) that holds Inf or NaN is set to NaN (the other elements are computed
 * as usual).  The host-pointer form then returns JB_ERR_NONFINITE; the _dev
 * form does not synchronise and returns JB_OK, and the condition is reported by
 * jb_nonfinite_check().  The _dev form can be captured into a hipGraph once a
 * call of the same or a larger size has run outside the capture (the
 * workspace does not grow during capture); replays must not overlap other
 * calls of the library's emulated methods. */
int jb_dgemm_ozaki(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                   const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                   double* C, int64_t ldc, int nslices);
int jb_dgemm_ozaki_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                       const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                       double* C, int64_t ldc, int nslices, void* stream);

/* Allocation-free form of jb_dgemm_ozaki_dev.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and bits as jb_dgemm_ozaki_dev; it calls no hipMalloc,
 * hipFree or allocating routine.  Its workspace is a static device array of
 * JB_NOALLOC_WORKSPACE_BYTES per GPU, held from the library's load until the
 * process ends; the workspace of an m x n x k call with s slices is the sum of
 *   2*s*Mp*Kp, 2*s*Np*Kp, 8*Mp, 8*Np, 4*Mp, 4*Np and 8 bytes, each term
 *   rounded up to a multiple of 256 (Mp = m rounded up to 128, Np = n to 64,
 *   Kp = k to 64), so m = n = k = 1600 fits at s = 6 (62.7 MB) and 1664 does
 *   not.  A larger call returns JB_ERR_UNSUPPORTED
 * before any HIP call.  The library must be initialised first (jb_init;
 * JB_ERR_STATE otherwise).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two streams, or two threads whose streams
 * are not ordered) race and give wrong results; calls ordered on a stream do
 * not.  Non-finite operands: as jb_dgemm_ozaki_dev (NaN elements, reported by
 * jb_nonfinite_check).  No per-kernel timing events are recorded. */
#define JB_NOALLOC_WORKSPACE_BYTES ((size_t)64 << 20)
int jb_dgemm_ozaki_noalloc_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                               const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                               int64_t ldc, int nslices, void* stream);

/* Emulated FP64 on INT8 slices: 8-bit digit weights (2^(-7-8t) times the
 * row/column exponent) with each digit a signed byte of the operand rounded to
 * 8*nslices-1 bits below its row/column scale, slice pairs (t+u <= nslices+1)
 * on v_mfma_i32_32x32x32_i8 with exact int32 level sums (k-chunks of 8192),
 * level sums combined in compensated FP64 as above.  No counterpart in the
 * reference (the drop-in's faster emulated method); nslices 1..8 (argument 14). */
int jb_dgemm_ozaki_i8(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                      const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                      double* C, int64_t ldc, int nslices);
int jb_dgemm_ozaki_i8_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                          const double* A, int64_t lda, const double* B, int64_t ldb, double beta,
                          double* C, int64_t ldc, int nslices, void* stream);

/* FP32 on FP32 MFMA. */
int jb_sgemm(char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha,
             const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C,
             int64_t ldc);
int jb_sgemm_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha,
                 const float* A, int64_t lda, const float* B, int64_t ldb, float beta, float* C,
                 int64_t ldc, void* stream);

/* FP32 emulated with 3 BF16 slices per element (hi, mid, lo) on BF16 MFMA and
 * FP32 accumulation; nprod in {3, 6, 9} slice products (6 = all pairs t+u <= 4).
 * Finite inputs up to FLT_MAX are split exactly.  An Inf or NaN operand element
 * becomes NaN in every product it enters (where native FP32 could give +-Inf)
 * and is reported as for the emulated FP64 methods (JB_ERR_NONFINITE from the
 * host-pointer form, jb_nonfinite_check() after the _dev form). */
int jb_sgemm_bf16x3(char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha,
                    const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
                    float* C, int64_t ldc, int nprod);
int jb_sgemm_bf16x3_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha,
                        const float* A, int64_t lda, const float* B, int64_t ldb, float beta,
                        float* C, int64_t ldc, int nprod, void* stream);

/* Reports Inf/NaN operands met by emulated _dev calls (split-BF16, split-INT8,
 * 3xBF16) since the previous check, on any stream: synchronises `stream`, then
 * returns JB_ERR_NONFINITE (and clears the condition) if any such call met one
 * (the host-pointer forms, which report their own, raise it too),
 * else JB_OK.  Not allowed while `stream` is being captured (JB_ERR_STATE).  No
 * counterpart in the reference (its jmul! has no error reporting). */
int jb_nonfinite_check(void* stream);

/* ---- batched / grouped GEMM ---------------------------------------------------- */
/* One problem of a grouped launch.  A, B, C are DEVICE pointers of the method's
 * element type (double for F64 methods, float for F32 methods). */
typedef struct jb_gemm_desc {
  char transA;
  char transB;
  char pad_[6];
  int64_t m, n, k;
  double alpha, beta;
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  void* C;
  int64_t ldc;
} jb_gemm_desc;

/* All `count` problems in one persistent launch (tile shape and split-K per
 * problem from the autotuning table).  `descs` is a HOST array.
 * Methods: JB_F64_NATIVE, JB_F32_BF16X3, JB_F32_NATIVE.  This one-shot form
 * uploads the descriptors and synchronises `stream` before it returns (use a
 * plan for asynchronous, graph-replayed launches). */
int jb_gemm_grouped_dev(int method, int64_t count, const jb_gemm_desc* descs, void* stream);

/* A grouped plan: descriptors uploaded once, the persistent launch captured in
 * a hipGraph; jb_grouped_plan_launch replays it on `stream`.  Replays of one
 * plan share its work counters, so a replay waits (on the device, through an
 * event) for the previous replay of the same plan, whatever its stream. */
typedef struct jb_grouped_plan jb_grouped_plan;
int jb_grouped_plan_create(int method, int64_t count, const jb_gemm_desc* descs, void* stream,
                           jb_grouped_plan** plan);
int jb_grouped_plan_launch(jb_grouped_plan* plan, void* stream);
int jb_grouped_plan_destroy(jb_grouped_plan* plan);

/* ---- autotuner ------------------------------------------------------------------ */
/* Tunes (tile shape x split-K) per size bucket present in `descs`, in context:
 * every trial times the whole group (fastest of `reps` launches), buckets are
 * visited once in decreasing order of their flops (those under 0.5 % of the
 * total keep the static rule), and a candidate is kept only if it makes the
 * group >= 1 % faster; the tuned choices are then kept only if the group,
 * timed again alternately with the static rule, is still >= 1 % faster.
 * Improved buckets go into the in-memory table (replaces
 * the static pick_kernel_size / blocking_structure choice).  The trials write
 * to scratch copies of the outputs: the group's C matrices are neither read
 * nor written.  Synchronous. */
int jb_autotune_grouped(int method, int64_t count, const jb_gemm_desc* descs, int reps,
                        void* stream);
/* Persist / restore the table as JSON (keyed by device, arch, method, bucket). */
int jb_autotune_save(const char* path);
int jb_autotune_load(const char* path);
/* Query the choice for one problem shape on its own: tile id (0 = 64x64 MFMA
 * tile, 1 = 32x32 MFMA tile, 2 = the whole problem in one work item on the
 * vector ALUs for m, n, k <= 64 -- fastmul!'s form, src/kernels.jl:202-208 --,
 * and, FP64 and 3xBF16 only, each run by a persistent kernel of its own:
 * 3 = 128x128, 4 = 128x64, 5 = 64x128 MFMA tiles) and split-K, from the table
 * or the static rule (FP64: the tile with the least padded area per unit of its
 * measured rate; 3xBF16: 64x128; native FP32: 64x64 or 32x32).  A plan may raise the split of
 * statically chosen problems while the group has fewer work items than the
 * device holds resident workgroups (jb_grouped_choices reports the result). */
int jb_autotune_lookup(int method, int64_t m, int64_t n, int64_t k, int* tile, int* splitk);
/* Clear the table (static defaults apply again). */
int jb_autotune_reset(void);
/* The choice a plan of this group makes now, per problem (arrays of `count`):
 * tile id (as above; -1 for a problem with m or n == 0), split count, and k per
 * split (a multiple of 32: split s sums k-values [s*kchunk, min(k, (s+1)*kchunk)),
 * and the splits' partial results are added to C in split order).  No launch.
 * No counterpart in the reference (its blocking is static); for tests and
 * callers that need to reproduce the grouped launch's summation order. */
int jb_grouped_choices(int method, int64_t count, const jb_gemm_desc* descs, int* tile, int* splitk,
                       int64_t* kchunk);

/* ---- differentiation (ChainRulesCore.rrule of gemm!) ---------------------------- */
/* For C = alpha*op(A)*op(B) + beta*C_in and upstream gradient dC (m x n):
 *   dA (shape of A) = alpha * dC * op(B)^T   (transA 'N')  or  alpha * op(B) * dC^T ('T')
 *   dB (shape of B) = alpha * op(A)^T * dC   (transB 'N')  or  alpha * dC^T * op(A) ('T')
 * Adjoint GEMMs with m*n >= 256 tiles of 128x128 (and k >= 256) run on the
 * single-problem kernels, one after the other on the stream (asynchronously);
 * smaller ones share one one-shot grouped launch (transposed-layout loaders),
 * which synchronises `stream` like jb_gemm_grouped_dev.
 * dA or dB may be NULL to skip it.  method: JB_F64_NATIVE, JB_F32_NATIVE, JB_F32_BF16X3. */
int jb_gemm_pullback_dev(int method, char transA, char transB, int64_t m, int64_t n, int64_t k,
                         double alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                         const void* dC, int64_t lddc, void* dA, int64_t ldda, void* dB,
                         int64_t lddb, void* stream);
int jb_gemm_pullback(int method, char transA, char transB, int64_t m, int64_t n, int64_t k,
                     double alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
                     const void* dC, int64_t lddc, void* dA, int64_t ldda, void* dB,
                     int64_t lddb);

/* ---- seeded device generator (replaces mrandn, src/randmat.jl:11-14) ------------- */
/* out[i] = 2*((x_i >> 11) * 2^-53) - 1, x_i = SplitMix64(seed) output i; bit-identical
 * to the host generator. */
int jb_fill_uniform_dev(uint64_t seed, int64_t n, double* out, void* stream);
/* out[i] = u_i * 2^e_i, u from stream index 2i, e = emin + floor((emax-emin+1) v), v
 * from index 2i+1. */
int jb_fill_wide_dev(uint64_t seed, int64_t n, int emin, int emax, double* out, void* stream);
/* float32 variant of the uniform fill (u rounded to float). */
int jb_fill_uniform_f32_dev(uint64_t seed, int64_t n, float* out, void* stream);
/* float32 variant of the wide-range fill: the FP64 value of jb_fill_wide_dev
 * rounded to float (to nearest).  Bit-identical to the host generator. */
int jb_fill_wide_f32_dev(uint64_t seed, int64_t n, int emin, int emax, float* out, void* stream);
/* Normal draws (mrandn's randn(), src/randmat.jl:5-14) by Box-Muller on the same
 * stream: out[i] = scale * sqrt(-2 ln u1) cos(2 pi u2), u1 = ((x_2i >> 11) + 1) 2^-53,
 * u2 = (x_2i+1 >> 11) 2^-53.  The device's log and cos make these agree with the
 * host generator to a few ulp (not bitwise).  The f32 form rounds scale*z to float. */
int jb_fill_normal_dev(uint64_t seed, int64_t n, double scale, double* out, void* stream);
int jb_fill_normal_f32_dev(uint64_t seed, int64_t n, double scale, float* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* JBLAS_HIP_H */
