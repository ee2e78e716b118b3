// jb_internal.h -- shared declarations of the HIP kernels' host launchers and
// the runtime (not part of the public C ABI; see include/jblas_hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace jb {

// Element type and transposition of one GEMM operand.
struct GemmArgs {
  bool ta, tb;          // op(A) = A^T, op(B) = B^T
  int64_t m, n, k;
  double alpha, beta;
  const void* A;
  int64_t lda;
  const void* B;
  int64_t ldb;
  void* C;
  int64_t ldc;
};

// ---- FP64 -----------------------------------------------------------------
// Native FP64 MFMA GEMM (128x128 block tile, LDS double buffer).
hipError_t launch_dgemm_mfma(const GemmArgs& g, hipStream_t s);
// Sequential-fma FP64 GEMM (bit-exact jmul! recurrence).
hipError_t launch_dgemm_recurrence(const GemmArgs& g, hipStream_t s);
// kernel! (D += A*X) on the FP64 MFMA kernel: each chain starts from C and
// every step is an fma; alpha/beta unused.
hipError_t launch_dgemm_mfma_acc(const GemmArgs& g, hipStream_t s);
// C = beta*C (alpha == 0 or k == 0 quick path); beta == 0 writes zeros.
hipError_t launch_scale_f64(int64_t m, int64_t n, double beta, double* C, int64_t ldc, hipStream_t s);
hipError_t launch_scale_f32(int64_t m, int64_t n, float beta, float* C, int64_t ldc, hipStream_t s);
// Whole-matrix small GEMM D = A*X (fastmul!), one fma chain per element.
hipError_t launch_fastmul_f64(int64_t M, int64_t N, int64_t P, double* D, const double* A, const double* X,
                              hipStream_t s);
hipError_t launch_fastmul_f32(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X,
                              hipStream_t s);

// ---- split-BF16 emulated FP64 --------------------------------------------
// Workspace bytes needed for an m x n x k problem with s slices.
size_t ozaki_workspace_bytes(int64_t m, int64_t n, int64_t k, int s);
// Split + slice-pair GEMM + compensated combine.  `ws` is device memory of
// ozaki_workspace_bytes().  A row of op(A) or column of op(B) holding Inf or
// NaN makes the C elements it feeds NaN and sets two device ints to 1: this
// call's flag, an int at byte offset emulated_flag_offset() of `ws` (cleared
// by the call itself), and `sticky`, which the call never clears.  No host
// synchronisation.
hipError_t launch_dgemm_ozaki(const GemmArgs& g, int s, void* ws, int* sticky, hipStream_t st);
// The same with INT8 slices (v_mfma_i32_32x32x32_i8, exact int32 level sums).
size_t ozaki_i8_workspace_bytes(int64_t m, int64_t n, int64_t k, int s);
hipError_t launch_dgemm_ozaki_i8(const GemmArgs& g, int s, void* ws, int* sticky, hipStream_t st);
// Byte offset of the call's non-finite flag in the workspace of the emulated
// DGEMM (BF16 or INT8 slices) and of the 3xBF16 SGEMM (whose planes have their
// own layout).
size_t emulated_flag_offset(int64_t m, int64_t n, int64_t k, int s, bool i8);
size_t bf16x3_flag_offset(int64_t m, int64_t n, int64_t k);

// ---- FP32 -----------------------------------------------------------------
hipError_t launch_sgemm_mfma(const GemmArgs& g, hipStream_t s);
// Sequential-fma FP32 GEMM (the jmul! recurrence in Float32, bit-exact).
hipError_t launch_sgemm_recurrence(const GemmArgs& g, hipStream_t s);
size_t bf16x3_workspace_bytes(int64_t m, int64_t n, int64_t k);
// Non-finite inputs become NaN slices (every product they enter is NaN) and set
// the two flags as in launch_dgemm_ozaki.
hipError_t launch_sgemm_bf16x3(const GemmArgs& g, int nprod, void* ws, int* sticky, hipStream_t s);

// ---- generator ------------------------------------------------------------
hipError_t launch_fill_uniform(uint64_t seed, int64_t n, double* out, hipStream_t s);
hipError_t launch_fill_uniform_f32(uint64_t seed, int64_t n, float* out, hipStream_t s);
hipError_t launch_fill_wide(uint64_t seed, int64_t n, int emin, int emax, double* out, hipStream_t s);
hipError_t launch_fill_wide_f32(uint64_t seed, int64_t n, int emin, int emax, float* out, hipStream_t s);
hipError_t launch_fill_normal(uint64_t seed, int64_t n, double scale, double* out, hipStream_t s);
hipError_t launch_fill_normal_f32(uint64_t seed, int64_t n, double scale, float* out, hipStream_t s);

// Compute units of the runtime's device (from the device query at init);
// launch geometry that depends on the device reads it.
int runtime_compute_units();

// Per-kernel timing (off unless jb_kernel_timing_enable(1)): HIP events
// recorded on the launch stream around one kernel launch.  Not recorded while
// the stream is being captured into a graph.
void timing_begin(hipStream_t s, const char* kernel);
void timing_end(hipStream_t s, const char* kernel);

// Pointers that a kernel reads from memory (a descriptor, a struct argument)
// carry no address space the compiler can infer, so accesses through them are
// flat accesses, which also count against the LDS wait counter (a wait for an
// LDS read then waits for outstanding global traffic too).  Casting them to
// the global address space gives global loads and stores.
template <typename T> using gptr = __attribute__((address_space(1))) T*;
template <typename T> __device__ __forceinline__ gptr<T> as_global(T* p) { return (gptr<T>)p; }

// XCD-aware block swizzle: hardware places workgroup w on XCD (w % nxcd).
// Returns a tile id such that each XCD owns a contiguous range of tile ids.
__device__ __forceinline__ int xcd_remap(int w, int nblocks, int nxcd) {
  const int q = nblocks / nxcd, r = nblocks % nxcd;
  const int x = w % nxcd, local = w / nxcd;
  return (x < r) ? x * (q + 1) + local : r * (q + 1) + (x - r) * q + local;
}

// Grouped ordering of an (tm x tn) tile grid: GROUP tile-rows walked column by
// column so that concurrently running tiles share A row panels and B column panels.
__device__ __forceinline__ void tile_coords(int tile, int tiles_m, int tiles_n, int group, int& tm, int& tn) {
  const int per_group = group * tiles_n;
  const int gid = tile / per_group;
  const int first_m = gid * group;
  const int gsize = min(tiles_m - first_m, group);
  const int in_group = tile - gid * per_group;
  tm = first_m + in_group % gsize;
  tn = in_group / gsize;
}

}  // namespace jb
