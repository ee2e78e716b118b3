// k_ozaki.hip -- split-BF16 and split-INT8 emulated DGEMM and 3xBF16 SGEMM on gfx950.
//
// The reference has no emulated path (its GEMM is the FP64/FP32 register tile
// of src/gemm.jl:244-348); this file implements the method of the emulated
// DGEMM the drop-in adds:
//
//   1. absmax  (HBM-bound): per row of op(A) / column of op(B), the bit pattern
//      of max|x| (an unsigned max of |x| bits orders finite < inf < NaN).
//   2. split   (HBM-bound): E = frexp exponent of that max; x = ldexp(a, -E);
//      balanced digits d_1 = rint(x*2^8), d_t = rint(r*2^9) of the running
//      residual, |d| <= 256, each exact in BF16 (an integer of magnitude <= 256
//      has at most 8 significant bits), so a slice carries 9 bits.  Slice planes are zero-padded
//      and stored tile-contiguous: for a row block of RB rows (128 for op(A),
//      64 for op(B)) and a k-stage of 16, the RB x 16 digits form one
//      contiguous block (row-major, 32 bytes per row), so a GEMM stage is a
//      handful of contiguous 1 KB pieces per plane.
//   3. slice-pair GEMM (MFMA-bound): 8 waves per block (2 per SIMD), each
//      owning a 32 x 32 tile of the 128 x 64 block tile, one FP32 accumulator
//      set per level l = t+u (pairs t+u <= s+1) on v_mfma_f32_32x32x16_bf16,
//      pairs in square order (step j: the new slices A_j, B_j against all
//      earlier ones, so a step needs two new fragments).  Operand stages stream
//      into an LDS ring of 4 slots (s <= 6; 3 above, whose stages 4 would not
//      fit) by direct global->LDS loads (global_load_lds_dwordx4), three (two)
//      stages ahead (past the last stage the last one again, so every stage
//      has the same loads in flight and no branch around them), one piece per
//      step, one s_barrier per stage.  Every kc_l k-values (kc_l*(l+1) <= 256 for the
//      l+1 pairs of level l, products |d_t d_u| <= 2^16, so a level sum stays
//      within 2^24 and is exact whatever the MFMA's internal order) level l is
//      flushed into an FP64 group accumulator after the stage's last MFMA
//      (behind one branch on the stage's chunk-end mask; the two-stage levels
//      flush branch-free in the next stage), G_g = fma(S_l, 2^-9j, G_g) with
//      l = 3g+j (exact: G_g is a multiple of 2^-18 below 2^35 while
//      k*s <= 2^18; for s > 6, l = 4g+j).  The epilogue combines the groups by
//      TwoSum from the smallest, rescales by 2^(E_i+F_j-16) and applies
//      alpha/beta.  These are the oracle restatement's operations in the same
//      order, so the result is bit-reproducible.
//   INT8 mode: signed-byte digits of 8 bits (split 2c), weights 2^(-7-8(t-1)),
//   the same LDS image (a stage of 32 int8 k has the bytes of a BF16 stage of 16),
//   v_mfma_i32_32x32x32_i8 with exact int32 level sums over k-chunks of 8192,
//   the same flush into groups of four levels (G_g = fma(S_l, 2^-8j, G_g),
//   l = 4g+j) and the same combine with the rescale 2^(E_i+F_j-14).
//   BF16X3 mode (FP32 inputs): per-element split a = a1 + a2 + a3 (bf16 round
//   to nearest of the running residual, exact), products t+u <= LMAX, level 2
//   in one FP32 accumulator and the smaller levels in a second.
#include <algorithm>
#include <type_traits>

#include "jb_internal.h"

namespace jb {
namespace {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));
typedef int i4v __attribute__((ext_vector_type(4)));
typedef int i16v __attribute__((ext_vector_type(16)));
#define JB_LDS __attribute__((address_space(3)))
#define JB_INLINE __attribute__((always_inline))  // lambdas of the kernels: never outlined

constexpr int QBM = 128, QBN = 64;  // block tile (4/WB x 2 waves of 32*WB x 32)
constexpr int QBK = 16;             // k per LDS stage = one MFMA k-step
constexpr int QSLOTS = 3;           // LDS ring depth (prefetch distance 2; 4 and 3 in MODE 0 with s <= 6)
constexpr int QPAD_K = 64;          // k padding of the planes (split tile width)
#ifndef OZ_WB
#define OZ_WB 1  // emulated FP64: 8 waves of 32 x 32 (the FP64 group accumulators need the registers)
#endif
#ifndef OZ_GROUP
#define OZ_GROUP 2  // tile-rows per group of the tile order (measured best of 2, 4, 8, 16)
#endif
#ifndef X3_WB
#define X3_WB 2  // 3xBF16: waves of 64 x 32
#endif

// k-chunk of 0-based level L (L + 1 slice pairs) over which its FP32 sum of
// products of digits |d| <= 256 stays within 2^24, i.e. exact.
__host__ __device__ constexpr int level_chunk(int L) {
  int kc = 256;
  while (kc * (L + 1) > 256) kc >>= 1;
  return kc;
}

// Exponent word of a row of op(A) / column of op(B) that holds Inf or NaN: the
// epilogue stores NaN in every element of C that such a row or column feeds.
constexpr int EXP_NONFINITE = 0x40000000;

// Element offset of (row, k) in one tile-contiguous plane with row block RB.
__device__ __forceinline__ int64_t plane_off(int64_t row, int64_t k, int RB, int64_t Kpad) {
  return (row / RB) * (RB * Kpad) + (k / QBK) * (RB * QBK) + (row % RB) * QBK + (k % QBK);
}

__device__ __forceinline__ int exponent_of_bits(unsigned long long bits) {
  if (bits == 0ull) return 0;
  int e;
  frexp(__longlong_as_double((long long)bits), &e);
  return e;
}

// ---------------------------------------------------------------------------
// 1. absmax: maxbits[r] = max_k bits(|X(r,k)|), X(r,k) = P[r*sr + k*sk].  For
//    non-negative doubles the bit pattern orders like the value, and Inf/NaN
//    patterns exceed every finite one, so they reach the split's check.
//    Register-only streaming reductions, one atomicMax per row and block:
//    rows contiguous (sr == 1): lane = row, each thread AM_RKC columns;
//    otherwise (k contiguous): a wave takes AM_WROWS rows, reading AM_KK
//    consecutive k of a row per pass (64 lanes x 8 bytes), and reduces across
//    its lanes.
// ---------------------------------------------------------------------------
constexpr int AM_RKC = 256;   // columns per thread, rows contiguous
constexpr int AM_WROWS = 16;  // rows per wave, k contiguous
constexpr int AM_KK = 1024;   // k per block, k contiguous

__device__ __forceinline__ unsigned long long abs_bits(double x) {
  return (unsigned long long)__double_as_longlong(fabs(x));
}

__global__ __launch_bounds__(256) void absmax_rows_kernel(const double* __restrict__ P, int64_t sk, int64_t R,
                                                          int64_t K, unsigned long long* __restrict__ maxbits) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= R) return;
  const int64_t k0 = (int64_t)blockIdx.y * AM_RKC;
  const int64_t nk = min<int64_t>(AM_RKC, K - k0);
  const double* __restrict__ p = P + r + k0 * sk;
  unsigned long long m = 0ull;
  int64_t k = 0;
  for (; k + 16 <= nk; k += 16) {
    double v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = p[(k + u) * sk];
#pragma unroll
    for (int u = 0; u < 16; ++u) m = max(m, abs_bits(v[u]));
  }
  for (; k < nk; ++k) m = max(m, abs_bits(p[k * sk]));
  atomicMax(maxbits + r, m);
}

__global__ __launch_bounds__(256) void absmax_kcontig_kernel(const double* __restrict__ P, int64_t sr, int64_t sk,
                                                             int64_t R, int64_t K,
                                                             unsigned long long* __restrict__ maxbits) {
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int64_t rw = (int64_t)blockIdx.x * (4 * AM_WROWS) + (int64_t)wave * AM_WROWS;
  const int64_t k0 = (int64_t)blockIdx.y * AM_KK;
#pragma unroll 2
  for (int i = 0; i < AM_WROWS; ++i) {
    const int64_t r = rw + i;
    if (r >= R) break;  // wave-uniform
    const double* __restrict__ p = P + r * sr;
    unsigned long long m = 0ull;
#pragma unroll
    for (int j = 0; j < AM_KK / 64; ++j) {
      const int64_t k = k0 + lane + 64 * j;
      if (k < K) m = max(m, abs_bits(p[k * sk]));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned long long)__shfl_xor(m, off, 64));
    if (lane == 0) atomicMax(maxbits + r, m);
  }
}

// k contiguous with unit stride, 16-byte aligned rows: each lane loads two
// doubles per instruction, four rows' loads in flight per wave (the general
// form above issues 8-byte loads through a runtime stride).
__global__ __launch_bounds__(256) void absmax_kunit_kernel(const double* __restrict__ P, int64_t sr, int64_t R,
                                                           int64_t K, unsigned long long* __restrict__ maxbits) {
  typedef double V2 __attribute__((ext_vector_type(2)));
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int64_t rw = (int64_t)blockIdx.x * (4 * AM_WROWS) + (int64_t)wave * AM_WROWS;
  const int64_t k0 = (int64_t)blockIdx.y * AM_KK;
  const bool full = k0 + AM_KK <= K;
  for (int i0 = 0; i0 < AM_WROWS; i0 += 4) {
    unsigned long long m[4] = {0ull, 0ull, 0ull, 0ull};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t r = rw + i0 + i;
      if (r >= R) break;  // wave-uniform
      const double* __restrict__ p = P + r * sr + k0;
      if (full) {
        V2 v[AM_KK / 128];
#pragma unroll
        for (int j = 0; j < AM_KK / 128; ++j) v[j] = *reinterpret_cast<const V2*>(p + 2 * lane + 128 * j);
#pragma unroll
        for (int j = 0; j < AM_KK / 128; ++j) m[i] = max(m[i], max(abs_bits(v[j][0]), abs_bits(v[j][1])));
      } else {
        for (int64_t k = lane; k0 + k < K && k < AM_KK; k += 64) m[i] = max(m[i], abs_bits(p[k]));
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t r = rw + i0 + i;
      if (r >= R) break;
      unsigned long long x = m[i];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) x = max(x, (unsigned long long)__shfl_xor(x, off, 64));
      if (lane == 0) atomicMax(maxbits + r, x);
    }
  }
}

void launch_absmax(const double* P, int64_t sr, int64_t sk, int64_t R, int64_t K, unsigned long long* mb,
                   hipStream_t st) {
  if (R <= 0 || K <= 0) return;
  if (sr != 1 && sk == 1 && (reinterpret_cast<uintptr_t>(P) & 15) == 0 && sr % 2 == 0)
    hipLaunchKernelGGL(absmax_kunit_kernel,
                       dim3((unsigned)((R + 4 * AM_WROWS - 1) / (4 * AM_WROWS)), (unsigned)((K + AM_KK - 1) / AM_KK)),
                       dim3(256), 0, st, P, sr, R, K, mb);
  else if (sr == 1)
    hipLaunchKernelGGL(absmax_rows_kernel, dim3((unsigned)((R + 255) / 256), (unsigned)((K + AM_RKC - 1) / AM_RKC)),
                       dim3(256), 0, st, P, sk, R, K, mb);
  else
    hipLaunchKernelGGL(absmax_kcontig_kernel,
                       dim3((unsigned)((R + 4 * AM_WROWS - 1) / (4 * AM_WROWS)), (unsigned)((K + AM_KK - 1) / AM_KK)),
                       dim3(256), 0, st, P, sr, sk, R, K, mb);
}

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

__device__ __forceinline__ uint16_t bf16_bits_exact(double d) {
  // d is an integer with |d| <= 256: exact in float, whose low 16 bits are then zero.
  return (uint16_t)(__float_as_uint((float)d) >> 16);
}

// ---------------------------------------------------------------------------
// 2a. FP64 split into S balanced-digit planes (+ exponents).  The grid covers
//     the padded Rpad x Kpad extent in 64 x 64 tiles; RB divides by 64.
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void split_f64_kernel(const double* __restrict__ P, int64_t sr, int64_t sk, int64_t R,
                                                        int64_t K, const unsigned long long* __restrict__ maxbits,
                                                        int* __restrict__ exps, uint16_t* __restrict__ planes,
                                                        int64_t plane_stride, int64_t Kpad, int RB,
                                                        int* __restrict__ call_flag, int* __restrict__ sticky) {
  __shared__ double tile[64][65];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  const bool rows_contig = (sr == 1);
  {
    // all 16 loads of the thread in flight before the LDS stores
    double v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int r = rows_contig ? (t % 64) : (t / 64 + 4 * j);
      const int k = rows_contig ? (t / 64 + 4 * j) : (t % 64);
      v[j] = (r0 + r < R && k0 + k < K) ? P[(r0 + r) * sr + (k0 + k) * sk] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int r = rows_contig ? (t % 64) : (t / 64 + 4 * j);
      const int k = rows_contig ? (t / 64 + 4 * j) : (t % 64);
      tile[r][k] = v[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int grp = t + 256 * q;  // 512 groups of 8 consecutive k
    const int row = grp / 8, kg = 8 * (grp % 8);
    int E = 0;
    if (r0 + row < R) {
      const unsigned long long mb = maxbits[r0 + row];
      const bool bad = mb >= 0x7FF0000000000000ull;  // Inf or NaN in this row/column
      if (bad) {
        *call_flag = 1;
        *sticky = 1;
      } else {
        E = exponent_of_bits(mb);
      }
      if (blockIdx.y == 0 && kg == 0) exps[r0 + row] = bad ? EXP_NONFINITE : E;
    }
    double res[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) res[e] = ldexp(tile[row][kg + e], -E) * 256.0;
    const int64_t off = plane_off(r0 + row, k0 + kg, RB, Kpad);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      uint32_t w[4];
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        if (s > 0) {
          res[e] *= 512.0;
          res[e + 1] *= 512.0;
        }
        const double d0 = rint(res[e]), d1 = rint(res[e + 1]);
        res[e] -= d0;
        res[e + 1] -= d1;
        w[e / 2] = (uint32_t)bf16_bits_exact(d0) | ((uint32_t)bf16_bits_exact(d1) << 16);
      }
      *reinterpret_cast<u4*>(planes + s * plane_stride + off) = u4{w[0], w[1], w[2], w[3]};
    }
  }
}

// ---------------------------------------------------------------------------
// 2c. FP64 split into S INT8 digit planes (+ exponents).  E = frexp exponent of
//     the row max, plus one when max*2^-E > 126.5/128; X = rint(x * 2^(8S-1))
//     with x = ldexp(a, -E); the digits are X's signed base-256 digits (least
//     significant from plane S-1 up, each in [-128, 127]), so that
//     a ~= 2^E * sum_t d_t 2^(-7-8t).  A plane
//     stage is 32 int8 of k per row (one v_mfma_i32_32x32x32_i8 k-step): its
//     byte image has the shape of a BF16 stage of 16.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t plane_off8(int64_t row, int64_t k, int RB, int64_t Kpad) {
  return (row / RB) * (RB * Kpad) + (k / 32) * (RB * 32) + (row % RB) * 32 + (k % 32);
}

template <int S>
__global__ __launch_bounds__(256) void split_f64_i8_kernel(const double* __restrict__ P, int64_t sr, int64_t sk,
                                                           int64_t R, int64_t K,
                                                           const unsigned long long* __restrict__ maxbits,
                                                           int* __restrict__ exps, uint8_t* __restrict__ planes,
                                                           int64_t plane_stride, int64_t Kpad, int RB,
                                                           int* __restrict__ call_flag, int* __restrict__ sticky) {
  __shared__ double tile[64][65];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  const bool rows_contig = (sr == 1);
  {
    // all 16 loads of the thread in flight before the LDS stores
    double v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int r = rows_contig ? (t % 64) : (t / 64 + 4 * j);
      const int k = rows_contig ? (t / 64 + 4 * j) : (t % 64);
      v[j] = (r0 + r < R && k0 + k < K) ? P[(r0 + r) * sr + (k0 + k) * sk] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int r = rows_contig ? (t % 64) : (t / 64 + 4 * j);
      const int k = rows_contig ? (t / 64 + 4 * j) : (t % 64);
      tile[r][k] = v[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int grp = t + 256 * q;  // 512 groups of 8 consecutive k
    const int row = grp / 8, kg = 8 * (grp % 8);
    int E = 0;
    if (r0 + row < R) {
      const unsigned long long mb = maxbits[r0 + row];
      const bool bad = mb >= 0x7FF0000000000000ull;  // Inf or NaN in this row/column
      if (bad) {
        *call_flag = 1;
        *sticky = 1;
      } else if (mb != 0ull) {
        E = exponent_of_bits(mb);
        if (ldexp(__longlong_as_double((long long)mb), -E) > 126.5 / 128.0) E += 1;
      }
      if (blockIdx.y == 0 && kg == 0) exps[r0 + row] = bad ? EXP_NONFINITE : E;
    }
    long long X[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) X[e] = (long long)rint(ldexp(tile[row][kg + e], 8 * S - 1 - E));
    const int64_t off = plane_off8(r0 + row, k0 + kg, RB, Kpad);
#pragma unroll
    for (int s = S - 1; s >= 0; --s) {
      uint32_t w[2] = {0u, 0u};
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const long long d = (s > 0) ? (long long)(int8_t)(uint8_t)(X[e] & 255) : X[e];
        if (s > 0) X[e] = (X[e] - d) >> 8;
        w[e / 4] |= ((uint32_t)(uint8_t)(int8_t)d) << (8 * (e % 4));
      }
      *reinterpret_cast<uint2*>(planes + s * plane_stride + off) = uint2{w[0], w[1]};
    }
  }
}

// ---------------------------------------------------------------------------
// 2b. FP32 split into 3 BF16 planes: a1 = bf16(a), a2 = bf16(a - a1), a3 = bf16(rest).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t bf16_rn_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__global__ __launch_bounds__(256) void split_f32_kernel(const float* __restrict__ P, int64_t sr, int64_t sk, int64_t R,
                                                        int64_t K, uint16_t* __restrict__ planes, int64_t plane_stride,
                                                        int64_t Kpad, int RB, int* __restrict__ call_flag,
                                                        int* __restrict__ sticky) {
  __shared__ float tile[64][65];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  const bool rows_contig = (sr == 1);
  {
    // all 16 loads of the thread in flight before the LDS stores
    float v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int r = rows_contig ? (t % 64) : (t / 64 + 4 * j);
      const int k = rows_contig ? (t / 64 + 4 * j) : (t % 64);
      v[j] = (r0 + r < R && k0 + k < K) ? P[(r0 + r) * sr + (k0 + k) * sk] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int r = rows_contig ? (t % 64) : (t / 64 + 4 * j);
      const int k = rows_contig ? (t / 64 + 4 * j) : (t % 64);
      tile[r][k] = v[j];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int grp = t + 256 * q;
    const int row = grp / 8, kg = 8 * (grp % 8);
    uint32_t w[3][4];
    bool bad = false;
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
      uint16_t b[3][2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float r = tile[row][kg + e + h];
        if (!isfinite(r)) {
          // Inf or NaN: one NaN slice, so that every product it enters is NaN
          bad = true;
          b[0][h] = 0x7FC0;
          b[1][h] = b[2][h] = 0;
          continue;
        }
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          b[s][h] = bf16_rn_bits(r);
          // |a| within half a BF16 ulp of FLT_MAX rounds to Inf: truncate instead
          // (the residual then has at most 16 significant bits, still exact)
          if (s == 0 && (b[s][h] & 0x7F80u) == 0x7F80u) b[s][h] = (uint16_t)(__float_as_uint(r) >> 16);
          r = r - bf16_to_f(b[s][h]);
        }
      }
#pragma unroll
      for (int s = 0; s < 3; ++s) w[s][e / 2] = (uint32_t)b[s][0] | ((uint32_t)b[s][1] << 16);
    }
    if (bad) {
      *call_flag = 1;
      *sticky = 1;
    }
    const int64_t off = plane_off(r0 + row, k0 + kg, RB, Kpad);
#pragma unroll
    for (int s = 0; s < 3; ++s)
      *reinterpret_cast<u4*>(planes + s * plane_stride + off) = u4{w[s][0], w[s][1], w[s][2], w[s][3]};
  }
}

// ---------------------------------------------------------------------------
// 3. slice-pair GEMM.  MODE 0: emulated FP64 (S BF16 slices of 9 bits, exact
//    FP32 level sums, FP64 groups of three levels).  MODE 1: 3xBF16 FP32
//    (S = 3, pairs t+u <= LMAX).  MODE 2: emulated FP64 on S INT8 slices of 8
//    bits (v_mfma_i32_32x32x32_i8, exact int32 level sums, 32 k per stage, same
//    LDS image as MODE 0, FP64 groups of four levels).
// ---------------------------------------------------------------------------
struct PairArgs {
  const uint16_t* pa;  // op(A) planes, row block 128
  const uint16_t* pb;  // op(B) planes, row block 64
  int64_t Kpad, strideA, strideB;
  const int* eA;
  const int* eB;
  int64_t M, N;
  double alpha, beta;
  void* C;
  int64_t ldc;
  int tiles_m, tiles_n;
};

__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
  const double x = a + b, z = x - a;
  e = (a - (x - z)) + (b - z);
  s = x;
}

// s_waitcnt with only the vector-memory counter constrained (gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

// Workgroup barrier without the fence __syncthreads() carries: the fence
// would wait for every in-flight global->LDS load, draining the ring.  The
// empty asm statements keep the compiler from moving LDS accesses across it.
__device__ __forceinline__ void block_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// WB = 32x32 MFMA blocks per wave along m: WB = 1 -> waves of 32 x 32, WB = 2
// -> waves of 64 x 32; block tile 128 x BN (BN = 64: 8 or 4 waves; the 3xBF16
// mode: BN = 128 or 256, 8 or 16 waves of 64 x 32, fewer B bytes per product).
template <int MODE, int S, int LMAX, int WB, int BN = QBN>
__global__ __launch_bounds__((64 * (QBM / (32 * WB)) * (BN / 32)), 1) void slice_pair_gemm_kernel(PairArgs p) {
  constexpr int WM = QBM / (32 * WB), NWAVES = WM * (BN / 32);  // waves along m, waves
  constexpr int PLANE_B = (QBM + BN) * QBK * 2;  // bytes of one plane in one stage (6 KB for BN = 64)
  constexpr int STAGE_B = S * PLANE_B;
  constexpr int NWL = (QBM / 32 + BN / 32) * S;  // 1 KB wave-loads per stage
  constexpr int WL_MIN = NWL / NWAVES;            // wave-loads per stage of waves >= NWL % NWAVES
  constexpr int WL_MAX = (NWL + NWAVES - 1) / NWAVES;  // ... of waves < NWL % NWAVES
  constexpr int NL = (MODE == 1) ? 2 : S;         // accumulator sets (levels)
  constexpr int DB = (MODE == 2) ? 8 : 9;         // bits per digit (level weight 2^-DB)
  constexpr int GL = (MODE == 2 || S > 6) ? 4 : 3;  // levels per FP64 group (three: exact groups; four
                                                     // above six levels, whose registers three would spill)
  constexpr int NG = (S + GL - 1) / GL;           // FP64 groups (MODES 0, 2)
  // stages per exact chunk of level L: FP32 level sums within 2^24 (MODE 0:
  // level_chunk(L) k); int32 sums of 8192 k (|S_l| <= 8 * 8192 * 2^14 = 2^30)
  // in MODE 2, whose stages hold 32 k.  All powers of two.
  auto kc_stages = [](int L) { return (MODE == 2) ? 8192 / 32 : level_chunk(L) / QBK; };
  typedef typename std::conditional<MODE == 2, i16v, f16v>::type AccV;
  // LDS ring depth: four slots (three stages of loads in flight) where they fit
  // the 160 KB of LDS in the emulated-FP64 mode (s <= 6: 2.5-3 % faster at
  // 8192^3 than three), else three
  constexpr int NSLOT = (MODE == 0 && 4 * STAGE_B <= 160 * 1024) ? 4 : QSLOTS;
  constexpr int PD = NSLOT - 1;  // prefetch distance
  __shared__ __attribute__((aligned(1024))) char ring[NSLOT * STAGE_B];

  const int nblocks = p.tiles_m * p.tiles_n;
  const int tile = xcd_remap(blockIdx.x, nblocks, 8);
  int tm, tn;
  tile_coords(tile, p.tiles_m, p.tiles_n, OZ_GROUP, tm, tn);
  const int64_t m0 = (int64_t)tm * QBM, n0 = (int64_t)tn * BN;
  const int tid = threadIdx.x, lane = tid % 64;
  // wave-uniform (SGPR) so that the DMA issue below is scalar code
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
  const int wm0 = 32 * WB * (wave % WM), wn0 = 32 * (wave / WM);
  const int l32 = lane % 32, hl = lane / 32;
  const int phl = hl ^ ((l32 >> 3) & 1);  // physical half of this lane's fragment in the LDS image

  // LDS image of a plane stage: 32 bytes per row (two 16-byte halves of 8 k),
  // with the halves of rows whose bit 3 is set swapped, so that the 16 lanes
  // of one ds_read_b128 phase (16 rows, same half) hit 16 distinct 4-bank
  // groups of the 64-bank LDS.  A DMA lane writes LDS chunk `lane` of its
  // 1 KB piece (row lane/2, physical half lane%2) and so fetches the logical
  // half (lane%2) ^ bit 3 of its row from global memory.
  const int dma_off = (lane / 2) * QBK + 8 * ((lane % 2) ^ ((lane >> 4) & 1));
  // Wave-uniform bases: a piece's address is the uniform part (formed in scalar
  // registers) plus the lane's offset, added last.  Adding the lane offset
  // first put a 64-bit vector add (and selects) in front of every load; this
  // form took the s = 6 kernel from 19.9 to 19.0 ms at 8192^3.
  const uint16_t* gA = p.pa + (int64_t)tm * QBM * p.Kpad;
  const uint16_t* gB = p.pb + (int64_t)tn * BN * p.Kpad;
  // Piece i of this wave's share of stage ks (pieces j = wave + NWAVES*i; A:
  // plane t, 4 pieces of 1 KB; B: plane t, BN/32 pieces) -> LDS slot `slot`.
  auto issue_piece = [&](int ks, int slot, int i) JB_INLINE {
    const int j = wave + NWAVES * i;  // uniform
    if (j >= NWL) return;
    const uint16_t* src;
    int loff;
    if (j < 4 * S) {
      const int t = j / 4, q = j % 4;
      src = gA + t * p.strideA + (int64_t)ks * (QBM * QBK) + q * 512;
      loff = t * PLANE_B + q * 1024;
    } else {
      const int jj = j - 4 * S, t = jj / (BN / 32), q = jj % (BN / 32);
      src = gB + t * p.strideB + (int64_t)ks * (BN * QBK) + q * 512;
      loff = t * PLANE_B + QBM * QBK * 2 + q * 1024;
    }
    __builtin_amdgcn_global_load_lds((const void*)(src + (uint32_t)dma_off), (JB_LDS void*)(ring + slot * STAGE_B + loff),
                                     16, 0, 0);
  };
  auto issue = [&](int ks, int slot) JB_INLINE {
#pragma unroll
    for (int i = 0; i < WL_MAX; ++i) issue_piece(ks, slot, i);
  };

  AccV acc[NL][WB];
#pragma unroll
  for (int l = 0; l < NL; ++l)
#pragma unroll
    for (int b = 0; b < WB; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[l][b][e] = 0;
  double G[MODE != 1 ? NG : 1][WB][16];
  if (MODE != 1) {
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
      for (int b = 0; b < WB; ++b)
#pragma unroll
        for (int e = 0; e < 16; ++e) G[g][b][e] = 0.0;
  }

  // Arguments that only the epilogue reads are pulled into SGPRs here: a scalar
  // load still in flight across the main loop would turn every LDS wait inside
  // it into a full lgkmcnt(0) (scalar loads return out of order).
  int64_t eN = p.N, eldc = p.ldc;
  double ealpha = p.alpha, ebeta = p.beta;
  void* eC = p.C;
  asm volatile("" : "+s"(eN), "+s"(eldc), "+s"(ealpha), "+s"(ebeta), "+s"(eC));

  // Flush of level j: G_g += S_j * 2^(-DB*(j%GL)) (S_j an exact integer, |S_j| <= 2^24);
  // the update itself is exact while G_g fits 53 bits (groups of three, k*s <= 2^18).
  auto flush = [&](const int j, const bool zero) JB_INLINE {
    const double wj = __builtin_ldexp(1.0, -DB * (j % GL));  // folds to a constant
#pragma unroll
    for (int b = 0; b < WB; ++b)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        G[j / GL][b][e] = fma((double)acc[j][b][e], wj, G[j / GL][b][e]);
        if (zero) acc[j][b][e] = 0;
      }
  };
  // "Fast" levels (MODE 0: an exact chunk of two stages; the highest levels, the
  // ones with the most pairs) flush without a branch: stages run in pairs
  // (phase 0, phase 1), a fast level's chunk ends with every phase-1 stage, and
  // its flush is deferred into the next phase-0 stage, just before the level's
  // first MFMA there, which then starts from a zero accumulator (no zeroing
  // pass).  Per chunk end the flushes stay in increasing level order.
  auto fast = [&](const int L) { return MODE == 0 && S <= 6 && kc_stages(L) == 2; };  // (S > 6 would spill)

  const int nst = (int)(p.Kpad / QBK);
  // The prefetch is unconditional: past the last stage it loads the last stage
  // again, into the slot of a stage every wave is done with, so every stage has
  // the same loads in flight and the stage code has no branch around its loads
  // (branches split the MFMA stream into basic blocks the scheduler cannot
  // overlap).  The kernel waits for those loads before it ends.
#pragma unroll
  for (int i = 0; i < PD; ++i) issue(min(i, nst - 1), i);
  // chunk-end mask of the non-fast levels after stage st (bit l: level l's exact
  // chunk ends here and it is flushed), the stage loop's runtime form
  auto chunk_mask = [&](const int st) JB_INLINE {
    int m = 0;
#pragma unroll
    for (int l = 0; l < NL; ++l)
      if (MODE != 1 && !fast(l) && (st + 1 == nst || ((st + 1) & (kc_stages(l) - 1)) == 0)) m |= 1 << l;
    return m;
  };
  auto stage = [&](const int st, auto phase, auto cmask) JB_INLINE {
    constexpr int PH = decltype(phase)::value;
    const int mask = cmask;  // a constant where the caller knows the stage index statically
    // this stage landed; the PD - 1 later stages' loads (this wave's count each) may be in flight
    if (wave < NWL % NWAVES) wait_vmcnt<(PD - 1) * WL_MAX>();
    else wait_vmcnt<(PD - 1) * WL_MIN>();
    block_barrier();  // every wave's loads of stage st landed, every wave is done with st-1
    // Stage st+PD goes into the slot of st-1, one piece per step below: a burst of
    // loads right after the barrier fills the memory pipeline's queue and holds
    // up the wave's MFMA issue behind it.
    const int pstage = min(st + PD, nst - 1);
    const int pslot = (st + PD) % NSLOT;
    const u4* L = reinterpret_cast<const u4*>(ring + (st % NSLOT) * STAGE_B);
    // Pairs run in square order: step j multiplies the new slices A_j and B_j
    // with all earlier ones (pairs with max(t, u) == j).  A step needs only two
    // new fragments, read one step ahead, so a stage starts after two LDS reads
    // instead of all of them.  Level l = t + u is complete after step l; its
    // first pair is (t, u) = ((l+1)/2, l-(l+1)/2), at step (l+1)/2.
    constexpr int PPS = (WL_MAX + S - 1) / S;  // load pieces per step
    auto mfma = [&](const int t, const int u, const bf8& b_frag, const bf8 (&a_frag)[WB], const bool first_of_level) JB_INLINE {
      if (t + u + 2 > LMAX) return;
      const int l = (MODE != 1) ? (t + u) : (t + u == 0 ? 0 : 1);
      // first MFMA of a fast level in a phase-0 stage: flush the chunk that ended
      // with the previous stage (zeros at st == 0), then start from zero
      const bool restart = PH == 0 && fast(l) && first_of_level;
      if (restart) flush(l, false);
      // X = op(B) plane (p -> n), Y = op(A) plane (q -> m): item r of lane l holds
      // C[m0+wm0+32b+l%32][n0+wn0+8(r/4)+4(l/32)+r%4].
#pragma unroll
      for (int b = 0; b < WB; ++b)
        if constexpr (MODE == 2)
          acc[l][b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(__builtin_bit_cast(i4v, b_frag),
                                                            __builtin_bit_cast(i4v, a_frag[b]), acc[l][b], 0, 0, 0);
        else
          acc[l][b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b_frag, a_frag[b], restart ? AccV{} : acc[l][b], 0, 0, 0);
    };
    auto frag_a = [&](const int t, const int b) JB_INLINE {
      return __builtin_bit_cast(bf8, L[t * (PLANE_B / 16) + (wm0 + 32 * b + l32) * 2 + phl]);
    };
    auto frag_b = [&](const int u) JB_INLINE { return __builtin_bit_cast(bf8, L[u * (PLANE_B / 16) + (QBM + wn0 + l32) * 2 + phl]); };
    {
      bf8 af[S][WB], bf[S];
#pragma unroll
      for (int b = 0; b < WB; ++b) af[0][b] = frag_a(0, b);
      bf[0] = frag_b(0);
#pragma unroll
      for (int j = 0; j < S; ++j) {
#pragma unroll
        for (int i = PPS * j; i < PPS * (j + 1) && i < WL_MAX; ++i) issue_piece(pstage, pslot, i);
        if (j + 1 < S) {
#pragma unroll
          for (int b = 0; b < WB; ++b) af[j + 1][b] = frag_a(j + 1, b);
          bf[j + 1] = frag_b(j + 1);
        }
#pragma unroll
        for (int q = 0; q < 2 * j + 1; ++q) {
          const int t = q <= j ? j : q - j - 1;  // (j, 0..j), then (0..j-1, j)
          const int u = q <= j ? q : j;
          const int l = t + u;
          mfma(t, u, bf[u], af[t], t == (l + 1) / 2 && u == l - (l + 1) / 2);
        }
      }
    }
    // The levels whose exact chunk ends with this stage, in increasing level order,
    // after the stage's last MFMA (their own last MFMAs were issued steps earlier):
    // one branch on the mask, taken at chunk ends only.
    if (MODE != 1 && mask != 0) {
#pragma unroll
      for (int l = 0; l < NL; ++l)
        if ((mask >> l) & 1) flush(l, true);
    }
  };
  if constexpr (MODE == 0) {
    int st = 0;
    // nst is a multiple of 2 (the planes' k padding is a multiple of 64); the
    // guard keeps an odd count correct too (the last chunk then ends in phase 0)
    for (; st < nst; st += 2) {
      stage(st, std::integral_constant<int, 0>(), chunk_mask(st));
      if (st + 1 < nst) stage(st + 1, std::integral_constant<int, 1>(), chunk_mask(st + 1));
    }
  } else {
    // no fast levels: one stage per iteration (pairing the stages cost the
    // 3xBF16 kernel ~20 %)
    for (int st = 0; st < nst; ++st) stage(st, std::integral_constant<int, 1>(), chunk_mask(st));
  }
  wait_vmcnt<0>();  // the last stages' repeated prefetch lands before the workgroup's LDS is released
  if (MODE == 0) {
#pragma unroll
    for (int j = 0; j < NL; ++j)
      if (fast(j)) flush(j, false);  // the fast levels' last chunk
  }

  if (MODE != 1) {
    gptr<double> C = as_global(static_cast<double*>(eC));
    gptr<const int> eAg = as_global(p.eA), eBg = as_global(p.eB);
#pragma unroll
    for (int b = 0; b < WB; ++b) {
      const int64_t m = m0 + wm0 + 32 * b + l32;
      if (m >= p.M) continue;
      const int ea = eAg[m];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t n = n0 + wn0 + 8 * (e / 4) + 4 * hl + (e % 4);
        if (n < eN) {
          const int eb = eBg[n];
          double hi = 0.0, lo = 0.0;
#pragma unroll
          for (int g = NG - 1; g >= 0; --g) {
            double err;
            two_sum(hi, ldexp(G[g][b][e], -DB * GL * g), hi, err);
            lo += err;
          }
          const double v = ldexp(hi + lo, ea + eb - (MODE == 2 ? 14 : 16));
          double c = ealpha * v;
          if (ebeta != 0.0) c = fma(ebeta, C[m + n * eldc], c);
          if (ea == EXP_NONFINITE || eb == EXP_NONFINITE) c = __builtin_nan("");
          C[m + n * eldc] = c;
        }
      }
    }
  } else {
    gptr<float> C = as_global(static_cast<float*>(eC));
    const float alpha = (float)ealpha, beta = (float)ebeta;
#pragma unroll
    for (int b = 0; b < WB; ++b) {
      const int64_t m = m0 + wm0 + 32 * b + l32;
      if (m >= p.M) continue;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t n = n0 + wn0 + 8 * (e / 4) + 4 * hl + (e % 4);
        if (n < eN) {
          float c = alpha * (acc[0][b][e] + acc[1][b][e]);
          if (beta != 0.0f) c = fmaf(beta, C[m + n * eldc], c);
          C[m + n * eldc] = c;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
inline int64_t rup(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

struct Layout {
  int64_t Mpad, Npad, Kpad;
  size_t offA, offB, offMA, offMB, offEA, offEB, offFlag, total;
};

// elt: bytes per digit (2: BF16 planes, 1: INT8 planes); bn: block tile width
// of the slice-pair kernel (op(B) row block of the planes).
Layout make_layout(int64_t m, int64_t n, int64_t k, int s, int elt = 2, int bn = QBN) {
  Layout L;
  L.Mpad = rup(std::max<int64_t>(m, 1), QBM);
  L.Npad = rup(std::max<int64_t>(n, 1), bn);
  L.Kpad = rup(std::max<int64_t>(k, 1), QPAD_K);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t off = 0;
  L.offA = off;
  off = al(off + (size_t)s * L.Mpad * L.Kpad * elt);
  L.offB = off;
  off = al(off + (size_t)s * L.Npad * L.Kpad * elt);
  L.offMA = off;
  off = al(off + (size_t)L.Mpad * 8);
  L.offMB = off;
  off = al(off + (size_t)L.Npad * 8);
  L.offEA = off;
  off = al(off + (size_t)L.Mpad * 4);
  L.offEB = off;
  off = al(off + (size_t)L.Npad * 4);
  L.offFlag = off;  // this call's non-finite flag
  off = al(off + 8);
  L.total = off;
  return L;
}

// op(A) row r, k: A[r + k*lda] ('N') or A[k + r*lda] ('T'); op(B) column r, k:
// B[k + r*ldb] ('N') or B[r + k*ldb] ('T').
inline void strides_a(bool ta, int64_t lda, int64_t& sr, int64_t& sk) {
  sr = ta ? lda : 1;
  sk = ta ? 1 : lda;
}
inline void strides_b(bool tb, int64_t ldb, int64_t& sr, int64_t& sk) {
  sr = tb ? 1 : ldb;
  sk = tb ? ldb : 1;
}

template <int S>
void launch_split_f64(const double* P, int64_t sr, int64_t sk, int64_t R, int64_t K, int64_t Rpad, int64_t Kpad, int RB,
                      unsigned long long* mb, int* ex, uint16_t* planes, int* call_flag, int* sticky, hipStream_t st) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)std::min<int64_t>((Rpad + 255) / 256, 1024)), dim3(256), 0, st,
                     mb, Rpad);
  launch_absmax(P, sr, sk, R, K, mb, st);
  hipLaunchKernelGGL(split_f64_kernel<S>, dim3((unsigned)(Rpad / 64), (unsigned)(Kpad / 64)), dim3(256), 0, st, P, sr,
                     sk, R, K, mb, ex, planes, Rpad * Kpad, Kpad, RB, call_flag, sticky);
}

// Clears this call's non-finite flag (a kernel, not a memset node, so that the
// call can be captured into a graph).
void clear_call_flag(int* call_flag, hipStream_t st) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned long long*>(call_flag),
                     (int64_t)1);
}

// elt as in make_layout: the kernel addresses planes in 16-bit units, so an
// INT8 plane of Kpad k is a BF16-shaped plane of Kpad/2.
PairArgs pair_args(const GemmArgs& g, const Layout& L, const uint16_t* pA, const uint16_t* pB, const int* eA,
                   const int* eB, int elt = 2, int bn = QBN) {
  PairArgs p;
  p.pa = pA;
  p.pb = pB;
  p.Kpad = L.Kpad * elt / 2;
  p.strideA = L.Mpad * p.Kpad;
  p.strideB = L.Npad * p.Kpad;
  p.eA = eA;
  p.eB = eB;
  p.M = g.m;
  p.N = g.n;
  p.alpha = g.alpha;
  p.beta = g.beta;
  p.C = g.C;
  p.ldc = g.ldc;
  p.tiles_m = (int)(L.Mpad / QBM);
  p.tiles_n = (int)(L.Npad / bn);
  return p;
}

template <int S>
hipError_t run_ozaki(const GemmArgs& g, const Layout& L, char* ws, int* sticky, hipStream_t st) {
  int64_t sr, sk;
  int* flag = reinterpret_cast<int*>(ws + L.offFlag);
  auto* mbA = reinterpret_cast<unsigned long long*>(ws + L.offMA);
  auto* mbB = reinterpret_cast<unsigned long long*>(ws + L.offMB);
  auto* eA = reinterpret_cast<int*>(ws + L.offEA);
  auto* eB = reinterpret_cast<int*>(ws + L.offEB);
  auto* pA = reinterpret_cast<uint16_t*>(ws + L.offA);
  auto* pB = reinterpret_cast<uint16_t*>(ws + L.offB);
  timing_begin(st, "ozaki_split");
  clear_call_flag(flag, st);
  strides_a(g.ta, g.lda, sr, sk);
  launch_split_f64<S>(static_cast<const double*>(g.A), sr, sk, g.m, g.k, L.Mpad, L.Kpad, QBM, mbA, eA, pA, flag, sticky,
                      st);
  strides_b(g.tb, g.ldb, sr, sk);
  launch_split_f64<S>(static_cast<const double*>(g.B), sr, sk, g.n, g.k, L.Npad, L.Kpad, QBN, mbB, eB, pB, flag, sticky,
                      st);
  timing_end(st, "ozaki_split");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const PairArgs p = pair_args(g, L, pA, pB, eA, eB);
  timing_begin(st, "slice_pair_gemm");
  hipLaunchKernelGGL((slice_pair_gemm_kernel<0, S, S + 1, OZ_WB>), dim3(p.tiles_m * p.tiles_n), dim3(512 / OZ_WB), 0, st, p);
  timing_end(st, "slice_pair_gemm");
  return hipGetLastError();
}

template <int S>
void launch_split_f64_i8(const double* P, int64_t sr, int64_t sk, int64_t R, int64_t K, int64_t Rpad, int64_t Kpad,
                         int RB, unsigned long long* mb, int* ex, uint8_t* planes, int* call_flag, int* sticky,
                         hipStream_t st) {
  hipLaunchKernelGGL(zero_u64_kernel, dim3((unsigned)std::min<int64_t>((Rpad + 255) / 256, 1024)), dim3(256), 0, st,
                     mb, Rpad);
  launch_absmax(P, sr, sk, R, K, mb, st);
  hipLaunchKernelGGL(split_f64_i8_kernel<S>, dim3((unsigned)(Rpad / 64), (unsigned)(Kpad / 64)), dim3(256), 0, st, P,
                     sr, sk, R, K, mb, ex, planes, Rpad * Kpad, Kpad, RB, call_flag, sticky);
}

template <int S>
hipError_t run_ozaki_i8(const GemmArgs& g, const Layout& L, char* ws, int* sticky, hipStream_t st) {
  int64_t sr, sk;
  int* flag = reinterpret_cast<int*>(ws + L.offFlag);
  auto* mbA = reinterpret_cast<unsigned long long*>(ws + L.offMA);
  auto* mbB = reinterpret_cast<unsigned long long*>(ws + L.offMB);
  auto* eA = reinterpret_cast<int*>(ws + L.offEA);
  auto* eB = reinterpret_cast<int*>(ws + L.offEB);
  auto* pA = reinterpret_cast<uint8_t*>(ws + L.offA);
  auto* pB = reinterpret_cast<uint8_t*>(ws + L.offB);
  timing_begin(st, "ozaki_i8_split");
  clear_call_flag(flag, st);
  strides_a(g.ta, g.lda, sr, sk);
  launch_split_f64_i8<S>(static_cast<const double*>(g.A), sr, sk, g.m, g.k, L.Mpad, L.Kpad, QBM, mbA, eA, pA, flag,
                         sticky, st);
  strides_b(g.tb, g.ldb, sr, sk);
  launch_split_f64_i8<S>(static_cast<const double*>(g.B), sr, sk, g.n, g.k, L.Npad, L.Kpad, QBN, mbB, eB, pB, flag,
                         sticky, st);
  timing_end(st, "ozaki_i8_split");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const PairArgs p = pair_args(g, L, reinterpret_cast<const uint16_t*>(pA), reinterpret_cast<const uint16_t*>(pB), eA,
                               eB, 1);
  timing_begin(st, "slice_pair_gemm_i8");
  hipLaunchKernelGGL((slice_pair_gemm_kernel<2, S, S + 1, OZ_WB>), dim3(p.tiles_m * p.tiles_n), dim3(512 / OZ_WB), 0,
                     st, p);
  timing_end(st, "slice_pair_gemm_i8");
  return hipGetLastError();
}

}  // namespace

size_t ozaki_workspace_bytes(int64_t m, int64_t n, int64_t k, int s) { return make_layout(m, n, k, s).total; }

size_t emulated_flag_offset(int64_t m, int64_t n, int64_t k, int s, bool i8) {
  return make_layout(m, n, k, s, i8 ? 1 : 2).offFlag;
}

size_t ozaki_i8_workspace_bytes(int64_t m, int64_t n, int64_t k, int s) { return make_layout(m, n, k, s, 1).total; }

hipError_t launch_dgemm_ozaki_i8(const GemmArgs& g, int s, void* ws, int* sticky, hipStream_t st) {
  const Layout L = make_layout(g.m, g.n, g.k, s, 1);
  char* w = static_cast<char*>(ws);
  switch (s) {
    case 1: return run_ozaki_i8<1>(g, L, w, sticky, st);
    case 2: return run_ozaki_i8<2>(g, L, w, sticky, st);
    case 3: return run_ozaki_i8<3>(g, L, w, sticky, st);
    case 4: return run_ozaki_i8<4>(g, L, w, sticky, st);
    case 5: return run_ozaki_i8<5>(g, L, w, sticky, st);
    case 6: return run_ozaki_i8<6>(g, L, w, sticky, st);
    case 7: return run_ozaki_i8<7>(g, L, w, sticky, st);
    case 8: return run_ozaki_i8<8>(g, L, w, sticky, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_dgemm_ozaki(const GemmArgs& g, int s, void* ws, int* sticky, hipStream_t st) {
  const Layout L = make_layout(g.m, g.n, g.k, s);
  char* w = static_cast<char*>(ws);
  switch (s) {
    case 1: return run_ozaki<1>(g, L, w, sticky, st);
    case 2: return run_ozaki<2>(g, L, w, sticky, st);
    case 3: return run_ozaki<3>(g, L, w, sticky, st);
    case 4: return run_ozaki<4>(g, L, w, sticky, st);
    case 5: return run_ozaki<5>(g, L, w, sticky, st);
    case 6: return run_ozaki<6>(g, L, w, sticky, st);
    case 7: return run_ozaki<7>(g, L, w, sticky, st);
    case 8: return run_ozaki<8>(g, L, w, sticky, st);
    default: return hipErrorInvalidValue;
  }
}

// Block width of the 3xBF16 kernel (and row block of its op(B) planes): 256
// (sixteen waves, one workgroup per CU, a quarter fewer plane bytes per product
// than 128) where that grid holds at least two tiles per CU, else 128 (eight
// waves, two workgroups per CU).  At 8192^3: 277 -> 305 TF/s with 3 products,
// 197 -> 207 with 6; 4096 x 2048 x 4096 with 6: 165 -> 171; 2048^3 (128 tiles
// of 128 x 256): 149 -> 119, hence the 128-wide tile there.
int x3_bn(int64_t m, int64_t n) {
  const int64_t t256 = ((m + QBM - 1) / QBM) * ((n + 255) / 256);
  return t256 >= 2 * (int64_t)std::max(runtime_compute_units(), 1) ? 256 : 128;
}

size_t bf16x3_workspace_bytes(int64_t m, int64_t n, int64_t k) {
  return make_layout(m, n, k, 3, 2, x3_bn(m, n)).total;
}
size_t bf16x3_flag_offset(int64_t m, int64_t n, int64_t k) { return make_layout(m, n, k, 3, 2, x3_bn(m, n)).offFlag; }

template <int BN>
void launch_x3_pairs(const PairArgs& p, int nprod, hipStream_t st) {
  const dim3 grid(p.tiles_m * p.tiles_n);
  const dim3 block(64 * (QBM / (32 * X3_WB)) * (BN / 32));
  if (nprod == 3) hipLaunchKernelGGL((slice_pair_gemm_kernel<1, 3, 3, X3_WB, BN>), grid, block, 0, st, p);
  else if (nprod == 6) hipLaunchKernelGGL((slice_pair_gemm_kernel<1, 3, 4, X3_WB, BN>), grid, block, 0, st, p);
  else hipLaunchKernelGGL((slice_pair_gemm_kernel<1, 3, 6, X3_WB, BN>), grid, block, 0, st, p);
}

hipError_t launch_sgemm_bf16x3(const GemmArgs& g, int nprod, void* ws, int* sticky, hipStream_t st) {
  const int bn = x3_bn(g.m, g.n);
  const Layout L = make_layout(g.m, g.n, g.k, 3, 2, bn);
  char* w = static_cast<char*>(ws);
  auto* pA = reinterpret_cast<uint16_t*>(w + L.offA);
  auto* pB = reinterpret_cast<uint16_t*>(w + L.offB);
  int* flag = reinterpret_cast<int*>(w + L.offFlag);
  int64_t sr, sk;
  timing_begin(st, "bf16x3_split");
  clear_call_flag(flag, st);
  strides_a(g.ta, g.lda, sr, sk);
  hipLaunchKernelGGL(split_f32_kernel, dim3((unsigned)(L.Mpad / 64), (unsigned)(L.Kpad / 64)), dim3(256), 0, st,
                     static_cast<const float*>(g.A), sr, sk, g.m, g.k, pA, L.Mpad * L.Kpad, L.Kpad, QBM, flag, sticky);
  strides_b(g.tb, g.ldb, sr, sk);
  hipLaunchKernelGGL(split_f32_kernel, dim3((unsigned)(L.Npad / 64), (unsigned)(L.Kpad / 64)), dim3(256), 0, st,
                     static_cast<const float*>(g.B), sr, sk, g.n, g.k, pB, L.Npad * L.Kpad, L.Kpad, bn, flag, sticky);
  timing_end(st, "bf16x3_split");
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const PairArgs p = pair_args(g, L, pA, pB, nullptr, nullptr, 2, bn);
  timing_begin(st, "bf16x3_gemm");
  if (bn == 256) launch_x3_pairs<256>(p, nprod, st);
  else launch_x3_pairs<128>(p, nprod, st);
  timing_end(st, "bf16x3_gemm");
  return hipGetLastError();
}

}  // namespace jb
