// k_dgemm.hip -- FP64 GEMM kernels for gfx950.
//
// (1) dgemm_mfma: C = alpha*op(A)*op(B) + beta*C on v_mfma_f64_16x16x4f64.
//     Replaces jmul!'s register-tiled AVX loop (src/gemm.jl:244-348): the CPU
//     register tile (pick_kernel_size, src/kernel_structure.jl:76-99) becomes a
//     64x64 wave tile of 4x4 MFMA blocks, and the software prefetch
//     (src/memory_management.jl:181-277) becomes a double-buffered LDS stage
//     fed by coalesced 16-byte global loads.  op(A)/op(B) transposition is
//     absorbed by the loader's LDS write pattern.
// (2) dgemm_recurrence: one fma chain per element in increasing k, the exact
//     order of src/gemm.jl:86 (first product) and :165 (fma k-step), for
//     bit-level parity with the reference recurrence.
//
// LDS holds each operand tile k-major ([k][row], rows contiguous) so that an
// MFMA fragment read is 16 consecutive doubles per 16-lane group.
#include "jb_internal.h"
#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace jb {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

// 16-byte vector of T.
template <typename T> struct Vec16;
template <> struct Vec16<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct Vec16<float> { typedef float type __attribute__((ext_vector_type(4))); };

constexpr int NT = 256;  // threads per block (4 waves)
constexpr int BK = 16;   // k extent of one LDS stage

// Column XOR of k-row k in a swizzled k-major LDS tile: flips bits 2-4 of the
// row index, so a 16-row fragment read stays 16 consecutive elements.
__device__ __forceinline__ int kswz(int k) { return 4 * ((k >> 1) & 7); }

// Loads one R x BK operand tile (R rows of op(A) or columns of op(B)) into
// registers, then writes it k-major into LDS.
//   MN_CONTIG: element (r, kk) at P[(r0+r) + (k0+kk)*ld]  (A 'N', B 'T')
//   otherwise: element (r, kk) at P[(k0+kk) + (r0+r)*ld]  (A 'T', B 'N')
// VEC: 16-byte loads are legal (base 16B aligned, ld a multiple of 16/sizeof(T)).
// NEGZ: elements outside the operand read as -0.0 instead of +0.0 (see the
// FP64 MFMA kernel: a -0 padding product leaves every chain value unchanged).
template <typename T, int R, bool MN_CONTIG, bool VEC, bool NEGZ = false>
struct TileLoader {
  static constexpr T PADV = NEGZ ? T(-0.0) : T(0);
  static constexpr int VW = 16 / sizeof(T);          // elements per 16-byte vector
  static constexpr int PER = R * BK / (VW * NT);     // vectors per thread
  static_assert(PER >= 1 && PER * VW * NT == R * BK, "tile must split into whole 16-byte vectors per thread");
  typedef typename Vec16<T>::type V;
  V v[PER];

  // full: the whole R x BK tile lies inside the operand (block-uniform); one
  // uniform branch picks the unguarded or the guarded loads.
  __device__ __forceinline__ void load(const T* __restrict__ P, int64_t ld, int64_t r0, int64_t k0, int64_t Rtot,
                                       int64_t K, bool full) {
    if (full)
      load_tile<true>(P, ld, r0, k0, Rtot, K);
    else
      load_tile<false>(P, ld, r0, k0, Rtot, K);
  }

  template <bool FULL>
  __device__ __forceinline__ void load_tile(const T* __restrict__ P, int64_t ld, int64_t r0, int64_t k0,
                                            int64_t Rtot, int64_t K) {
    const int t = threadIdx.x;
    if (MN_CONTIG) {
      constexpr int RV = R / VW;  // vectors per k column
      const int r = VW * (t % RV), kb = t / RV;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int kk = kb + (NT / RV) * i;
        const T* src = P + (r0 + r) + (k0 + kk) * ld;
        if (FULL && VEC) {
          v[i] = *reinterpret_cast<const V*>(src);
        } else if (FULL) {
#pragma unroll
          for (int e = 0; e < VW; ++e) v[i][e] = src[e];
        } else {
          const bool kok = (k0 + kk) < K;
#pragma unroll
          for (int e = 0; e < VW; ++e) v[i][e] = (kok && r0 + r + e < Rtot) ? src[e] : PADV;
        }
      }
    } else {
      constexpr int KV = BK / VW;  // vectors per row
      const int kk = VW * (t % KV), rb = t / KV;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int r = rb + (NT / KV) * i;
        const T* src = P + (k0 + kk) + (r0 + r) * ld;
        if (FULL && VEC) {
          v[i] = *reinterpret_cast<const V*>(src);
        } else if (FULL) {
#pragma unroll
          for (int e = 0; e < VW; ++e) v[i][e] = src[e];
        } else {
          const bool rok = (r0 + r) < Rtot;
#pragma unroll
          for (int e = 0; e < VW; ++e) v[i][e] = (rok && k0 + kk + e < K) ? src[e] : PADV;
        }
      }
    }
  }

  // SWZ (k-contiguous loads only): row r of k-row k is stored at column
  // r ^ kswz(k), which spreads the 8 k-rows one store instruction writes
  // over distinct banks; readers apply the same XOR.
  template <int STRIDE, bool SWZ = false>
  __device__ __forceinline__ void store(T* S) const {
    const int t = threadIdx.x;
    if (MN_CONTIG) {
      constexpr int RV = R / VW;
      const int r = VW * (t % RV), kb = t / RV;
#pragma unroll
      for (int i = 0; i < PER; ++i) *reinterpret_cast<V*>(S + (kb + (NT / RV) * i) * STRIDE + r) = v[i];
    } else {
      constexpr int KV = BK / VW;
      const int kk = VW * (t % KV), rb = t / KV;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int r = rb + (NT / KV) * i;
#pragma unroll
        for (int e = 0; e < VW; ++e) S[(kk + e) * STRIDE + (SWZ ? (r ^ kswz(kk + e)) : r)] = v[i][e];
      }
    }
  }
};

// ---------------------------------------------------------------------------
// (1) MFMA GEMM: 128x128 block tile, 4 waves of 64x64 (4x4 MFMA blocks of
//     16x16), or 64x64 with waves of 32x32 when the 128-tile grid would
//     leave the device's last round mostly idle, or (FP64) 64x32 / 32x32 when even
//     the 64-tiles are fewer than four per CU; BK=16 per LDS stage, two
//     LDS stages; global loads are issued one full stage (128-tile) or four
//     stages (64-tile) before their LDS store (the load latency, not the MFMA
//     work, bounded the earlier form that loaded within the stage).  T = double uses
//     v_mfma_f64_16x16x4f64, T = float v_mfma_f32_16x16x4f32.
// ---------------------------------------------------------------------------
#ifndef DG_PAD_BYTES
#define DG_PAD_BYTES 32
#endif

template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t mma(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
  // Row p of the 16x16 result held in item r of a lane of group g = lane/16.
  static __device__ __forceinline__ int prow(int r, int g) { return 4 * r + g; }
};
template <> struct Mfma<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t mma(float x, float y, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int prow(int r, int g) { return 4 * g + r; }
};

// BTM x BTN block tile, 4 waves of (BTM/2) x (BTN/2): 128 x 128, 64 x 64 (for
// problems with too few 128-tiles to fill the device) or 64 x 32 (about one
// 64-tile per CU: two 64 x 32 workgroups per CU then hide each other's
// barrier and LDS latency).
template <typename T, bool TA, bool TB, bool VEC, int BTM, int BTN>
__global__ __launch_bounds__(NT, 2) void gemm_mfma_kernel(int64_t M, int64_t N, int64_t K, T alpha,
                                                          const T* __restrict__ A, int64_t lda,
                                                          const T* __restrict__ B, int64_t ldb, T beta,
                                                          T* __restrict__ C, int64_t ldc, int tiles_m, int tiles_n,
                                                          int acc_from_c) {
  // LDS row stride in elements.  FP64: tile extent + 16 (144, 80 or 48), so
  // the k-rows read by lane groups g and g+1 of a ds_read_b64 start 32 banks
  // apart (no bank conflicts); the k-contiguous operands (A 'T', B 'N') are
  // stored with the row XOR kswz(k), without which their 8-k-row stores would
  // hit the same banks at this stride.  FP32: extent plus DG_PAD_BYTES.
  constexpr bool F64 = sizeof(T) == 8;
  constexpr int STRIDE_A = F64 ? BTM + 16 : BTM + DG_PAD_BYTES / (int)sizeof(T);
  constexpr int STRIDE_B = F64 ? BTN + 16 : BTN + DG_PAD_BYTES / (int)sizeof(T);
  constexpr int WIM = BTM / 32, WIN = BTN / 32;  // 16 x 16 MFMA blocks per wave along m, n
  constexpr bool SWA = F64 && TA, SWB = F64 && !TB;
  __shared__ __attribute__((aligned(16))) T As[2][BK * STRIDE_A];
  __shared__ __attribute__((aligned(16))) T Bs[2][BK * STRIDE_B];
  typedef typename Mfma<T>::acc_t acc_t;

  const int nblocks = tiles_m * tiles_n;
  const int tile = xcd_remap(blockIdx.x, nblocks, 8);
  int tm, tn;
  tile_coords(tile, tiles_m, tiles_n, 8, tm, tn);
  const int64_t m0 = (int64_t)tm * BTM, n0 = (int64_t)tn * BTN;

  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int wm0 = (wave % 2) * (BTM / 2), wn0 = (wave / 2) * (BTN / 2);
  const int g = lane / 16, l16 = lane % 16;

  // PD register sets carry the next PD stages: a stage's loads are issued PD
  // stages before its LDS store.  The smaller tiles run where the grid has
  // about one tile per CU and a stage is only 16 MFMAs per wave or fewer, far
  // shorter than a global load, so they keep 4 stages in flight; the 128-tile
  // keeps 1.
  constexpr int PD = BTM == 128 ? 1 : 4;
  TileLoader<T, BTM, !TA, VEC, F64> la[PD];  // FP64: A's k-padding is -0.0 (below)
  TileLoader<T, BTN, TB, VEC> lb[PD];
  const bool fullA = m0 + BTM <= M, fullB = n0 + BTN <= N;

  // FP64 chains start from -0.0: the first MFMA step fma(a, x, -0.0) is then
  // exactly fl(a * x) for every input, the plain first product of the
  // reference recurrence (src/gemm.jl:86), signed zeros included (a start of
  // +0.0 would turn an all -0 chain into +0).  The k-steps past K that the
  // last 16-k stage pads must leave the chain unchanged too: A pads with -0.0
  // and B with +0.0, so each padding product is -0 and x + (-0) == x for
  // every x, zeros included.  Callers route k == 0 elsewhere.
  // acc_from_c (kernel! semantics, src/kernels.jl:225-233): each chain starts
  // from the element of C instead, every k-step an fma; the caller passes
  // alpha = 1, beta = 0, so the epilogue stores the chain value itself.
  constexpr T acc0 = F64 ? T(-0.0) : T(0);
  acc_t acc[WIM][WIN];
#pragma unroll
  for (int i = 0; i < WIM; ++i)
#pragma unroll
    for (int j = 0; j < WIN; ++j) acc[i][j] = acc_t{acc0, acc0, acc0, acc0};
  if (acc_from_c) {
#pragma unroll
    for (int i = 0; i < WIM; ++i) {
      const int64_t m = m0 + wm0 + 16 * i + l16;
#pragma unroll
      for (int j = 0; j < WIN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t n = n0 + wn0 + 16 * j + Mfma<T>::prow(r, g);
          if (m < M && n < N) acc[i][j][r] = C[m + n * ldc];
        }
    }
  }

  const int nkt = (int)((K + BK - 1) / BK);
  // Stage j travels in register set j % PD: loaded as soon as stage j - PD
  // left that set for LDS, stored to LDS buffer j & 1 near the end of stage
  // j - 1.
  auto load_stage = [&](auto& xa, auto& xb, int kt) {
    const int64_t k0 = (int64_t)kt * BK;
    const bool fk = k0 + BK <= K;
    xa.load(A, lda, m0, k0, M, K, fullA && fk);
    xb.load(B, ldb, n0, k0, N, K, fullB && fk);
  };
  load_stage(la[0], lb[0], 0);
  la[0].template store<STRIDE_A, SWA>(As[0]);
  lb[0].template store<STRIDE_B, SWB>(Bs[0]);
#pragma unroll
  for (int j = 1; j <= PD; ++j)
    if (j < nkt) load_stage(la[j % PD], lb[j % PD], j);
  __syncthreads();

  // MFMA t of a stage covers k = 4t..4t+3; lane group g supplies k = 4t+g.
  // Swizzled column (w0 + 16i + l16) ^ kswz(4t + g): with t < 4, kswz(4t + g) =
  // 16 (t >> 1) + 8 (t & 1) + 4 (g >> 1), and w0 is a multiple of 16, so it is
  // 16 ((w0/16 + i) ^ (t >> 1)) + (l16 ^ (8 (t & 1) + 4 (g >> 1))): a
  // compile-time block offset (per wave) plus one of two per-lane columns.
  static_assert(BK == 16, "swizzle decomposition assumes four MFMA k-groups per stage");
  const int sw0 = l16 ^ (4 * (g >> 1)), sw1 = l16 ^ (8 + 4 * (g >> 1));
  const int bm0 = wm0 / 16, bn0 = wn0 / 16;
  auto frags = [&](const T* as, const T* bs, int t, T (&a)[WIM], T (&b)[WIN]) {
    const int swc = (t & 1) ? sw1 : sw0;
#pragma unroll
    for (int i = 0; i < WIM; ++i)
      a[i] = as[(4 * t + g) * STRIDE_A + (SWA ? 16 * ((bm0 + i) ^ (t >> 1)) + swc : wm0 + 16 * i + l16)];
#pragma unroll
    for (int j = 0; j < WIN; ++j)
      b[j] = bs[(4 * t + g) * STRIDE_B + (SWB ? 16 * ((bn0 + j) ^ (t >> 1)) + swc : wn0 + 16 * j + l16)];
  };

  for (int kt0 = 0; kt0 < nkt; kt0 += PD)
#pragma unroll
  for (int d = 0; d < PD; ++d) {
    const int kt = kt0 + d;
    if (kt >= nkt) break;
    const int nset = (d + 1) % PD;  // register set of stage kt+1 (compile-time)
    const int cur = kt & 1;
    const T* as = As[cur];
    const T* bs = Bs[cur];
    T a[2][WIM], b[2][WIN];
    frags(as, bs, 0, a[0], b[0]);
#pragma unroll
    for (int t = 0; t < BK / 4; ++t) {
      if (t + 1 < BK / 4) frags(as, bs, t + 1, a[(t + 1) & 1], b[(t + 1) & 1]);
      // D'[p][q] = sum_k X[p][k] Y[k][q] with X = op(B)^T block (p -> n) and Y = op(A)^T
      // block (q -> m), so a 16-lane group holds 16 consecutive rows of one column of C
      // (coalesced column-major stores).  The f64 MFMA is a sequential fma chain over
      // its four k values in lane-group order, so each dot product is one fma chain in
      // increasing k, as in the reference recurrence.
#pragma unroll
      for (int i = 0; i < WIM; ++i)
#pragma unroll
        for (int j = 0; j < WIN; ++j) acc[i][j] = Mfma<T>::mma(b[t & 1][j], a[t & 1][i], acc[i][j]);
      // The next stage goes to the other LDS buffer (last read before the
      // previous barrier) one MFMA group before the end, so its writes
      // complete under the last group instead of in front of the barrier;
      // the registers then take the stage after it.
      if (t == BK / 4 - 2 && kt + 1 < nkt) {
        la[nset].template store<STRIDE_A, SWA>(As[cur ^ 1]);
        lb[nset].template store<STRIDE_B, SWB>(Bs[cur ^ 1]);
        if (kt + 1 + PD < nkt) load_stage(la[nset], lb[nset], kt + 1 + PD);
      }
    }
    __syncthreads();
  }

  if (m0 + BTM <= M && n0 + BTN <= N) {
    // Interior tile: for each 16-row strip, the beta*C loads are issued
    // together, then the results stored.
#pragma unroll
    for (int i = 0; i < WIM; ++i) {
      T* Ci = C + (m0 + wm0 + 16 * i + l16);
      T cv[WIN][4];
#pragma unroll
      for (int j = 0; j < WIN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          cv[j][r] = beta != T(0) ? Ci[(n0 + wn0 + 16 * j + Mfma<T>::prow(r, g)) * ldc] : T(0);
#pragma unroll
      for (int j = 0; j < WIN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          T c = alpha * acc[i][j][r];
          if (beta != T(0)) c = fma(beta, cv[j][r], c);
          Ci[(n0 + wn0 + 16 * j + Mfma<T>::prow(r, g)) * ldc] = c;
        }
    }
    return;
  }
#pragma unroll
  for (int i = 0; i < WIM; ++i) {
    const int64_t m = m0 + wm0 + 16 * i + l16;
#pragma unroll
    for (int j = 0; j < WIN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t n = n0 + wn0 + 16 * j + Mfma<T>::prow(r, g);
        if (m < M && n < N) {
          T c = alpha * acc[i][j][r];
          if (beta != T(0)) c = fma(beta, C[m + n * ldc], c);
          C[m + n * ldc] = c;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// (2) Sequential-fma GEMM (bit-exact recurrence, FP64 or FP32): 64x64 block tile,
//     each thread a 4x4 register tile, BK=16 LDS stage.
// ---------------------------------------------------------------------------
constexpr int RB = 64;
// LDS row stride: one 16-byte vector of padding (66 doubles, 68 floats), so rows
// stay 16-byte aligned for the vector stores and start on different banks
template <typename T>
constexpr int rstride() { return RB + 16 / (int)sizeof(T); }

template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(NT, 2) void gemm_recurrence_kernel(int64_t M, int64_t N, int64_t K, T alpha,
                                                                const T* __restrict__ A, int64_t lda,
                                                                const T* __restrict__ B, int64_t ldb, T beta,
                                                                T* __restrict__ C, int64_t ldc, int tiles_m,
                                                                int tiles_n) {
  constexpr int RSTRIDE = rstride<T>();
  __shared__ __attribute__((aligned(16))) T As[BK * RSTRIDE];
  __shared__ __attribute__((aligned(16))) T Bs[BK * RSTRIDE];
  const int nblocks = tiles_m * tiles_n;
  const int tile = xcd_remap(blockIdx.x, nblocks, 8);
  int tm, tn;
  tile_coords(tile, tiles_m, tiles_n, 8, tm, tn);
  const int64_t m0 = (int64_t)tm * RB, n0 = (int64_t)tn * RB;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;

  TileLoader<T, RB, !TA, false> la;
  TileLoader<T, RB, TB, false> lb;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      acc[i][j] = T(0);
    }

  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    la.load(A, lda, m0, k0, M, K, false);
    lb.load(B, ldb, n0, k0, N, K, false);
    __syncthreads();  // previous stage fully consumed
    la.template store<RSTRIDE>(As);
    lb.template store<RSTRIDE>(Bs);
    __syncthreads();
    const int kmax = (int)min<int64_t>(BK, K - k0);
    int kk = 0;
    if (k0 == 0) {  // acc_1 = fl(a_1 * x_1)   (src/gemm.jl:86)
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = a[i] * b[j];
      kk = 1;
    }
    for (; kk < kmax; ++kk) {  // acc_n = fma(a_n, x_n, acc_{n-1})   (src/gemm.jl:165)
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk * RSTRIDE + tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk * RSTRIDE + ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int64_t m = m0 + tx + 16 * i, n = n0 + ty + 16 * j;
      if (m < M && n < N) {
        T c = alpha * acc[i][j];
        if (beta != T(0)) c = fma(beta, C[m + n * ldc], c);
        C[m + n * ldc] = c;
      }
    }
}

template <typename T>
__global__ void scale_kernel(int64_t m, int64_t n, T beta, T* __restrict__ C, int64_t ldc) {
  const int64_t total = m * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx % m, j = idx / m;
    T* p = C + i + j * ldc;
    *p = (beta == T(0)) ? T(0) : beta * *p;
  }
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace

template <typename T, int BTM, int BTN>
static hipError_t launch_gemm_mfma_bt(const GemmArgs& g, hipStream_t s, int acc_from_c) {
  const int tiles_m = (int)((g.m + BTM - 1) / BTM), tiles_n = (int)((g.n + BTN - 1) / BTN);
  const int nblocks = tiles_m * tiles_n;
  constexpr int VW = 16 / sizeof(T);
  const bool vec = aligned16(g.A) && aligned16(g.B) && (g.lda % VW == 0) && (g.ldb % VW == 0);
  auto A = static_cast<const T*>(g.A);
  auto B = static_cast<const T*>(g.B);
  auto C = static_cast<T*>(g.C);
  const T alpha = (T)g.alpha, beta = (T)g.beta;
  const char* tname = sizeof(T) == 8 ? "dgemm_mfma" : "sgemm_mfma";
#define JB_LAUNCH(TA, TB, V)                                                                           \
  hipLaunchKernelGGL((gemm_mfma_kernel<T, TA, TB, V, BTM, BTN>), dim3(nblocks), dim3(NT), 0, s, g.m, g.n, g.k, alpha, \
                     A, g.lda, B, g.ldb, beta, C, g.ldc, tiles_m, tiles_n, acc_from_c)
  timing_begin(s, tname);
  if (vec) {
    if (!g.ta && !g.tb) JB_LAUNCH(false, false, true);
    else if (!g.ta && g.tb) JB_LAUNCH(false, true, true);
    else if (g.ta && !g.tb) JB_LAUNCH(true, false, true);
    else JB_LAUNCH(true, true, true);
  } else {
    if (!g.ta && !g.tb) JB_LAUNCH(false, false, false);
    else if (!g.ta && g.tb) JB_LAUNCH(false, true, false);
    else if (g.ta && !g.tb) JB_LAUNCH(true, false, false);
    else JB_LAUNCH(true, true, false);
  }
#undef JB_LAUNCH
  timing_end(s, tname);
  return hipGetLastError();
}

// Tile choice.  128-tiles run two blocks per CU (512 on the device at a time),
// 64-tiles at least four and ~5 % below the 128-tile rate when both fill the
// device; FP64 problems too small for four 64-tiles per CU take 64 x 32.  The
// 128-tile grid is used when its last round of resident workgroups (512 for
// FP64 on 256 CUs) is at least 90 % full (measured at n = 1536..8192:
// 64-tiles win at 1536-3584 and 5120, 128-tiles at 4096, 6144, 8192).
//
// Round size of the 128-tile grid: the two workgroups per CU that the launch
// bound guarantees, times the device's CU count.  (The occupancy query gives
// FP32 three per CU; rounds of that size moved FP32 8192^3 to the 64-tile and
// lost 6 %, so the rule stays on the guaranteed residency it was measured with.)
static int64_t resident_128() { return 2 * (int64_t)std::max(runtime_compute_units(), 1); }

// The 128-tile unless its last round of resident workgroups would be under
// 90 % full; then the 64-tile (four times the tiles) fills the device better.
template <typename T>
static hipError_t launch_gemm_mfma(const GemmArgs& g, hipStream_t s, int acc_from_c = 0) {
  const int64_t tiles128 = ((g.m + 127) / 128) * ((g.n + 127) / 128);
  const int64_t res = resident_128();
  const int64_t slots = ((tiles128 + res - 1) / res) * res;
  // testing: pin one variant ("128", "64", "64x32", "32")
  const char* force = getenv("JB_GEMM_TILE");
  if (force && *force) {
    if (!strcmp(force, "128")) return launch_gemm_mfma_bt<T, 128, 128>(g, s, acc_from_c);
    if (!strcmp(force, "64")) return launch_gemm_mfma_bt<T, 64, 64>(g, s, acc_from_c);
    if constexpr (sizeof(T) == 8) {  // FP32 64 x 32 and 32 x 32 would leave threads without a whole vector
      if (!strcmp(force, "64x32")) return launch_gemm_mfma_bt<T, 64, 32>(g, s, acc_from_c);
      if (!strcmp(force, "32")) return launch_gemm_mfma_bt<T, 32, 32>(g, s, acc_from_c);
    }
  }
  if ((double)tiles128 >= 0.9 * (double)slots) return launch_gemm_mfma_bt<T, 128, 128>(g, s, acc_from_c);
  // FP64 with fewer than four 64-tiles per CU: 64 x 32 tiles (two per CU hide
  // each other's barrier and LDS waits; measured +40 % at 512^3, +10 % at
  // 1536^3, even at 2048^3 where the 64-tile grid has 4 per CU), and 32 x 32
  // tiles where even the 64 x 32 grid has fewer than four per CU (waves of
  // 16 x 16, four workgroups per CU: +31 % at 512^3, +25 % at 768^3, +4 % at
  // 1024^3 over 64 x 32)
  if constexpr (sizeof(T) == 8) {
    const int64_t cus4 = 4 * (int64_t)std::max(runtime_compute_units(), 1);
    const int64_t tiles64 = ((g.m + 63) / 64) * ((g.n + 63) / 64);
    const int64_t tiles64x32 = ((g.m + 63) / 64) * ((g.n + 31) / 32);
    if (tiles64 < cus4)
      return tiles64x32 < cus4 ? launch_gemm_mfma_bt<T, 32, 32>(g, s, acc_from_c)
                               : launch_gemm_mfma_bt<T, 64, 32>(g, s, acc_from_c);
  }
  return launch_gemm_mfma_bt<T, 64, 64>(g, s, acc_from_c);
}

hipError_t launch_dgemm_mfma(const GemmArgs& g, hipStream_t s) { return launch_gemm_mfma<double>(g, s); }
// kernel!: C = C + A*X with every chain starting from C (g.alpha, g.beta ignored)
hipError_t launch_dgemm_mfma_acc(const GemmArgs& g, hipStream_t s) {
  GemmArgs h = g;
  h.alpha = 1.0;
  h.beta = 0.0;
  return launch_gemm_mfma<double>(h, s, 1);
}
hipError_t launch_sgemm_mfma(const GemmArgs& g, hipStream_t s) { return launch_gemm_mfma<float>(g, s); }

template <typename T>
hipError_t launch_gemm_recurrence(const GemmArgs& g, hipStream_t s, const char* name) {
  const int tiles_m = (int)((g.m + RB - 1) / RB), tiles_n = (int)((g.n + RB - 1) / RB);
  const int nblocks = tiles_m * tiles_n;
  auto A = static_cast<const T*>(g.A);
  auto B = static_cast<const T*>(g.B);
  auto C = static_cast<T*>(g.C);
  const T alpha = (T)g.alpha, beta = (T)g.beta;
#define JB_LAUNCH(TA, TB)                                                                                    \
  hipLaunchKernelGGL((gemm_recurrence_kernel<T, TA, TB>), dim3(nblocks), dim3(NT), 0, s, g.m, g.n, g.k, alpha, \
                     A, g.lda, B, g.ldb, beta, C, g.ldc, tiles_m, tiles_n)
  timing_begin(s, name);
  if (!g.ta && !g.tb) {
    JB_LAUNCH(false, false);
  } else if (!g.ta && g.tb) {
    JB_LAUNCH(false, true);
  } else if (g.ta && !g.tb) {
    JB_LAUNCH(true, false);
  } else {
    JB_LAUNCH(true, true);
  }
#undef JB_LAUNCH
  timing_end(s, name);
  return hipGetLastError();
}

hipError_t launch_dgemm_recurrence(const GemmArgs& g, hipStream_t s) {
  return launch_gemm_recurrence<double>(g, s, "dgemm_recurrence");
}

hipError_t launch_sgemm_recurrence(const GemmArgs& g, hipStream_t s) {
  return launch_gemm_recurrence<float>(g, s, "sgemm_recurrence");
}


hipError_t launch_scale_f64(int64_t m, int64_t n, double beta, double* C, int64_t ldc, hipStream_t s) {
  if (m == 0 || n == 0) return hipSuccess;
  const int64_t total = m * n;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(scale_kernel<double>, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc);
  return hipGetLastError();
}
hipError_t launch_scale_f32(int64_t m, int64_t n, float beta, float* C, int64_t ldc, hipStream_t s) {
  if (m == 0 || n == 0) return hipSuccess;
  const int64_t total = m * n;
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(scale_kernel<float>, dim3(blocks), dim3(256), 0, s, m, n, beta, C, ldc);
  return hipGetLastError();
}

}  // namespace jb
