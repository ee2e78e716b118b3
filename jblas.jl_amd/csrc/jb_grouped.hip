// jb_grouped.hip -- batched / grouped GEMM in one persistent launch, hipGraph
// plans, the autotuner, and the gemm! pullback.
//
// The reference has no batching: its blocking is static (pick_kernel_size,
// src/kernel_structure.jl:76-99; blocking_structure, src/memory_management.jl:
// 78-140, which nothing calls).  Here every problem of a group gets a tile
// shape and a split-K factor from a measured table (autotuner), and all
// problems run in ONE persistent launch:
//   * work items (problem, output tile, k-split) sorted by decreasing cost
//     are pulled from a device counter by a grid of CU-filling workgroups,
//     so uneven problems load-balance;
//   * tile shapes 64x64 and 32x32 (4 waves each) are device functions of
//     one kernel, chosen per work item; the 128-row/column tiles (FP64 and
//     3xBF16) run in kernels of their own, one per tile shape and transpose
//     combination, launched concurrently with it;
//   * split-K partials go to a workspace slot per (tile, split); the last
//     split to finish (per-tile arrival counter) sums them in split order,
//     so results are deterministic;
//   * methods: FP64 on FP64 MFMA, FP32 on FP32 MFMA, FP32 as 3xBF16 (split on
//     the fly into hi/mid/lo BF16 planes in LDS, 6 products on BF16 MFMA).
// A plan uploads descriptors once and captures [counter reset kernel,
// persistent kernels] in a hipGraph; launching the plan replays the graph.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "jb_runtime.h"

namespace jb {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int GNT = 256;  // 4 waves
// (Operand and result pointers come from problem descriptors in memory: every
// access goes through as_global, jb_internal.h.)
// Work-item configurations: 0 = 64x64 MFMA tile, 1 = 32x32 MFMA tile, 2 = the
// whole problem on the vector ALUs (m, n, k <= SMALL_MAX, one split); FP64 and
// 3xBF16 only, run by persistent kernels of their own whose register budget is
// the large tile's (so the others keep four workgroups per CU): 3 = 128x128,
// 4 = 128x64, 5 = 64x128 MFMA tiles.
constexpr int NCFG = 6;
constexpr int CFG_BM[NCFG] = {64, 32, 64, 128, 128, 64};  // tile rows
constexpr int CFG_BN[NCFG] = {64, 32, 64, 128, 64, 128};  // tile columns
constexpr int CFG_SMALL = 2;
constexpr int CFG_LARGE = 3;
constexpr int CFG_TALL = 4;
constexpr int CFG_WIDE = 5;
constexpr int SMALL_MAX = 64;
__host__ __device__ constexpr bool large_cfg(int cfg) { return cfg >= CFG_LARGE; }

// Device descriptor of one problem.
struct GProb {
  const void* A;
  const void* B;
  void* C;
  void* part;         // split-K partials: tiles * splitk slabs of BM*BN elements
  int64_t m, n, k, lda, ldb, ldc;
  double alpha, beta;
  int ta, tb, cfg, splitk;
  int tiles_m, tiles_n;
  int64_t kchunk;     // k per split (multiple of 32)
  int counter_base;   // index of this problem's first per-tile arrival counter
  int pad_;
};
struct GItem {
  int prob, tile, split, pad_;
};

// ---------------------------------------------------------------------------
// operand staging (generic strides, coalesced along the contiguous index)
// ---------------------------------------------------------------------------
// The thread index as an opaque value, read once per work item: offsets derived
// from it are then formed inside the work loop of the persistent kernel rather
// than hoisted out of it and held in registers for all eight tile variants at
// once (which exhausts the register file).
__device__ __forceinline__ int work_item_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// RC: the R index is the contiguous one (stride 1), else the k index is.
// NEGZ: elements outside the operand read as -0.0 (FP64 A operand: a -0
// padding product leaves every chain value unchanged, zeros included).
template <typename T, int R, int BK, bool RC, bool NEGZ = false>
struct GLoader {
  static constexpr int PER = R * BK / GNT;
  T v[PER];
  int tid;  // thread index (see work_item_tid)
  __device__ explicit GLoader(int t) : tid(t) {}
  static __device__ __forceinline__ int row_of(int e) { return RC ? (e % R) : (e / BK); }
  static __device__ __forceinline__ int k_of(int e) { return RC ? (e / R) : (e % BK); }
  __device__ __forceinline__ void load(gptr<const T> P, int64_t ld, int64_t r0, int64_t k0, int64_t Rtot,
                                       int64_t Kend) {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = tid + GNT * i;
      const int64_t gr = r0 + row_of(e), gk = k0 + k_of(e);
      v[i] = (gr < Rtot && gk < Kend) ? (RC ? P[gr + gk * ld] : P[gk + gr * ld]) : (NEGZ ? T(-0.0) : T(0));
    }
  }
  // the same with the vector-loader interface (element loads in every mode)
  static __device__ __forceinline__ int mode(bool, bool, bool) { return 0; }
  __device__ __forceinline__ void load(gptr<const T> P, int64_t ld, int64_t r0, int64_t k0, int64_t Rtot,
                                       int64_t Kend, int) {
    load(P, ld, r0, k0, Rtot, Kend);
  }
  template <int STRIDE, bool SWZ>
  __device__ __forceinline__ void store(T* S) const {
    static_assert(!SWZ, "element staging is not swizzled");
    store_kmajor<STRIDE>(S);
  }
  // k-major [kk][r] (MFMA fp64/fp32 operands)
  template <int STRIDE>
  __device__ __forceinline__ void store_kmajor(T* S) const {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int e = tid + GNT * i;
      S[k_of(e) * STRIDE + row_of(e)] = v[i];
    }
  }
};

// Staging of one R x 16 FP64 operand tile in 16-byte vectors (two doubles
// along the contiguous index), k range [k0, Kend):
//   RC (rows contiguous): vector i of thread t holds rows 2 (t % (R/2)) + {0,1}
//                         of k-row t / (R/2) + (2 GNT / R) i
//   otherwise:            vector i holds k 2 (t % 8) + {0,1} of row t / 8 + 32 i
// Load modes (uniform per work item and stage):
//   LD_FULL  the tile lies inside the operand, 16-byte loads are aligned;
//   LD_VPRED aligned, and the operand's extent along the vector is even, so a
//            vector (even start) is wholly inside or wholly outside: one
//            predicate per vector;
//   LD_ELEM  each element loaded on its own.
// Elements outside the operand read as PADV.
enum { LD_FULL = 0, LD_VPRED = 1, LD_ELEM = 2 };
template <int R, bool RC, bool NEGZ>
struct GVecLoader {
  static constexpr int BK = 16;
  static constexpr int PER = R * BK / (2 * GNT);
  static_assert(PER >= 1 && PER * 2 * GNT == R * BK, "tile must split into whole vectors per thread");
  static constexpr double PADV = NEGZ ? -0.0 : 0.0;
  typedef double V __attribute__((ext_vector_type(2)));
  V v[PER];
  int tid;
  __device__ explicit GVecLoader(int t) : tid(t) {}
  __device__ __forceinline__ int r_of(int i) const { return RC ? 2 * (tid % (R / 2)) : tid / 8 + (2 * GNT / BK) * i; }
  __device__ __forceinline__ int k_of(int i) const { return RC ? tid / (R / 2) + (2 * GNT / R) * i : 2 * (tid % 8); }
  // the mode of a stage: vec = aligned vectors, even = even extent along the vector
  static __device__ __forceinline__ int mode(bool vec, bool even, bool full) {
    return vec && full ? LD_FULL : (vec && even ? LD_VPRED : LD_ELEM);
  }
  // FULLPATH = false: LD_FULL is served by the predicated path (fewer live
  // registers in a kernel at its register limit)
  template <bool FULLPATH = true>
  __device__ __forceinline__ void load(gptr<const double> P, int64_t ld, int64_t r0, int64_t k0, int64_t Rtot,
                                       int64_t Kend, int md) {
    if (FULLPATH && md == LD_FULL) {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int64_t gr = r0 + r_of(i), gk = k0 + k_of(i);
        v[i] = *(gptr<const V>)(RC ? P + gr + gk * ld : P + gk + gr * ld);
      }
    } else if (md != LD_ELEM) {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int64_t gr = r0 + r_of(i), gk = k0 + k_of(i);
        v[i] = (gr < Rtot && gk < Kend) ? *(gptr<const V>)(RC ? P + gr + gk * ld : P + gk + gr * ld)
                                        : V{PADV, PADV};
      }
    } else {
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int64_t gr = r0 + r_of(i), gk = k0 + k_of(i);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int64_t r = RC ? gr + e : gr, k = RC ? gk : gk + e;
          v[i][e] = (r < Rtot && k < Kend) ? (RC ? P[r + k * ld] : P[k + r * ld]) : PADV;
        }
      }
    }
  }
  // k-major [k][row] at row stride STRIDE; SWZ (k-contiguous operands): row r
  // of k-row k sits at column r ^ lswz(k), so the two k-rows of one store
  // instruction's lanes fall in distinct banks.
  template <int STRIDE, bool SWZ>
  __device__ __forceinline__ void store(double* S) const {
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int r = r_of(i), k = k_of(i);
      if (RC) {
        *reinterpret_cast<V*>(S + k * STRIDE + r) = v[i];
      } else {
#pragma unroll
        for (int e = 0; e < 2; ++e) S[(k + e) * STRIDE + (SWZ ? (r ^ lswz(k + e)) : r)] = v[i][e];
      }
    }
  }
  static __device__ __forceinline__ int lswz(int k) { return 4 * ((k >> 1) & 7); }
};

// The vector mode conditions of one operand: `vec` (16-byte aligned base, even
// leading dimension) and `even` (even extent along the vector: rows for RC,
// the k range end otherwise).
template <bool RC>
__device__ __forceinline__ void vec_conditions(const void* P, int64_t ld, int64_t rows, int64_t kend, bool& vec,
                                               bool& even) {
  vec = ((reinterpret_cast<uintptr_t>(P) & 15) == 0) && (ld % 2 == 0);
  even = RC ? rows % 2 == 0 : kend % 2 == 0;
}

// Operand staging of the MFMA tiles: 16-byte vectors for FP64, elements for FP32.
template <typename T, int R, bool RC, bool NEGZ>
struct TileLoaderOf {
  typedef GLoader<T, R, 16, RC, NEGZ> type;
};
template <int R, bool RC, bool NEGZ>
struct TileLoaderOf<double, R, RC, NEGZ> {
  typedef GVecLoader<R, RC, NEGZ> type;
};

typedef __bf16 bf2 __attribute__((ext_vector_type(2)));

template <typename T> struct GMfma;
template <> struct GMfma<double> {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t mma(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int prow(int r, int g) { return 4 * r + g; }
};
template <> struct GMfma<float> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t mma(float x, float y, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int prow(int r, int g) { return 4 * g + r; }
};

// Result element (i, j) of tile (tm, tn): with one split it goes straight to
// C through the alpha/beta epilogue, otherwise to this split's partial slab.
template <typename T, int BM, int BN>
__device__ __forceinline__ void put_result(const GProb& P, int tm, int tn, int split, int i, int j, T v) {
  if (P.splitk == 1) {
    const int64_t m = (int64_t)tm * BM + i, n = (int64_t)tn * BN + j;
    if (m < P.m && n < P.n) {
      gptr<T> C = as_global(static_cast<T*>(P.C));
      T c;
      if (P.k == 0) {  // no products (k == 0 or alpha == 0): C = beta*C, 0 for beta == 0
        c = (T)P.beta != T(0) ? (T)P.beta * C[m + n * P.ldc] : T(0);
      } else {
        c = (T)P.alpha * v;
        if ((T)P.beta != T(0)) c = fma((T)P.beta, C[m + n * P.ldc], c);
      }
      C[m + n * P.ldc] = c;
    }
  } else {
    const int tile = tm * P.tiles_n + tn;
    gptr<T> slab = as_global(static_cast<T*>(P.part)) + ((int64_t)tile * P.splitk + split) * (BM * BN);
    slab[i + j * BM] = v;
  }
}

// After every thread stored its partials: the last split of the tile to
// arrive sums all partials in split order and applies the epilogue.
template <typename T, int BM, int BN, int NTH = GNT>
__device__ void splitk_combine(const GProb& P, int tm, int tn, int* counters, int* sflag) {
  if (P.splitk == 1) return;
  const int tile = tm * P.tiles_n + tn;
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) *sflag = atomicAdd(counters + P.counter_base + tile, 1);
  __syncthreads();
  if (*sflag != P.splitk - 1) return;  // not the last split of this tile
  __threadfence();
  gptr<const T> slab = as_global(static_cast<const T*>(P.part)) + ((int64_t)tile * P.splitk) * (BM * BN);
  gptr<T> C = as_global(static_cast<T*>(P.C));
  const T alpha = (T)P.alpha, beta = (T)P.beta;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  for (int e = threadIdx.x; e < BM * BN; e += NTH) {
    const int i = e % BM, j = e / BM;
    const int64_t m = m0 + i, n = n0 + j;
    if (m < P.m && n < P.n) {
      T s = slab[e];
      for (int q = 1; q < P.splitk; ++q) s += slab[(int64_t)q * (BM * BN) + e];
      T c = alpha * s;
      if (beta != T(0)) c = fma(beta, C[m + n * P.ldc], c);
      C[m + n * P.ldc] = c;
    }
  }
}

// Native MFMA tile (T = double: f64 MFMA, T = float: f32 MFMA).
template <typename T, int BM, int BN, bool TA, bool TB>
__device__ void native_tile(const GProb& P, int tm, int tn, int split, char* smem, int* counters, int* sflag) {
  constexpr int BK = 16;
  constexpr int SA = BM + 16 / (int)sizeof(T) * 2, SB = BN + 16 / (int)sizeof(T) * 2;
  constexpr int IM = BM / 32, JN = BN / 32;  // 16x16 blocks per wave (2 x 2 waves)
  typedef typename GMfma<T>::acc_t acc_t;
  T* As = reinterpret_cast<T*>(smem);
  T* Bs = As + 2 * BK * SA;
  const int tid = work_item_tid();
  const int lane = tid % 64, wave = tid / 64;
  const int wm0 = (wave % 2) * (BM / 2), wn0 = (wave / 2) * (BN / 2);
  const int g = lane / 16, l16 = lane % 16;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t kb = (int64_t)split * P.kchunk;
  const int64_t ke = min(P.k, kb + P.kchunk);
  gptr<const T> A = as_global(static_cast<const T*>(P.A));
  gptr<const T> B = as_global(static_cast<const T*>(P.B));
  // op(A) rows contiguous when A is not transposed, op(B) columns when B is;
  // FP64 operands travel in 16-byte vectors
  typename TileLoaderOf<T, BM, !TA, sizeof(T) == 8>::type la(tid);
  typename TileLoaderOf<T, BN, TB, false>::type lb(tid);
  bool va, ea, vb, eb;
  vec_conditions<!TA>(P.A, P.lda, P.m, ke, va, ea);
  vec_conditions<TB>(P.B, P.ldb, P.n, ke, vb, eb);
  const bool fa = m0 + BM <= P.m, fb = n0 + BN <= P.n;
  auto load_stage = [&](int64_t k0) {
    const bool fk = k0 + BK <= ke;
    la.load(A, P.lda, m0, k0, P.m, ke, la.mode(va, ea, fa && fk));
    lb.load(B, P.ldb, n0, k0, P.n, ke, lb.mode(vb, eb, fb && fk));
  };
  // FP64 chains start from -0.0 and A's k-padding is -0.0: the first step is
  // exactly the plain product and padding products leave the chain unchanged,
  // so each split is the reference recurrence bit for bit, zeros included.
  constexpr T acc0 = sizeof(T) == 8 ? T(-0.0) : T(0);
  acc_t acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[i][j] = acc_t{acc0, acc0, acc0, acc0};
  const int nkt = (int)((ke - kb + BK - 1) / BK);
  __syncthreads();  // LDS may still be read by the previous work item
  load_stage(kb);
  la.template store<SA, false>(As);
  lb.template store<SB, false>(Bs);
  // Stage kt+1 travels in registers during stage kt: loaded right after stage
  // kt went to LDS, stored to the other buffer before stage kt's last MFMA
  // group, so a load has a whole stage to arrive and the LDS writes finish
  // under MFMAs.
  if (nkt > 1) load_stage(kb + BK);
  __syncthreads();
  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    const T* as = As + cur * BK * SA;
    const T* bs = Bs + cur * BK * SB;
#pragma unroll
    for (int t = 0; t < BK / 4; ++t) {
      T a[IM], b[JN];
#pragma unroll
      for (int i = 0; i < IM; ++i) a[i] = as[(4 * t + g) * SA + wm0 + 16 * i + l16];
#pragma unroll
      for (int j = 0; j < JN; ++j) b[j] = bs[(4 * t + g) * SB + wn0 + 16 * j + l16];
#pragma unroll
      for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j) acc[i][j] = GMfma<T>::mma(b[j], a[i], acc[i][j]);
      if (t == BK / 4 - 2 && kt + 1 < nkt) {
        la.template store<SA, false>(As + (cur ^ 1) * BK * SA);
        lb.template store<SB, false>(Bs + (cur ^ 1) * BK * SB);
        if (kt + 2 < nkt) load_stage(kb + (int64_t)(kt + 2) * BK);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        put_result<T, BM, BN>(P, tm, tn, split, wm0 + 16 * i + l16, wn0 + 16 * j + GMfma<T>::prow(r, g), acc[i][j][r]);
  splitk_combine<T, BM, BN>(P, tm, tn, counters, sflag);
}

// BM x BN FP64 tile (128x128, 128x64 or 64x128): four waves of (BM/2)x(BN/2)
// (blocks of v_mfma_f64_16x16x4f64), 16 k per LDS stage, two LDS buffers,
// the next stage's loads one stage ahead in registers.  Same chain semantics
// as native_tile: chains start from -0.0, A's padding is -0.0, each k-split
// is one fma chain in increasing k.
template <int BM, int BN, bool TA, bool TB>
__device__ void large_tile_f64(const GProb& P, int tm, int tn, int split, char* smem, int* counters, int* sflag) {
  // LDS row strides extent + 16 doubles: lane groups g, g+1 of a fragment read start 32 banks apart
  constexpr int BK = 16, STA = BM + 16, STB = BN + 16;
  constexpr int IM = BM / 32, JN = BN / 32;
  constexpr bool SWA = TA, SWB = !TB;  // k-contiguous operands are stored swizzled
  double* const As = reinterpret_cast<double*>(smem);
  double* const Bs = As + 2 * BK * STA;
  const int tid = work_item_tid();
  const int lane = tid % 64, wave = tid / 64;
  const int wm0 = (wave % 2) * (BM / 2), wn0 = (wave / 2) * (BN / 2);
  const int g = lane / 16, l16 = lane % 16;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t kb = (int64_t)split * P.kchunk;
  const int64_t ke = min(P.k, kb + P.kchunk);
  gptr<const double> A = as_global(static_cast<const double*>(P.A));
  gptr<const double> B = as_global(static_cast<const double*>(P.B));
  GVecLoader<BM, !TA, true> la(tid);
  GVecLoader<BN, TB, false> lb(tid);
  // 16-byte loads: aligned bases and even leading dimensions (k0 and the
  // vectors' start rows are even)
  bool va, ea, vb, eb;
  vec_conditions<!TA>(P.A, P.lda, P.m, ke, va, ea);
  vec_conditions<TB>(P.B, P.ldb, P.n, ke, vb, eb);
  const bool fa = m0 + BM <= P.m, fb = n0 + BN <= P.n;
  d4 acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[i][j] = d4{-0.0, -0.0, -0.0, -0.0};
  const int nkt = (int)((ke - kb + BK - 1) / BK);
  auto load_stage = [&](int kt) {
    const int64_t k0 = kb + (int64_t)kt * BK;
    const bool fk = k0 + BK <= ke;
    if (fa && fb && fk && va && vb) {
      la.template load<true>(A, P.lda, m0, k0, P.m, ke, LD_FULL);
      lb.template load<true>(B, P.ldb, n0, k0, P.n, ke, LD_FULL);
    } else {
      la.template load<false>(A, P.lda, m0, k0, P.m, ke, la.mode(va, ea, false));
      lb.template load<false>(B, P.ldb, n0, k0, P.n, ke, lb.mode(vb, eb, false));
    }
  };
  __syncthreads();  // LDS may still be read by the previous work item
  load_stage(0);
  la.template store<STA, SWA>(As);
  lb.template store<STB, SWB>(Bs);
  if (nkt > 1) load_stage(1);
  __syncthreads();
  // Swizzled column of (w0 + 16 i + l16) in k-row 4t + g: 16 ((w0/16 + i) ^ (t >> 1))
  // + (l16 ^ (8 (t & 1) + 4 (g >> 1))) for t < 4 (w0 a multiple of 16).
  const int sw0 = l16 ^ (4 * (g >> 1)), sw1 = l16 ^ (8 + 4 * (g >> 1));
  const int bm0 = wm0 / 16, bn0 = wn0 / 16;
  auto frags = [&](const double* as, const double* bs, int t, double (&a)[IM], double (&b)[JN]) {
    const int swc = (t & 1) ? sw1 : sw0;
#pragma unroll
    for (int i = 0; i < IM; ++i)
      a[i] = as[(4 * t + g) * STA + (SWA ? 16 * ((bm0 + i) ^ (t >> 1)) + swc : wm0 + 16 * i + l16)];
#pragma unroll
    for (int j = 0; j < JN; ++j)
      b[j] = bs[(4 * t + g) * STB + (SWB ? 16 * ((bn0 + j) ^ (t >> 1)) + swc : wn0 + 16 * j + l16)];
  };
  for (int kt = 0; kt < nkt; ++kt) {
    const int cur = kt & 1;
    const double* as = As + cur * BK * STA;
    const double* bs = Bs + cur * BK * STB;
    double a[2][IM], b[2][JN];
    frags(as, bs, 0, a[0], b[0]);
#pragma unroll
    for (int t = 0; t < BK / 4; ++t) {
      if (t + 1 < BK / 4) frags(as, bs, t + 1, a[(t + 1) & 1], b[(t + 1) & 1]);
#pragma unroll
      for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[t & 1][j], a[t & 1][i], acc[i][j], 0, 0, 0);
      if (t == BK / 4 - 2 && kt + 1 < nkt) {
        la.template store<STA, SWA>(As + (cur ^ 1) * BK * STA);
        lb.template store<STB, SWB>(Bs + (cur ^ 1) * BK * STB);
        if (kt + 2 < nkt) load_stage(kt + 2);
      }
    }
    __syncthreads();
  }
  // results: a 16-lane group holds 16 consecutive rows of one column
  if (P.splitk == 1) {
    gptr<double> C = as_global(static_cast<double*>(P.C));
    const double alpha = P.alpha, beta = P.beta;
#pragma unroll
    for (int i = 0; i < IM; ++i) {
      const int64_t m = m0 + wm0 + 16 * i + l16;
      if (m >= P.m) continue;
#pragma unroll
      for (int j = 0; j < JN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int64_t n = n0 + wn0 + 16 * j + GMfma<double>::prow(r, g);
          if (n < P.n) {  // as put_result: k == 0 (or alpha == 0) gives beta*C, +0 for beta == 0
            gptr<double> c = C + m + n * P.ldc;
            double v;
            if (P.k == 0) v = beta != 0.0 ? beta * *c : 0.0;
            else v = beta != 0.0 ? fma(beta, *c, alpha * acc[i][j][r]) : alpha * acc[i][j][r];
            *c = v;
          }
        }
    }
  } else {
    gptr<double> slab =
        as_global(static_cast<double*>(P.part)) + ((int64_t)(tm * P.tiles_n + tn) * P.splitk + split) * (BM * BN);
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          slab[(wm0 + 16 * i + l16) + (wn0 + 16 * j + GMfma<double>::prow(r, g)) * BM] = acc[i][j][r];
    splitk_combine<double, BM, BN>(P, tm, tn, counters, sflag);
  }
}

// LDS rows of the 3xBF16 tile hold BK = 32 BF16 (64 bytes, four 16-byte chunks,
// no padding); chunk c of row r is stored at chunk c ^ ((r >> 2) & 3), so the 16
// rows that one 16-byte fragment read covers (same c) hit 16 distinct 4-bank
// groups of the 64-bank LDS.
__device__ __forceinline__ int bf_chunk(int r, int c) { return c ^ ((r >> 2) & 3); }

// FP32 operand staging of the 3xBF16 tile in runs of four consecutive k of one
// row (R x BK tile, NR runs per thread, NTH threads), split into hi/mid/lo BF16
// planes ([plane][row][BK], chunks swizzled) with one 8-byte LDS store per plane
// and run.
//   RC (rows contiguous): run q of thread t is row t % R, k-run t / R + (NTH / R) q
//   otherwise:            row t / (BK / 4) + (4 NTH / BK) q, k-run t % (BK / 4)
// Lanes with consecutive t take consecutive addresses of the contiguous index.
template <int R, int BK, bool RC, int NTH = GNT>
struct GRun4 {
  static constexpr int NR = R * BK / (4 * NTH);
  static_assert(NR >= 1 && R * BK == 4 * NTH * NR, "tile must split into whole runs per thread");
  static_assert(RC ? (NTH % R == 0 && (NTH / R) * NR == BK / 4) : ((4 * NTH / BK) * NR == R), "run map");
  typedef float V4 __attribute__((ext_vector_type(4)));
  float v[NR][4];
  int tid;
  __device__ explicit GRun4(int t) : tid(t) {}
  __device__ __forceinline__ int row(int q) const { return RC ? tid % R : tid / (BK / 4) + (4 * NTH / BK) * q; }
  __device__ __forceinline__ int krun(int q) const { return RC ? tid / R + (NTH / R) * q : tid % (BK / 4); }
  __device__ __forceinline__ void load(gptr<const float> P, int64_t ld, int64_t r0, int64_t k0, int64_t Rtot,
                                       int64_t Kend) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const int64_t gr = r0 + row(q), gk = k0 + 4 * krun(q);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        v[q][e] = (gr < Rtot && gk + e < Kend) ? (RC ? P[gr + (gk + e) * ld] : P[gk + e + gr * ld]) : 0.0f;
    }
  }
  // Stages whose k range lies inside the operand: the lane part of each run's
  // address is formed once per work item (lane_base), with the row clamped
  // into the operand (a row past its end reads the last row and is zeroed);
  // a stage then adds only the wave-uniform k0 (times ld for rows-contiguous
  // operands), so a load costs one 64-bit add instead of a guarded 64-bit
  // multiply-add chain per element.  A k-contiguous run is one 16-byte load
  // when `vec` (16-byte aligned base, ld a multiple of 4).
  gptr<const float> lp[NR];
  bool rok[NR];
  __device__ __forceinline__ void lane_base(gptr<const float> P, int64_t ld, int64_t r0, int64_t Rtot) {
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const int64_t r = r0 + row(q);
      rok[q] = r < Rtot;
      const int64_t rc = rok[q] ? r : Rtot - 1;
      lp[q] = RC ? P + rc + (int64_t)(4 * krun(q)) * ld : P + rc * ld + 4 * krun(q);
    }
  }
  // rows_in: every row of the tile lies inside the operand (no zeroing)
  __device__ __forceinline__ void load_kfull(int64_t ld, int64_t k0, bool vec, bool rows_in) {
    if (!RC && vec) {
#pragma unroll
      for (int q = 0; q < NR; ++q) {
        const V4 x = *(gptr<const V4>)(lp[q] + k0);
#pragma unroll
        for (int e = 0; e < 4; ++e) v[q][e] = x[e];
      }
    } else {
#pragma unroll
      for (int q = 0; q < NR; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) v[q][e] = RC ? lp[q][(k0 + e) * ld] : lp[q][k0 + e];
    }
    if (!rows_in) {
#pragma unroll
      for (int q = 0; q < NR; ++q)
#pragma unroll
        for (int e = 0; e < 4; ++e) v[q][e] = rok[q] ? v[q][e] : 0.0f;
    }
  }
  // x = hi + mid + lo + rest: each plane is x rounded to BF16 (nearest even,
  // v_cvt_pk_bf16_f32), subtracted exactly before the next plane.
  __device__ __forceinline__ void store3(uint16_t* S) const {
    static_assert(BK == 32, "four 16-byte chunks per LDS row");
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      float x[4] = {v[q][0], v[q][1], v[q][2], v[q][3]};
      const int r = row(q), kr = krun(q);
      uint16_t* dst = S + r * BK + 8 * bf_chunk(r, kr >> 1) + 4 * (kr & 1);
      // |x| within half a BF16 ulp of FLT_MAX rounds to Inf: the first plane
      // rounds x clamped to the largest finite BF16, which for such an x is
      // its truncation (the residual then has at most 16 significant bits and
      // the split stays exact, as in the single-problem split) and for every
      // other finite x changes nothing.  An Inf or NaN x still leaves a NaN
      // residual, so every product it enters is NaN.
      const float BMAX = __uint_as_float(0x7F7F0000u);
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        float y[4] = {x[0], x[1], x[2], x[3]};
        if (p == 0) {
#pragma unroll
          for (int e = 0; e < 4; ++e) y[e] = __builtin_amdgcn_fmed3f(x[e], -BMAX, BMAX);
        }
        uint32_t w0 = __builtin_bit_cast(uint32_t, bf2{(__bf16)y[0], (__bf16)y[1]});
        uint32_t w1 = __builtin_bit_cast(uint32_t, bf2{(__bf16)y[2], (__bf16)y[3]});
        *reinterpret_cast<uint2*>(dst + p * R * BK) = uint2{w0, w1};
        if (p < 2) {
          // the packed BF16 values back in FP32 (low half shifted up, high half
          // masked): one packed conversion per two elements and plane.  Written
          // as (float) of each BF16, or with w0, w1 visible to the compiler, it
          // converted each low element a second time on its own; the empty asm
          // hides where w0 and w1 came from.
          asm("" : "+v"(w0), "+v"(w1));
          x[0] -= __uint_as_float(w0 << 16);
          x[1] -= __uint_as_float(w0 & 0xffff0000u);
          x[2] -= __uint_as_float(w1 << 16);
          x[3] -= __uint_as_float(w1 & 0xffff0000u);
        }
      }
    }
  }
};

// 3xBF16 tile for FP32 operands: the staged FP32 tile is split into hi/mid/lo
// BF16 planes in LDS ([plane][row][32 k]); v_mfma_f32_16x16x32_bf16 computes the
// 6 products t+u <= 4, level 2 in one accumulator and the rest in another.
template <int BM, int BN, bool TA, bool TB>
__device__ void bf16x3_tile(const GProb& P, int tm, int tn, int split, char* smem, int* counters, int* sflag) {
  constexpr int BK = 32;
  constexpr int RS = BK;  // bf16 per LDS row
  constexpr int IM = BM / 32, JN = BN / 32;
  constexpr int BUF = 3 * (BM + BN) * RS;  // one stage: A planes [3][BM][RS], then B planes [3][BN][RS]
  uint16_t* const S0 = reinterpret_cast<uint16_t*>(smem);
  const int tid = work_item_tid();
  const int lane = tid % 64, wave = tid / 64;
  const int wm0 = (wave % 2) * (BM / 2), wn0 = (wave / 2) * (BN / 2);
  const int g = lane / 16, l16 = lane % 16;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t kb = (int64_t)split * P.kchunk;
  const int64_t ke = min(P.k, kb + P.kchunk);
  gptr<const float> A = as_global(static_cast<const float*>(P.A));
  gptr<const float> B = as_global(static_cast<const float*>(P.B));
  GRun4<BM, BK, !TA> la(tid);
  GRun4<BN, BK, TB> lb(tid);
  f4 acc[2][IM][JN];
#pragma unroll
  for (int l = 0; l < 2; ++l)
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j) acc[l][i][j] = f4{0, 0, 0, 0};
  const int nkt = (int)((ke - kb + BK - 1) / BK);
  // Two LDS stages.  Stage kt+1 is split into the other buffer after stage
  // kt's products are issued, and stage kt+2 is then loaded into the freed
  // registers, so a load has a whole stage to arrive; one barrier per stage.
  if (nkt > 0) {
    __syncthreads();  // LDS may still be read by the previous work item
    la.load(A, P.lda, m0, kb, P.m, ke);
    lb.load(B, P.ldb, n0, kb, P.n, ke);
    la.store3(S0);
    lb.store3(S0 + 3 * BM * RS);
    if (nkt > 1) {
      la.load(A, P.lda, m0, kb + BK, P.m, ke);
      lb.load(B, P.ldb, n0, kb + BK, P.n, ke);
    }
    __syncthreads();
  }
  for (int kt = 0; kt < nkt; ++kt) {
    const uint16_t* As = S0 + (kt & 1) * BUF;
    const uint16_t* Bs = As + 3 * BM * RS;
    bf8 a[3][IM], b[3][JN];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
#pragma unroll
      for (int i = 0; i < IM; ++i)
        a[p][i] = __builtin_bit_cast(bf8, *reinterpret_cast<const u4*>(As + (p * BM + wm0 + 16 * i + l16) * RS +
                                                                          8 * bf_chunk(l16, g)));
#pragma unroll
      for (int j = 0; j < JN; ++j)
        b[p][j] = __builtin_bit_cast(bf8, *reinterpret_cast<const u4*>(Bs + (p * BN + wn0 + 16 * j + l16) * RS +
                                                                          8 * bf_chunk(l16, g)));
    }
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j) {
        // X = op(B) plane (p -> n), Y = op(A) plane (q -> m)
        acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[0][j], a[0][i], acc[0][i][j], 0, 0, 0);
        acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[1][j], a[0][i], acc[1][i][j], 0, 0, 0);
        acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[0][j], a[1][i], acc[1][i][j], 0, 0, 0);
        acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[2][j], a[0][i], acc[1][i][j], 0, 0, 0);
        acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[1][j], a[1][i], acc[1][i][j], 0, 0, 0);
        acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[0][j], a[2][i], acc[1][i][j], 0, 0, 0);
      }
    if (kt + 1 < nkt) {
      uint16_t* nxt = S0 + ((kt + 1) & 1) * BUF;  // last read in stage kt-1, before the previous barrier
      la.store3(nxt);
      lb.store3(nxt + 3 * BM * RS);
      if (kt + 2 < nkt) {
        la.load(A, P.lda, m0, kb + (int64_t)(kt + 2) * BK, P.m, ke);
        lb.load(B, P.ldb, n0, kb + (int64_t)(kt + 2) * BK, P.n, ke);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        put_result<float, BM, BN>(P, tm, tn, split, wm0 + 16 * i + l16, wn0 + 16 * j + 4 * g + r,
                                  acc[0][i][j][r] + acc[1][i][j][r]);
  splitk_combine<float, BM, BN>(P, tm, tn, counters, sflag);
}

// 3xBF16 tile of the 128-row/column shapes (128x128, 128x64, 64x128): WM x WN
// waves of (BM/WM) x (BN/WN), each a grid of 32x32 blocks of
// v_mfma_f32_32x32x16_bf16 (one 16-byte fragment per lane feeds twice the
// products of a 16x16x32 fragment, so the LDS reads per product halve).  The
// staging and the LDS image are those of bf16x3_tile ([plane][row][32 k],
// chunks swizzled by bf_chunk); a 32x32x16 fragment of k-step h is chunk
// 2h + lane/32 of row lane%32.  Same 6 products and two accumulators.  The
// in-register split costs VALU work per staged element, so the larger tiles
// (fewer staged elements per product) run more MFMAs per VALU instruction.
template <int BM, int BN, int WM, int WN, bool TA, bool TB>
__device__ void x3_large_tile(const GProb& P, int tm, int tn, int split, char* smem, int* counters, int* sflag) {
  constexpr int NTH = 64 * WM * WN;
  constexpr int BK = 32;
  constexpr int RS = BK;
  constexpr int IB = BM / WM / 32, JB = BN / WN / 32;  // 32x32 blocks per wave along m, n
  static_assert(IB >= 1 && JB >= 1 && IB * WM * 32 == BM && JB * WN * 32 == BN, "wave grid");
  constexpr int BUF = 3 * (BM + BN) * RS;
  typedef float f16v __attribute__((ext_vector_type(16)));
  uint16_t* const S0 = reinterpret_cast<uint16_t*>(smem);
  const int tid = work_item_tid();
  const int lane = tid % 64, wave = tid / 64;
  const int wm0 = (wave % WM) * (BM / WM), wn0 = (wave / WM) * (BN / WN);
  const int l32 = lane % 32, hl = lane / 32;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t kb = (int64_t)split * P.kchunk;
  const int64_t ke = min(P.k, kb + P.kchunk);
  gptr<const float> A = as_global(static_cast<const float*>(P.A));
  gptr<const float> B = as_global(static_cast<const float*>(P.B));
  GRun4<BM, BK, !TA, NTH> la(tid);
  GRun4<BN, BK, TB, NTH> lb(tid);
  // stages inside the operand along k take load_kfull: k-contiguous operands in
  // 16-byte vectors where the base is 16-byte aligned and ld a multiple of 4
  const bool fa = m0 + BM <= P.m, fb = n0 + BN <= P.n;
  const bool va = ((reinterpret_cast<uintptr_t>(P.A) & 15) == 0) && P.lda % 4 == 0;
  const bool vb = ((reinterpret_cast<uintptr_t>(P.B) & 15) == 0) && P.ldb % 4 == 0;
  la.lane_base(A, P.lda, m0, P.m);
  lb.lane_base(B, P.ldb, n0, P.n);
  auto load_stage = [&](int64_t k0) {
    if (k0 + BK <= ke) {
      la.load_kfull(P.lda, k0, va, fa);
      lb.load_kfull(P.ldb, k0, vb, fb);
    } else {
      la.load(A, P.lda, m0, k0, P.m, ke);
      lb.load(B, P.ldb, n0, k0, P.n, ke);
    }
  };
  f16v acc[2][IB][JB];
#pragma unroll
  for (int l = 0; l < 2; ++l)
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
      for (int j = 0; j < JB; ++j)
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[l][i][j][e] = 0.0f;
  const int nkt = (int)((ke - kb + BK - 1) / BK);
  // two LDS stages, one barrier per stage, as in bf16x3_tile
  if (nkt > 0) {
    __syncthreads();  // LDS may still be read by the previous work item
    load_stage(kb);
    la.store3(S0);
    lb.store3(S0 + 3 * BM * RS);
    if (nkt > 1) load_stage(kb + BK);
    __syncthreads();
  }
  for (int kt = 0; kt < nkt; ++kt) {
    const uint16_t* As = S0 + (kt & 1) * BUF;
    const uint16_t* Bs = As + 3 * BM * RS;
#pragma unroll
    for (int h = 0; h < BK / 16; ++h) {
      bf8 a[3][IB], b[3][JB];
#pragma unroll
      for (int p = 0; p < 3; ++p) {
#pragma unroll
        for (int i = 0; i < IB; ++i) {
          const int r = wm0 + 32 * i + l32;
          a[p][i] = __builtin_bit_cast(bf8, *reinterpret_cast<const u4*>(As + (p * BM + r) * RS + 8 * bf_chunk(r, 2 * h + hl)));
        }
#pragma unroll
        for (int j = 0; j < JB; ++j) {
          const int r = wn0 + 32 * j + l32;
          b[p][j] = __builtin_bit_cast(bf8, *reinterpret_cast<const u4*>(Bs + (p * BN + r) * RS + 8 * bf_chunk(r, 2 * h + hl)));
        }
      }
#pragma unroll
      for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) {
          // X = op(B) plane (p -> n), Y = op(A) plane (q -> m)
          acc[0][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[0][j], a[0][i], acc[0][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[1][j], a[0][i], acc[1][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[0][j], a[1][i], acc[1][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[2][j], a[0][i], acc[1][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[1][j], a[1][i], acc[1][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b[0][j], a[2][i], acc[1][i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nkt) {
      uint16_t* nxt = S0 + ((kt + 1) & 1) * BUF;  // last read in stage kt-1, before the previous barrier
      la.store3(nxt);
      lb.store3(nxt + 3 * BM * RS);
      if (kt + 2 < nkt) load_stage(kb + (int64_t)(kt + 2) * BK);
    }
    __syncthreads();
  }
  // item e of lane l in block (i, j): row wm0 + 32 i + l%32, column wn0 + 32 j + 8 (e/4) + 4 (l/32) + e%4
#pragma unroll
  for (int i = 0; i < IB; ++i)
#pragma unroll
    for (int j = 0; j < JB; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e)
        put_result<float, BM, BN>(P, tm, tn, split, wm0 + 32 * i + l32, wn0 + 32 * j + 8 * (e / 4) + 4 * hl + e % 4,
                                  acc[0][i][j][e] + acc[1][i][j][e]);
  splitk_combine<float, BM, BN, NTH>(P, tm, tn, counters, sflag);
}

// Small problem (m, n, k <= 64) in one work item, the form of fastmul!
// (src/kernels.jl:202-208, kernel_quote :42-201): no MFMA tile; lane = row of C
// (rows >= m recompute row m-1 and store nothing), wave w holds columns
// 16w..16w+15 in registers.  The workgroup first stages all of op(B) (k x n)
// in LDS, [p][j] with rows of SB_STRIDE elements, in two batches of loads;
// each lane then reads its row of op(A) in register chunks of 4 k (the next
// chunk in flight) and B(p, j0..j0+15) as LDS broadcasts, 8 columns at a time
// (registers stay within the budget of the MFMA tiles sharing this kernel).  Each element is the chain
// acc = A(i,0) B(0,j), acc = fma(A(i,p), B(p,j), acc) for p = 1..k-1 -- the
// plain first product of src/gemm.jl:86 and the k-steps of :165 -- then the
// alpha/beta epilogue of the tiles (put_result).
template <typename T, bool TA, bool TB>
__device__ __forceinline__ void small_tile(const GProb& P, char* smem) {
  constexpr int SC = 16;                          // columns per wave
  constexpr int SB_STRIDE = 64 + 16 / sizeof(T);  // 16-byte aligned rows, 4-way at most on the staging writes
  constexpr int AC = 4;                           // k per register chunk of A
  T* Bs = reinterpret_cast<T*>(smem);
  const int tid = work_item_tid();
  const int lane = tid % 64, wave = tid / 64;
  const int64_t m = P.m, n = P.n, k = P.k;
  gptr<const T> A = as_global(static_cast<const T*>(P.A));
  gptr<const T> B = as_global(static_cast<const T*>(P.B));
  // stage op(B): element e = tid + 256 q is (p, j) with the contiguous index fastest
  constexpr int NQ = 64 * 64 / GNT;  // 16 elements per thread, staged 8 at a time
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    T v[NQ / 2];
#pragma unroll
    for (int q = 0; q < NQ / 2; ++q) {
      const int e = tid + GNT * (q + h * NQ / 2);
      const int pp = TB ? e / 64 : e % 64, jj = TB ? e % 64 : e / 64;
      v[q] = (pp < k && jj < n) ? (TB ? B[jj + (int64_t)pp * P.ldb] : B[pp + (int64_t)jj * P.ldb]) : T(0);
    }
#pragma unroll
    for (int q = 0; q < NQ / 2; ++q) {
      const int e = tid + GNT * (q + h * NQ / 2);
      const int pp = TB ? e / 64 : e % 64, jj = TB ? e % 64 : e / 64;
      Bs[pp * SB_STRIDE + jj] = v[q];
    }
  }
  const int64_t i = lane < m ? lane : m - 1;
  gptr<const T> arow = TA ? A + i * P.lda : A + i;  // A(i, p) = arow[p * astep]
  const int64_t astep = TA ? 1 : P.lda;
  T a[2][AC];
#pragma unroll
  for (int u = 0; u < AC; ++u) a[0][u] = u < k ? arow[u * astep] : T(0);
  __syncthreads();
  const int j0 = SC * wave;
  if (j0 >= n || k == 0) {
    if (k == 0 && j0 < n && lane < m) {  // alpha == 0 or k == 0: C = beta*C, 0 for beta == 0
      gptr<T> C = as_global(static_cast<T*>(P.C));
#pragma unroll
      for (int c = 0; c < SC; ++c)
        if (j0 + c < n) {
          gptr<T> cp = C + i + (j0 + c) * P.ldc;
          *cp = (T)P.beta != T(0) ? (T)P.beta * *cp : T(0);
        }
    }
    return;  // no barrier below
  }
  const T* bp = Bs + j0;
  T acc[SC];
#pragma unroll
  for (int c0 = 0; c0 < SC; c0 += 8) {  // p = 0: the plain first product
    T b[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) b[c] = bp[c0 + c];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c0 + c] = a[0][0] * b[c];
  }
  // chunks of AC k, two per iteration so that the register buffer of a chunk
  // (cb & 1) is a compile-time index; chunk cb+1 is loaded while cb computes
#pragma unroll 1
  for (int ch = 0; ch * AC < k; ch += 2) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {  // guarded bodies, no break: the loops unroll fully
      const int cb = ch + h;
      if (cb * AC < k) {
        if ((cb + 1) * AC < k) {
#pragma unroll
          for (int u = 0; u < AC; ++u) {
            const int64_t pn = (int64_t)(cb + 1) * AC + u;
            a[h ^ 1][u] = pn < k ? arow[pn * astep] : T(0);
          }
        }
#pragma unroll
        for (int u = 0; u < AC; ++u) {
          const int pp = cb * AC + u;
          if (pp > 0 && pp < k) {
            const T* brow = bp + pp * SB_STRIDE;
#pragma unroll
            for (int c0 = 0; c0 < SC; c0 += 8) {
              T b[8];
#pragma unroll
              for (int c = 0; c < 8; ++c) b[c] = brow[c0 + c];
#pragma unroll
              for (int c = 0; c < 8; ++c) acc[c0 + c] = fma(a[h][u], b[c], acc[c0 + c]);
            }
          }
        }
      }
    }
  }
  if (lane >= m) return;
  gptr<T> C = as_global(static_cast<T*>(P.C));
  const T alpha = (T)P.alpha, beta = (T)P.beta;
#pragma unroll
  for (int c = 0; c < SC; ++c) {
    const int64_t j = j0 + c;
    if (j < n) {
      T v = alpha * acc[c];
      if (beta != T(0)) v = fma(beta, C[i + j * P.ldc], v);
      C[i + j * P.ldc] = v;
    }
  }
}

template <int KIND>
__device__ __forceinline__ void run_small(const GProb& P, int combo, char* smem) {
  typedef typename std::conditional<KIND == 0, double, float>::type T;
  switch (combo) {
    case 0: small_tile<T, false, false>(P, smem); break;
    case 1: small_tile<T, false, true>(P, smem); break;
    case 2: small_tile<T, true, false>(P, smem); break;
    default: small_tile<T, true, true>(P, smem); break;
  }
}

template <int KIND, int BM, bool TA, bool TB>
__device__ __forceinline__ void run_tile(const GProb& P, int tm, int tn, int split, char* smem, int* counters,
                                         int* sflag) {
  if (KIND == 0) native_tile<double, BM, BM, TA, TB>(P, tm, tn, split, smem, counters, sflag);
  else if (KIND == 1) native_tile<float, BM, BM, TA, TB>(P, tm, tn, split, smem, counters, sflag);
  else bf16x3_tile<BM, BM, TA, TB>(P, tm, tn, split, smem, counters, sflag);
}

template <int KIND, int BM>
__device__ __forceinline__ void run_combo(const GProb& P, int combo, int tm, int tn, int split, char* smem,
                                          int* counters, int* sflag) {
  switch (combo) {
    case 0: run_tile<KIND, BM, false, false>(P, tm, tn, split, smem, counters, sflag); break;
    case 1: run_tile<KIND, BM, false, true>(P, tm, tn, split, smem, counters, sflag); break;
    case 2: run_tile<KIND, BM, true, false>(P, tm, tn, split, smem, counters, sflag); break;
    default: run_tile<KIND, BM, true, true>(P, tm, tn, split, smem, counters, sflag); break;
  }
}

template <int KIND>
__device__ __forceinline__ void dispatch_tile(const GProb& P, int combo, int tm, int tn, int split, char* smem,
                                              int* counters, int* sflag) {
  if (P.cfg == 0) run_combo<KIND, 64>(P, combo, tm, tn, split, smem, counters, sflag);
  else if (P.cfg == 1) run_combo<KIND, 32>(P, combo, tm, tn, split, smem, counters, sflag);
  else run_small<KIND>(P, combo, smem);
}

// Work queues of one persistent launch, one per XCD: queue x holds items
// [qdesc[x], qdesc[x] + qdesc[NXCD + x]) of the launch's item array and is
// drained through counter qnext[x].  Workgroups are dispatched to the XCDs
// round-robin, so workgroup b pops from queue b % NXCD first: the tiles that
// share operand panels were put in one queue and meet in one XCD's L2.  Once
// its queue is drained a workgroup takes items from the others in turn.
// Returns the item index, or -1 after every queue was found drained (each
// workgroup then leaves; an item is taken exactly once, by the atomic on its
// queue's counter).
constexpr int NXCD = 8;
__device__ __forceinline__ int pop_item(const int* __restrict__ qdesc, int* __restrict__ qnext, int& q, int& tries,
                                        int* s_item) {
  for (;;) {
    if (threadIdx.x == 0) *s_item = atomicAdd(qnext + q, 1);
    __syncthreads();
    // uniform across the wave: descriptor loads become scalar loads into SGPRs
    const int it = __builtin_amdgcn_readfirstlane(*s_item);
    __syncthreads();
    if (it < qdesc[NXCD + q]) return qdesc[q] + it;
    if (++tries == NXCD) return -1;
    q = (q + 1) % NXCD;
  }
}

// KIND 0: FP64 native, 1: FP32 native, 2: FP32 3xBF16.
template <int KIND>
__global__ __launch_bounds__(GNT, 2) void grouped_kernel(const GProb* __restrict__ probs, const GItem* __restrict__ items,
                                                         const int* __restrict__ qdesc, int* __restrict__ qnext,
                                                         int* __restrict__ counters) {
  constexpr int SMEM = (KIND == 2) ? (2 * 3 * (64 + 64) * 32 * 2) : (2 * 16 * (64 + 8) * 2 * (KIND == 0 ? 8 : 4));
  __shared__ __attribute__((aligned(16))) char smem[SMEM];
  static_assert(SMEM >= 64 * (64 + 16 / (KIND == 0 ? 8 : 4)) * (KIND == 0 ? 8 : 4), "small path stages op(B) in LDS");
  __shared__ int s_item, s_flag;
  int q = blockIdx.x % NXCD, tries = 0;
  for (;;) {
    const int it = pop_item(qdesc, qnext, q, tries, &s_item);
    if (it < 0) break;
    const GItem w = items[it];
    const GProb P = probs[w.prob];
    const int tm = w.tile / P.tiles_n, tn = w.tile % P.tiles_n;
    const int combo = (P.ta ? 2 : 0) + (P.tb ? 1 : 0);
    dispatch_tile<KIND>(P, combo, tm, tn, w.split, smem, counters, &s_flag);
  }
}

// The FP64 items of one 128-row/column tile shape and transpose combination,
// pulled from their own queues (one kernel per shape and combination: a kernel
// holding several tile variants exceeds the register budget of two
// workgroups per CU).
template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(GNT, 2) void grouped_large_kernel(const GProb* __restrict__ probs,
                                                               const GItem* __restrict__ items,
                                                               const int* __restrict__ qdesc, int* __restrict__ qnext,
                                                               int* __restrict__ counters) {
  constexpr int SMEM = 2 * 16 * ((BM + 16) + (BN + 16)) * 8;  // two buffers of A and B: 16 k-rows of stride extent + 16
  __shared__ __attribute__((aligned(16))) char smem[SMEM];
  __shared__ int s_item, s_flag;
  int q = blockIdx.x % NXCD, tries = 0;
  for (;;) {
    const int it = pop_item(qdesc, qnext, q, tries, &s_item);
    if (it < 0) break;
    const GItem w = items[it];
    const GProb P = probs[w.prob];
    const int tm = w.tile / P.tiles_n, tn = w.tile % P.tiles_n;
    large_tile_f64<BM, BN, TA, TB>(P, tm, tn, w.split, smem, counters, &s_flag);
  }
}

// The 3xBF16 counterpart: items of one 128-row/column tile shape and transpose
// combination.  LDS: two stages of three BF16 planes of (BM + BN) rows x 32 k.
// 128x64 and 64x128: four waves of 64x32 / 32x64, 72 KB (two workgroups per
// CU); 128x128: eight waves of 64x32, 96 KB (one workgroup of 8 waves per CU).
template <int BM, int BN> struct X3Waves {
  static constexpr int WM = 2, WN = (BM == 128 && BN == 128) ? 4 : 2;
  static constexpr int NTH = 64 * WM * WN;
  static constexpr int MINB = NTH > GNT ? 1 : 2;  // workgroups per CU for the register budget
};
template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__((X3Waves<BM, BN>::NTH), (X3Waves<BM, BN>::MINB))
void grouped_large_x3_kernel(const GProb* __restrict__ probs, const GItem* __restrict__ items,
                             const int* __restrict__ qdesc, int* __restrict__ qnext, int* __restrict__ counters) {
  constexpr int SMEM = 2 * 3 * (BM + BN) * 32 * 2;
  __shared__ __attribute__((aligned(16))) char smem[SMEM];
  __shared__ int s_item, s_flag;
  int q = blockIdx.x % NXCD, tries = 0;
  for (;;) {
    const int it = pop_item(qdesc, qnext, q, tries, &s_item);
    if (it < 0) break;
    const GItem w = items[it];
    const GProb P = probs[w.prob];
    const int tm = w.tile / P.tiles_n, tn = w.tile % P.tiles_n;
    x3_large_tile<BM, BN, X3Waves<BM, BN>::WM, X3Waves<BM, BN>::WN, TA, TB>(P, tm, tn, w.split, smem, counters, &s_flag);
  }
}

__global__ void zero_i32_kernel(int* p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

int kind_of(int method) {
  if (method == JB_F64_NATIVE) return 0;
  if (method == JB_F32_NATIVE) return 1;
  if (method == JB_F32_BF16X3) return 2;
  return -1;
}

// ---------------------------------------------------------------------------
// autotuning table
// ---------------------------------------------------------------------------
struct TuneEntry {
  int cfg, splitk;
  double ms;
};
std::map<std::string, TuneEntry> g_table;

int ilog2(int64_t x) {
  int r = 0;
  while (x > 1) {
    x >>= 1;
    ++r;
  }
  return r;
}
// Size bucket: (log2 of the geometric mean of m and n, log2 k).
std::string bucket_key(int method, int64_t m, int64_t n, int64_t k) {
  const int mn = ilog2((int64_t)std::sqrt((double)std::max<int64_t>(m, 1) * (double)std::max<int64_t>(n, 1)));
  char buf[64];
  snprintf(buf, sizeof buf, "m%d|mn%d|k%d", method, mn, ilog2(std::max<int64_t>(k, 1)));
  return buf;
}

// Static defaults (no table entry).  Tile: the small path for m, n, k <= 64
// (one work item, B staged in LDS: on 1024 such problems it measured 43.0 us
// against 44.4 us for the tiles in FP64 and 34.7 against 37.6 us in 3xBF16);
// otherwise, for FP64, the tile with the least padded area per unit of
// measured throughput (TILE_RATE: each tile's FP64 rate on groups of
// 1024 x 1024 problems with k = 256 and 1024, relative to the 128x128 tile);
// for 3xBF16 the 64x128 tile (its split in registers makes every tile shape
// cost about the same per staged element: on the 1024-problem log-uniform set
// the 64x128 tile for every problem ran 7 % (NN) to 24 % (TN) faster than the
// 64x64 / 32x32 rule and within 2 % of the best single shape per transpose
// combination except NT, where 128x128 is 7.6 % faster -- a choice the tuner
// can make); for native FP32 the 64x64 tile from max(m, n) >= 96, else 32x32.
// Split-K is not chosen per problem: plan_build splits the longest work items
// when the whole group has fewer work items than the device holds workgroups
// (the blocking decision that depends on the device, blocking_structure's role
// in src/memory_management.jl:78-140).
constexpr double TILE_RATE[NCFG] = {0.90, 0.65, 0.0, 1.0, 1.0, 0.97};
void default_choice(int kind, int64_t m, int64_t n, int64_t k, int& cfg, int& splitk) {
  splitk = 1;
  if (m <= SMALL_MAX && n <= SMALL_MAX && k <= SMALL_MAX) {
    cfg = CFG_SMALL;
    return;
  }
  if (kind == 1) {
    cfg = std::max(m, n) >= 96 ? 0 : 1;
    return;
  }
  if (kind == 2) {
    cfg = CFG_WIDE;
    return;
  }
  double best = 0;
  for (int c : {CFG_LARGE, CFG_TALL, CFG_WIDE, 0, 1}) {
    const double padded = (double)((m + CFG_BM[c] - 1) / CFG_BM[c] * CFG_BM[c]) *
                          (double)((n + CFG_BN[c] - 1) / CFG_BN[c] * CFG_BN[c]);
    const double t = padded / TILE_RATE[c];
    if (c == CFG_LARGE || t < best) {
      best = t;
      cfg = c;
    }
  }
}

// A choice taken from a bucket (table, tuner trial, forced) may name the small
// path for a problem that does not fit it: such a problem takes the 32x32 tile;
// the 128-row/column tiles exist for FP64 and 3xBF16 (native FP32 takes 64x64).
void fit_choice(int kind, int64_t m, int64_t n, int64_t k, int& cfg, int& splitk) {
  if (cfg == CFG_SMALL) {
    if (m <= SMALL_MAX && n <= SMALL_MAX && k <= SMALL_MAX) splitk = 1;
    else cfg = 1;
  }
  if (large_cfg(cfg) && kind == 1) cfg = 0;
}

// Returns true when the choice is the static default (plan_build may then add split-K).
bool choose(int method, int64_t m, int64_t n, int64_t k, int& cfg, int& splitk) {
  const int kind = kind_of(method);
  auto it = g_table.find(bucket_key(method, m, n, k));
  bool dflt = false;
  if (it != g_table.end()) {
    cfg = it->second.cfg;
    splitk = it->second.splitk;
  } else {
    default_choice(kind, m, n, k, cfg, splitk);
    dflt = true;
  }
  fit_choice(kind, m, n, k, cfg, splitk);
  return dflt;
}

// ---------------------------------------------------------------------------
// plans
// ---------------------------------------------------------------------------
// persistent launches of a plan: L = 4 (cfg - CFG_LARGE) + 2 ta + tb for the
// 128-row/column FP64 tiles (one kernel per tile shape and transpose
// combination), then the kernel of the other tiles
constexpr int NLAUNCH = 13;

// Relative duration of one work item of problem g (MFMA work of one k-split
// of one tile plus a fixed epilogue share).
double item_cost(const GProb& g) {
  return (double)CFG_BM[g.cfg] * CFG_BN[g.cfg] * (double)std::max<int64_t>(g.kchunk, 32) + 4096.0;
}

// forced choice for the autotuner's trial launches: cfg >= 0 forces one choice
// on every problem; otherwise `over` (bucket -> choice) is consulted before the table
struct Forced {
  int cfg = -1, splitk = 1;
  const std::map<std::string, TuneEntry>* over = nullptr;
};

}  // namespace

struct GroupedPlanImpl {
  int method = 0, kind = 0;
  int nitems = 0, nprobs = 0, ncounters = 0;
  // items and grid of each launch (see NLAUNCH)
  int litems[NLAUNCH] = {}, lgrid[NLAUNCH] = {};
  void* dev = nullptr;  // probs | items | queue descriptors | counters | partials
  size_t dev_bytes = 0;
  GProb* probs = nullptr;
  GItem* items = nullptr;
  int* qdesc = nullptr;  // per launch: NXCD queue offsets into items, then NXCD queue lengths
  int* next = nullptr;   // per launch NXCD queue counters, then per-tile arrival counters (zeroed every launch)
  // the persistent launches run concurrently: all but the first on side
  // streams forked from the caller's stream after the counter reset and
  // joined back to it (a launch's tail overlaps the others' work)
  int nside = 0;
  hipStream_t side[NLAUNCH - 1] = {};
  hipEvent_t fork = nullptr, join[NLAUNCH - 1] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  // recorded after every replay: the next replay (on any stream) waits for it,
  // since all replays share the work and arrival counters
  hipEvent_t done = nullptr;
  bool launched = false;
  double flops = 0;
};

}  // namespace jb

struct jb_grouped_plan {
  jb::GroupedPlanImpl impl;
};

namespace jb {
namespace {

std::vector<jb_grouped_plan*> g_plans;

int validate_descs(int method, int64_t count, const jb_gemm_desc* d) {
  if (kind_of(method) < 0) return fail(-1, "method must be JB_F64_NATIVE, JB_F32_NATIVE or JB_F32_BF16X3");
  if (count < 0) return fail(-2, "count < 0");
  if (count > 0 && !d) return fail(-3, "descs is NULL");
  if (count > (1 << 24)) return fail(-2, "count above 2^24");
  for (int64_t i = 0; i < count; ++i) {
    GemmArgs g;
    const jb_gemm_desc& x = d[i];
    int rc = validate_gemm(x.transA, x.transB, x.m, x.n, x.k, x.alpha == 0.0, x.A, x.lda, x.B, x.ldb, x.C, x.ldc, &g);
    if (rc) return fail(-3, "descs[%lld]: %s", (long long)i, jb_last_error());
  }
  return JB_OK;
}

size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

// Resident workgroups per CU of a persistent kernel, from the device's
// register and LDS budget for that kernel (queried once per kernel and device,
// under the runtime lock; the persistent grid is this times the CU count).
// kind 0-2: the kernel of the 64x64 / 32x32 / small items of method kind 0-2;
// 3: the 128-row/column FP64 kernels (the 128x128 one's budget); 4: the
// 128x128 3xBF16 kernels; 5: the 128x64 and 64x128 3xBF16 kernels.
hipError_t grouped_occupancy(int kind, int* occ) {
  static int cached[6] = {0, 0, 0, 0, 0, 0};
  static int cached_device = -1;
  if (cached_device != rt().device) {
    for (int& c : cached) c = 0;
    cached_device = rt().device;
  }
  if (cached[kind] == 0) {
    int n = 0;
    hipError_t e;
    switch (kind) {
      case 0: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_kernel<0>, GNT, 0); break;
      case 1: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_kernel<1>, GNT, 0); break;
      case 2: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_kernel<2>, GNT, 0); break;
      case 3: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_large_kernel<128, 128, true, true>, GNT, 0); break;
      case 4:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_large_x3_kernel<128, 128, true, true>,
                                                         X3Waves<128, 128>::NTH, 0);
        break;
      default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, grouped_large_x3_kernel<128, 64, true, true>, GNT, 0); break;
    }
    if (e != hipSuccess) return e;
    cached[kind] = std::min(std::max(n, 1), 8);
  }
  *occ = cached[kind];
  return hipSuccess;
}

// The occupancy kind (grouped_occupancy) of launch L < NLAUNCH - 1 of a plan of
// method kind `kind`, and of the 128-row/column kernels as a whole (L < 0: the
// device fill of plan_choices).
int large_occ_kind(int kind, int L) {
  if (kind == 0) return 3;
  return (L >= 0 && L < 4) ? 4 : 5;
}

// Builds the device-side problem/work lists of a plan (no launch).
// Tile and split of every problem with m, n > 0 (in order), the plan's first
// two passes: the table / tuner / static choice per problem, then the device
// fill: while a persistent kernel (the 128x128 one, or the one for the other
// tiles) has fewer work items than the device holds of its resident
// workgroups (compute units x that kernel's occupancy), the split of its
// statically chosen problem with the longest work items doubles, down to 256 k
// per split.  Last, each split's k range is fixed.  flops: sum of 2 m n k.
int plan_choices(int method, int64_t count, const jb_gemm_desc* d, const Forced& forced, std::vector<GProb>& probs,
                 double* flops) {
  const int kind = kind_of(method);
  int occ = 0, occ_large = 1;
  hipError_t e = grouped_occupancy(kind, &occ);
  if (e == hipSuccess && kind != 1) e = grouped_occupancy(large_occ_kind(kind, -1), &occ_large);
  if (e != hipSuccess) return hip_fail(e, "grouped plan: occupancy query");
  std::vector<char> dflt;
  *flops = 0;
  for (int64_t i = 0; i < count; ++i) {
    const jb_gemm_desc& x = d[i];
    if (x.m == 0 || x.n == 0) continue;
    GProb g;
    memset(&g, 0, sizeof g);
    g.A = x.A;
    g.B = x.B;
    g.C = x.C;
    g.m = x.m;
    g.n = x.n;
    g.k = (x.alpha == 0.0) ? 0 : x.k;
    g.lda = x.lda;
    g.ldb = x.ldb;
    g.ldc = x.ldc;
    g.alpha = x.alpha;
    g.beta = x.beta;
    g.ta = !(x.transA == 'N' || x.transA == 'n');
    g.tb = !(x.transB == 'N' || x.transB == 'n');
    bool is_default = false;
    if (forced.cfg >= 0) {
      g.cfg = forced.cfg;
      g.splitk = forced.splitk;
    } else {
      const TuneEntry* o = nullptr;
      if (forced.over) {
        auto it = forced.over->find(bucket_key(method, g.m, g.n, g.k));
        if (it != forced.over->end()) o = &it->second;
      }
      if (o) {
        g.cfg = o->cfg;
        g.splitk = o->splitk;
      } else {
        is_default = choose(method, g.m, g.n, g.k, g.cfg, g.splitk);
      }
    }
    fit_choice(kind, g.m, g.n, g.k, g.cfg, g.splitk);
    g.tiles_m = (int)((g.m + CFG_BM[g.cfg] - 1) / CFG_BM[g.cfg]);
    g.tiles_n = (int)((g.n + CFG_BN[g.cfg] - 1) / CFG_BN[g.cfg]);
    *flops += 2.0 * (double)g.m * (double)g.n * (double)x.k;
    probs.push_back(g);
    dflt.push_back(is_default && g.cfg != CFG_SMALL);
  }
  const int64_t cus = rt().info.compute_units;
  const int64_t fill[2] = {cus * std::max(occ, 1), cus * std::max(occ_large, 1)};
  int64_t items[2] = {0, 0};  // work items of the two persistent kernels
  for (auto& g : probs) items[large_cfg(g.cfg)] += (int64_t)g.tiles_m * g.tiles_n * std::max(g.splitk, 1);
  for (;;) {
    int best = -1;
    double best_cost = 0;
    for (size_t i = 0; i < probs.size(); ++i) {
      const GProb& g = probs[i];
      const int q = large_cfg(g.cfg);
      if (!dflt[i] || items[q] >= fill[q] || g.splitk >= 64 || g.k / (2 * g.splitk) < 256) continue;
      const double cost = (double)CFG_BM[g.cfg] * CFG_BN[g.cfg] * (double)g.k / g.splitk;
      if (cost > best_cost) {
        best_cost = cost;
        best = (int)i;
      }
    }
    if (best < 0) break;
    GProb& g = probs[best];
    items[large_cfg(g.cfg)] += (int64_t)g.tiles_m * g.tiles_n * g.splitk;
    g.splitk *= 2;
  }
  // k per split: a multiple of 32, and no split comes out empty
  for (auto& g : probs) {
    int64_t kc = (g.k + g.splitk - 1) / std::max(g.splitk, 1);
    kc = std::max<int64_t>(32, (kc + 31) / 32 * 32);
    g.splitk = g.k == 0 ? 1 : (int)std::max<int64_t>(1, (g.k + kc - 1) / kc);
    g.kchunk = kc;
  }
  return JB_OK;
}

int plan_build(int method, int64_t count, const jb_gemm_desc* d, const Forced& forced, GroupedPlanImpl& P) {
  P.method = method;
  P.kind = kind_of(method);
  const size_t elt = P.kind == 0 ? 8 : 4;
  hipError_t e;
  int occ = 0;
  if ((e = grouped_occupancy(P.kind, &occ)) != hipSuccess) return hip_fail(e, "grouped plan: occupancy query");
  std::vector<GProb> probs;
  int rc = plan_choices(method, count, d, forced, probs, &P.flops);
  if (rc) return rc;
  // pass 3: arrival counters and partial slabs of the split problems
  size_t part_elems = 0;
  int counters = 0;
  std::vector<size_t> part_off(probs.size(), (size_t)-1);
  std::vector<double> prob_cost(probs.size());
  size_t nall = 0;
  for (size_t pi = 0; pi < probs.size(); ++pi) {
    GProb& g = probs[pi];
    const int ntiles = g.tiles_m * g.tiles_n;
    g.counter_base = counters;
    if (g.splitk > 1) {
      counters += ntiles;
      part_off[pi] = part_elems;
      part_elems += (size_t)ntiles * g.splitk * CFG_BM[g.cfg] * CFG_BN[g.cfg];
    }
    prob_cost[pi] = item_cost(g) * ntiles * g.splitk;
    nall += (size_t)ntiles * g.splitk;
  }
  if (nall >= ((size_t)1 << 31) || (size_t)counters >= ((size_t)1 << 30))
    return fail(JB_ERR_UNSUPPORTED, "grouped plan: %zu work items is too many for one launch", nall);
  // pass 4: work queues.  One launch per 128-row/column tile shape and
  // transpose combination, then one for the other items (NLAUNCH); each
  // launch has one queue per XCD.  A problem with at least 2 NXCD items is
  // cut into NXCD runs of consecutive items, one per queue: items are ordered
  // split by split, and within a split tile by tile in groups of `gr`
  // tile-rows (column by column inside a group), so a run is a compact block
  // of tiles sharing A row panels and B column panels.  A smaller problem goes
  // whole to the queue of its launch with the least work so far (problems
  // visited in decreasing order of their work).  Each queue runs its longest
  // items first.
  std::vector<std::pair<double, GItem>> q[NLAUNCH][NXCD];
  double load[NLAUNCH][NXCD] = {};
  std::vector<int> order(probs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return prob_cost[x] > prob_cost[y]; });
  std::vector<GItem> seq;
  for (int pi : order) {
    const GProb& g = probs[pi];
    const int L = large_cfg(g.cfg) ? 4 * (g.cfg - CFG_LARGE) + 2 * g.ta + g.tb : NLAUNCH - 1;
    const int ntiles = g.tiles_m * g.tiles_n;
    const int nitems = ntiles * g.splitk;
    const double c = item_cost(g);
    const int per_run = (nitems + NXCD - 1) / NXCD;
    const int gr = std::max(1, std::min(g.tiles_m, (int)std::lround(std::sqrt((double)per_run))));
    seq.clear();
    for (int s2 = 0; s2 < g.splitk; ++s2)
      for (int r0 = 0; r0 < g.tiles_m; r0 += gr)
        for (int tn = 0; tn < g.tiles_n; ++tn)
          for (int tm = r0; tm < std::min(r0 + gr, g.tiles_m); ++tm) seq.push_back(GItem{pi, tm * g.tiles_n + tn, s2, 0});
    if (nitems >= 2 * NXCD) {
      for (int i = 0; i < nitems; ++i) {
        const int x = (int)((int64_t)i * NXCD / nitems);
        q[L][x].push_back({c, seq[i]});
        load[L][x] += c;
      }
    } else {
      int x = 0;
      for (int y = 1; y < NXCD; ++y)
        if (load[L][y] < load[L][x]) x = y;
      for (const GItem& w : seq) q[L][x].push_back({c, w});
      load[L][x] += c * nitems;
    }
  }
  P.nprobs = (int)probs.size();
  P.nitems = (int)nall;
  P.ncounters = NLAUNCH * NXCD + counters;  // queue counters, then per-tile arrival counters
  const size_t ob_probs = 0;
  const size_t ob_items = al256(ob_probs + probs.size() * sizeof(GProb));
  const size_t ob_qdesc = al256(ob_items + nall * sizeof(GItem));
  const size_t ob_cnt = al256(ob_qdesc + (size_t)NLAUNCH * 2 * NXCD * sizeof(int));
  const size_t ob_part = al256(ob_cnt + (size_t)P.ncounters * sizeof(int));
  const size_t total = al256(ob_part + part_elems * elt) + 256;
  const char* fault = getenv("JB_FAULT_ALLOC");
  if (fault && *fault == '1') return fail(JB_ERR_ALLOC, "grouped plan: allocation refused (JB_FAULT_ALLOC)");
  e = hipMalloc(&P.dev, total);
  if (e != hipSuccess) {
    P.dev = nullptr;
    hipGetLastError();
    return fail(JB_ERR_ALLOC, "grouped plan: hipMalloc(%zu) failed: %s", total, hipGetErrorString(e));
  }
  P.dev_bytes = total;
  char* base = static_cast<char*>(P.dev);
  P.probs = reinterpret_cast<GProb*>(base + ob_probs);
  P.items = reinterpret_cast<GItem*>(base + ob_items);
  P.qdesc = reinterpret_cast<int*>(base + ob_qdesc);
  P.next = reinterpret_cast<int*>(base + ob_cnt);
  for (size_t i = 0; i < probs.size(); ++i)
    probs[i].part = part_off[i] == (size_t)-1 ? nullptr : base + ob_part + part_off[i] * elt;
  std::vector<GItem> items;
  std::vector<int> qdesc((size_t)NLAUNCH * 2 * NXCD, 0);
  items.reserve(nall);
  auto by_cost = [](const std::pair<double, GItem>& x, const std::pair<double, GItem>& y) { return x.first > y.first; };
  const int cus = rt().info.compute_units;
  for (int L = 0; L < NLAUNCH; ++L) {
    P.litems[L] = 0;
    for (int x = 0; x < NXCD; ++x) {
      std::stable_sort(q[L][x].begin(), q[L][x].end(), by_cost);
      qdesc[(size_t)L * 2 * NXCD + x] = (int)items.size();
      qdesc[(size_t)L * 2 * NXCD + NXCD + x] = (int)q[L][x].size();
      for (auto& w : q[L][x]) items.push_back(w.second);
      P.litems[L] += (int)q[L][x].size();
    }
    int occ_l = occ;
    if (L < NLAUNCH - 1 && P.litems[L] > 0 && (e = grouped_occupancy(large_occ_kind(P.kind, L), &occ_l)) != hipSuccess)
      return hip_fail(e, "grouped plan: occupancy query");
    P.lgrid[L] = std::max(1, std::min(P.litems[L], cus * occ_l));
  }
  if (!probs.empty() && (e = hipMemcpy(P.probs, probs.data(), probs.size() * sizeof(GProb), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "grouped plan: upload problems");
  if (!items.empty() && (e = hipMemcpy(P.items, items.data(), items.size() * sizeof(GItem), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "grouped plan: upload work list");
  if ((e = hipMemcpy(P.qdesc, qdesc.data(), qdesc.size() * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(e, "grouped plan: upload queue descriptors");
  int nlaunch = 0;
  for (int L = 0; L < NLAUNCH; ++L) nlaunch += P.litems[L] > 0;
  if (log_enabled()) {
    int per_cfg[NCFG] = {}, split = 0;
    for (const GProb& g : probs) {
      ++per_cfg[g.cfg];
      split += g.splitk > 1;
    }
    char launches[256] = "";
    size_t at = 0;
    for (int L = 0; L < NLAUNCH && at < sizeof launches; ++L)
      if (P.litems[L] > 0)
        at += snprintf(launches + at, sizeof launches - at, " [%d: %d items, grid %d]", L, P.litems[L], P.lgrid[L]);
    log_line("grouped plan: method %d, %d problems (64x64 %d, 32x32 %d, small %d, 128x128 %d, 128x64 %d, 64x128 %d;"
             " %d split), %d items, %.4g flops, %zu device bytes, launches%s",
             method, P.nprobs, per_cfg[0], per_cfg[1], per_cfg[2], per_cfg[3], per_cfg[4], per_cfg[5], split,
             P.nitems, P.flops, P.dev_bytes, launches);
  }
  if (nlaunch > 1) {
    if ((e = hipEventCreateWithFlags(&P.fork, hipEventDisableTiming)) != hipSuccess)
      return hip_fail(e, "grouped plan: event");
    for (; P.nside < nlaunch - 1; ++P.nside) {
      if ((e = hipStreamCreateWithFlags(&P.side[P.nside], hipStreamNonBlocking)) != hipSuccess)
        return hip_fail(e, "grouped plan: side stream");
      if ((e = hipEventCreateWithFlags(&P.join[P.nside], hipEventDisableTiming)) != hipSuccess) {
        hipStreamDestroy(P.side[P.nside]);
        P.side[P.nside] = nullptr;
        return hip_fail(e, "grouped plan: event");
      }
    }
  }
  return JB_OK;
}

typedef void (*GroupedKernelFn)(const GProb*, const GItem*, const int*, int*, int*);
const GroupedKernelFn kLargeKernels[NLAUNCH - 1] = {
    grouped_large_kernel<128, 128, false, false>, grouped_large_kernel<128, 128, false, true>,
    grouped_large_kernel<128, 128, true, false>,  grouped_large_kernel<128, 128, true, true>,
    grouped_large_kernel<128, 64, false, false>,  grouped_large_kernel<128, 64, false, true>,
    grouped_large_kernel<128, 64, true, false>,   grouped_large_kernel<128, 64, true, true>,
    grouped_large_kernel<64, 128, false, false>,  grouped_large_kernel<64, 128, false, true>,
    grouped_large_kernel<64, 128, true, false>,   grouped_large_kernel<64, 128, true, true>};
const GroupedKernelFn kLargeX3Kernels[NLAUNCH - 1] = {
    grouped_large_x3_kernel<128, 128, false, false>, grouped_large_x3_kernel<128, 128, false, true>,
    grouped_large_x3_kernel<128, 128, true, false>,  grouped_large_x3_kernel<128, 128, true, true>,
    grouped_large_x3_kernel<128, 64, false, false>,  grouped_large_x3_kernel<128, 64, false, true>,
    grouped_large_x3_kernel<128, 64, true, false>,   grouped_large_x3_kernel<128, 64, true, true>,
    grouped_large_x3_kernel<64, 128, false, false>,  grouped_large_x3_kernel<64, 128, false, true>,
    grouped_large_x3_kernel<64, 128, true, false>,   grouped_large_x3_kernel<64, 128, true, true>};

// Enqueues [reset counters, 128-row/column persistent kernels, persistent kernel] on s.
hipError_t plan_enqueue(const GroupedPlanImpl& P, hipStream_t s) {
  if (P.nitems == 0) return hipSuccess;
  hipLaunchKernelGGL(zero_i32_kernel, dim3((P.ncounters + 255) / 256 < 1024 ? (P.ncounters + 255) / 256 : 1024),
                     dim3(256), 0, s, P.next, P.ncounters);
  timing_begin(s, "grouped_gemm");
  hipError_t e = hipSuccess;
  if (P.nside > 0 && (e = hipEventRecord(P.fork, s)) != hipSuccess) return e;
  int* const tiles = P.next + NLAUNCH * NXCD;
  int j = -1;  // launch j + 1 > 0 goes to side stream j
  for (int L = 0; L < NLAUNCH; ++L) {
    if (P.litems[L] == 0) continue;
    hipStream_t ls = s;
    if (j >= 0) {
      ls = P.side[j];
      if ((e = hipStreamWaitEvent(ls, P.fork, 0)) != hipSuccess) return e;
    }
    const int* qd = P.qdesc + L * 2 * NXCD;
    int* qn = P.next + L * NXCD;
    // the 128x128 3xBF16 kernels run workgroups of 8 waves
    const dim3 grid(P.lgrid[L]), block(P.kind == 2 && L < 4 ? X3Waves<128, 128>::NTH : GNT);
    if (L < NLAUNCH - 1) {
      hipLaunchKernelGGL(P.kind == 0 ? kLargeKernels[L] : kLargeX3Kernels[L], grid, block, 0, ls, P.probs, P.items, qd, qn,
                         tiles);
    } else if (P.kind == 0) {
      hipLaunchKernelGGL(grouped_kernel<0>, grid, block, 0, ls, P.probs, P.items, qd, qn, tiles);
    } else if (P.kind == 1) {
      hipLaunchKernelGGL(grouped_kernel<1>, grid, block, 0, ls, P.probs, P.items, qd, qn, tiles);
    } else {
      hipLaunchKernelGGL(grouped_kernel<2>, grid, block, 0, ls, P.probs, P.items, qd, qn, tiles);
    }
    if (j >= 0) {
      if ((e = hipEventRecord(P.join[j], ls)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(s, P.join[j], 0)) != hipSuccess) return e;
    }
    ++j;
  }
  timing_end(s, "grouped_gemm");
  return hipGetLastError();
}

void plan_free(GroupedPlanImpl& P) {
  for (int i = 0; i < P.nside; ++i) {
    if (P.side[i]) hipStreamSynchronize(P.side[i]);
    if (P.side[i]) hipStreamDestroy(P.side[i]);
    if (P.join[i]) hipEventDestroy(P.join[i]);
  }
  if (P.fork) hipEventDestroy(P.fork);
  if (P.exec) hipGraphExecDestroy(P.exec);
  if (P.graph) hipGraphDestroy(P.graph);
  if (P.cap) hipStreamDestroy(P.cap);
  if (P.done) hipEventDestroy(P.done);
  if (P.dev) hipFree(P.dev);
  P = GroupedPlanImpl();
}

int plan_capture(GroupedPlanImpl& P) {
  if (getenv("JB_FAULT_GRAPH")) return fail(JB_ERR_GRAPH, "graph capture refused (JB_FAULT_GRAPH)");
  hipError_t e = hipStreamCreateWithFlags(&P.cap, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_fail(e, "grouped plan: capture stream");
  if ((e = hipStreamBeginCapture(P.cap, hipStreamCaptureModeThreadLocal)) != hipSuccess)
    return fail(JB_ERR_GRAPH, "hipStreamBeginCapture: %s", hipGetErrorString(e));
  hipError_t le = plan_enqueue(P, P.cap);
  e = hipStreamEndCapture(P.cap, &P.graph);
  if (le != hipSuccess || e != hipSuccess)
    return fail(JB_ERR_GRAPH, "graph capture failed: %s", hipGetErrorString(le != hipSuccess ? le : e));
  if ((e = hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0)) != hipSuccess)
    return fail(JB_ERR_GRAPH, "hipGraphInstantiate: %s", hipGetErrorString(e));
  if ((e = hipEventCreateWithFlags(&P.done, hipEventDisableTiming)) != hipSuccess)
    return hip_fail(e, "grouped plan: event");
  return JB_OK;
}

// One-shot grouped launch (no graph), synchronous.
int run_grouped_once(int method, int64_t count, const jb_gemm_desc* d, const Forced& forced, hipStream_t s) {
  GroupedPlanImpl P;
  int rc = plan_build(method, count, d, forced, P);
  if (rc == JB_OK) {
    hipError_t e = plan_enqueue(P, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = hip_fail(e, "grouped launch");
  }
  hipStreamSynchronize(s);
  plan_free(P);
  return rc;
}

}  // namespace

void grouped_release_all() {
  for (auto* p : g_plans) {
    plan_free(p->impl);
    delete p;
  }
  g_plans.clear();
}

// ---- JSON persistence of the table ------------------------------------------
static std::string table_prefix() {
  char buf[160];
  snprintf(buf, sizeof buf, "%s|%s|", rt().info.name, rt().info.arch);
  return buf;
}

int tune_save_locked(const char* path) {
  FILE* f = fopen(path, "w");
  if (!f) return fail(-1, "cannot open %s for writing", path);
  fprintf(f, "{\n  \"format\": \"jblas-tune-1\",\n  \"entries\": [\n");
  size_t i = 0;
  const std::string pre = table_prefix();
  for (auto& kv : g_table) {
    fprintf(f, "    {\"key\": \"%s%s\", \"cfg\": %d, \"splitk\": %d, \"ms\": %.6f}%s\n", pre.c_str(), kv.first.c_str(),
            kv.second.cfg, kv.second.splitk, kv.second.ms, ++i < g_table.size() ? "," : "");
  }
  fprintf(f, "  ]\n}\n");
  fclose(f);
  return JB_OK;
}

int tune_load_locked(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return fail(-1, "cannot open %s", path);
  std::string txt;
  char buf[4096];
  size_t nr;
  while ((nr = fread(buf, 1, sizeof buf, f)) > 0) txt.append(buf, nr);
  fclose(f);
  if (txt.find("\"jblas-tune-1\"") == std::string::npos) return fail(JB_ERR_STATE, "%s: not a jblas tuning table", path);
  const std::string pre = table_prefix();
  std::map<std::string, TuneEntry> loaded;
  size_t pos = 0;
  while ((pos = txt.find("{\"key\": \"", pos)) != std::string::npos) {
    pos += 9;
    const size_t q = txt.find('"', pos);
    if (q == std::string::npos) break;
    std::string key = txt.substr(pos, q - pos);
    int cfg = -1, sk = -1;
    double ms = 0;
    if (sscanf(txt.c_str() + q, "\", \"cfg\": %d, \"splitk\": %d, \"ms\": %lf", &cfg, &sk, &ms) != 3)
      return fail(JB_ERR_STATE, "%s: malformed entry", path);
    if (cfg < 0 || cfg >= NCFG || sk < 1 || sk > 64) return fail(JB_ERR_STATE, "%s: entry out of range", path);
    if (key.compare(0, pre.size(), pre) == 0) loaded[key.substr(pre.size())] = TuneEntry{cfg, sk, ms};
    pos = q;
  }
  for (auto& kv : loaded) g_table[kv.first] = kv.second;  // entries of other devices are ignored
  return JB_OK;
}

}  // namespace jb

using namespace jb;

extern "C" {

int jb_gemm_grouped_dev(int method, int64_t count, const jb_gemm_desc* descs, void* stream) {
  int rc = validate_descs(method, count, descs);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(rt().mu);
  if ((rc = ensure_init())) return rc;
  return run_grouped_once(method, count, descs, Forced(), (hipStream_t)stream);
}

int jb_grouped_plan_create(int method, int64_t count, const jb_gemm_desc* descs, void* stream,
                           jb_grouped_plan** plan) {
  (void)stream;
  if (!plan) return fail(-5, "plan is NULL");
  *plan = nullptr;
  int rc = validate_descs(method, count, descs);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(rt().mu);
  if ((rc = ensure_init())) return rc;
  auto* p = new jb_grouped_plan();
  rc = plan_build(method, count, descs, Forced(), p->impl);
  if (rc == JB_OK) rc = plan_capture(p->impl);
  if (rc != JB_OK) {
    plan_free(p->impl);
    delete p;
    return rc;
  }
  g_plans.push_back(p);
  *plan = p;
  return JB_OK;
}

int jb_grouped_plan_launch(jb_grouped_plan* plan, void* stream) {
  std::lock_guard<std::mutex> lk(rt().mu);
  if (!plan || std::find(g_plans.begin(), g_plans.end(), plan) == g_plans.end())
    return fail(JB_ERR_STATE, "unknown grouped plan");
  GroupedPlanImpl& P = plan->impl;
  if (P.nitems == 0) return JB_OK;
  hipError_t e = hipSuccess;
  // replays of one plan never overlap: they share its counters
  if (P.launched && (e = hipStreamWaitEvent((hipStream_t)stream, P.done, 0)) != hipSuccess)
    return hip_fail(e, "grouped plan: wait for the previous replay");
  if ((e = hipGraphLaunch(P.exec, (hipStream_t)stream)) != hipSuccess)
    return fail(JB_ERR_GRAPH, "hipGraphLaunch: %s", hipGetErrorString(e));
  if ((e = hipEventRecord(P.done, (hipStream_t)stream)) != hipSuccess) return hip_fail(e, "grouped plan: record replay");
  P.launched = true;
  return JB_OK;
}

int jb_grouped_plan_destroy(jb_grouped_plan* plan) {
  std::lock_guard<std::mutex> lk(rt().mu);
  auto it = std::find(g_plans.begin(), g_plans.end(), plan);
  if (!plan || it == g_plans.end()) return fail(JB_ERR_STATE, "unknown grouped plan");
  hipDeviceSynchronize();
  plan_free(plan->impl);
  delete plan;
  g_plans.erase(it);
  return JB_OK;
}

// Builds the plan of the whole group with the bucket choices `over`, runs it
// once to warm up and `reps` times under events; *ms = the fastest run.
static int autotune_trial(int method, int64_t count, const jb_gemm_desc* descs,
                          const std::map<std::string, TuneEntry>& over, int reps, hipStream_t s, hipEvent_t e0,
                          hipEvent_t e1, double* ms) {
  Forced fc;
  fc.over = &over;
  GroupedPlanImpl P;
  int rc = plan_build(method, count, descs, fc, P);
  if (rc) {
    plan_free(P);
    return rc;
  }
  hipError_t e = plan_enqueue(P, s);
  double best = 1e30;
  for (int r = 0; r < reps && e == hipSuccess; ++r) {
    if ((e = hipEventRecord(e0, s)) != hipSuccess) break;
    if ((e = plan_enqueue(P, s)) != hipSuccess) break;
    if ((e = hipEventRecord(e1, s)) != hipSuccess) break;
    if ((e = hipEventSynchronize(e1)) != hipSuccess) break;
    float t = 0;
    hipEventElapsedTime(&t, e0, e1);
    best = std::min(best, (double)t);
  }
  hipStreamSynchronize(s);
  plan_free(P);
  if (e != hipSuccess) return hip_fail(e, "autotune trial");
  *ms = best;
  return JB_OK;
}

int jb_autotune_grouped(int method, int64_t count, const jb_gemm_desc* descs, int reps, void* stream) {
  int rc = validate_descs(method, count, descs);
  if (rc) return rc;
  if (reps < 1 || reps > 1000) return fail(-4, "reps must be in 1..1000");
  std::lock_guard<std::mutex> lk(rt().mu);
  if ((rc = ensure_init())) return rc;
  hipStream_t s = (hipStream_t)stream;
  // Tuning happens in context: every trial times the whole group, so a
  // bucket's choice is judged by what it does to the group's time (a split-K
  // that helps a bucket alone can slow a group that already fills the GPU).
  // Coordinate descent, one pass, buckets in decreasing order of their flops;
  // a candidate replaces the current choice only if the group gets >= 1 % faster.
  struct Bucket {
    double flops = 0;
    int64_t k = 0;
  };
  std::map<std::string, Bucket> buckets;
  double total = 0;
  for (int64_t i = 0; i < count; ++i) {
    const jb_gemm_desc& d = descs[i];
    if (d.m <= 0 || d.n <= 0 || d.k <= 0) continue;
    Bucket& b = buckets[bucket_key(method, d.m, d.n, d.k)];
    const double f = 2.0 * (double)d.m * (double)d.n * (double)d.k;
    b.flops += f;
    b.k = std::max<int64_t>(b.k, d.k);
    total += f;
  }
  std::vector<std::pair<double, std::string>> order;
  for (auto& kv : buckets)
    if (kv.second.flops >= 0.005 * total) order.push_back({kv.second.flops, kv.first});
  std::sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
  // Trials write into a zeroed scratch copy of every C (ldc = m): the caller's
  // C is neither read nor written, and beta != 0 problems still read a C.
  const size_t elt = kind_of(method) == 0 ? 8 : 4;
  std::vector<jb_gemm_desc> trial(descs, descs + count);
  size_t cbytes = 0;
  for (auto& d : trial) {
    if (d.m <= 0 || d.n <= 0) continue;
    cbytes = al256(cbytes);
    d.C = reinterpret_cast<void*>(cbytes);  // offset for now
    d.ldc = d.m;
    cbytes += (size_t)d.m * (size_t)d.n * elt;
  }
  void* cscratch = nullptr;
  if (cbytes > 0) {
    const char* fault = getenv("JB_FAULT_ALLOC");
    if (fault && *fault == '1') return fail(JB_ERR_ALLOC, "autotune: allocation refused (JB_FAULT_ALLOC)");
    hipError_t e = hipMalloc(&cscratch, cbytes);
    if (e != hipSuccess) {
      hipGetLastError();
      return fail(JB_ERR_ALLOC, "autotune: hipMalloc(%zu) for the trial outputs failed", cbytes);
    }
    if ((e = hipMemsetAsync(cscratch, 0, cbytes, s)) != hipSuccess) {
      hipFree(cscratch);
      return hip_fail(e, "autotune: clear trial outputs");
    }
    for (auto& d : trial)
      if (d.m > 0 && d.n > 0) d.C = static_cast<char*>(cscratch) + reinterpret_cast<size_t>(d.C);
  }
  descs = trial.data();
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  std::map<std::string, TuneEntry> over;
  double cur = 0;
  // the first trial of a group runs slower (first touch of the operands, clock
  // ramp): a throwaway trial before the baseline keeps the descent from
  // crediting candidates with that difference
  rc = autotune_trial(method, count, descs, over, reps, s, e0, e1, &cur);
  if (rc == JB_OK) rc = autotune_trial(method, count, descs, over, reps, s, e0, e1, &cur);
  for (size_t bi = 0; bi < order.size() && rc == JB_OK; ++bi) {
    const std::string& key = order[bi].second;
    const int64_t kmax = buckets[key].k;
    TuneEntry best{-1, 1, cur};
    for (int cfg = 0; cfg < NCFG && rc == JB_OK; ++cfg)
      for (int sk = 1; sk <= 16 && rc == JB_OK; sk *= 2) {
        if (sk > 1 && kmax / sk < 128) continue;
        if (cfg == CFG_SMALL && (sk > 1 || kmax > SMALL_MAX)) continue;  // one split; k must fit
        if (large_cfg(cfg) && kind_of(method) == 1) continue;               // FP64 and 3xBF16 only
        over[key] = TuneEntry{cfg, sk, 0.0};
        double t = 0;
        rc = autotune_trial(method, count, descs, over, reps, s, e0, e1, &t);
        if (rc == JB_OK && t < 0.99 * best.ms) best = TuneEntry{cfg, sk, t};
      }
    if (best.cfg >= 0) {
      over[key] = best;
      cur = best.ms;
      log_line("autotune %s: tile %d, split %d, group %.4f ms", key.c_str(), best.cfg, best.splitk, best.ms);
    } else {
      over.erase(key);
    }
  }
  // Confirmation: the descent compared single trials, so a choice that only
  // looked >= 1 % faster through timing noise can slip in; the tuned choices
  // are kept only if the whole group, timed again alternately with the
  // static rule (twice as many launches each), is still >= 1 % faster.
  if (rc == JB_OK && !over.empty()) {
    const std::map<std::string, TuneEntry> none;
    double t_static = 1e30, t_tuned = 1e30;
    for (int round = 0; round < 2 && rc == JB_OK; ++round) {
      double t = 0;
      if ((rc = autotune_trial(method, count, descs, none, 2 * reps, s, e0, e1, &t)) != JB_OK) break;
      t_static = std::min(t_static, t);
      if ((rc = autotune_trial(method, count, descs, over, 2 * reps, s, e0, e1, &t)) != JB_OK) break;
      t_tuned = std::min(t_tuned, t);
    }
    if (rc == JB_OK) log_line("autotune confirmation: static %.4f ms, tuned %.4f ms: %s", t_static, t_tuned,
                              t_tuned < 0.99 * t_static ? "kept" : "dropped");
    if (rc == JB_OK && !(t_tuned < 0.99 * t_static)) over.clear();
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (cscratch) {
    hipStreamSynchronize(s);
    hipFree(cscratch);
  }
  if (rc) return rc;
  for (auto& kv : over) g_table[kv.first] = kv.second;
  return JB_OK;
}

int jb_autotune_save(const char* path) {
  if (!path) return fail(-1, "path is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  int rc = ensure_init();
  if (rc) return rc;
  return tune_save_locked(path);
}

int jb_autotune_load(const char* path) {
  if (!path) return fail(-1, "path is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  int rc = ensure_init();
  if (rc) return rc;
  return tune_load_locked(path);
}

int jb_autotune_lookup(int method, int64_t m, int64_t n, int64_t k, int* tile, int* splitk) {
  if (kind_of(method) < 0) return fail(-1, "unknown method");
  if (m < 0 || n < 0 || k < 0) return fail(-2, "negative size");
  if (!tile || !splitk) return fail(-5, "output pointer is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  choose(method, m, n, k, *tile, *splitk);
  return JB_OK;
}

int jb_grouped_choices(int method, int64_t count, const jb_gemm_desc* descs, int* tile, int* splitk,
                       int64_t* kchunk) {
  int rc = validate_descs(method, count, descs);
  if (rc) return rc;
  if (count > 0 && (!tile || !splitk || !kchunk)) return fail(-5, "output pointer is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  if ((rc = ensure_init())) return rc;
  std::vector<GProb> probs;
  double flops = 0;
  if ((rc = plan_choices(method, count, descs, Forced(), probs, &flops))) return rc;
  size_t j = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (descs[i].m == 0 || descs[i].n == 0) {
      tile[i] = -1;
      splitk[i] = 0;
      kchunk[i] = 0;
      continue;
    }
    tile[i] = probs[j].cfg;
    splitk[i] = probs[j].splitk;
    kchunk[i] = probs[j].kchunk;
    ++j;
  }
  return JB_OK;
}

int jb_autotune_reset(void) {
  std::lock_guard<std::mutex> lk(rt().mu);
  g_table.clear();
  return JB_OK;
}

// ---- pullback -------------------------------------------------------------
static int pullback_descs(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const void* A, int64_t lda,
                          const void* B, int64_t ldb, const void* dC, int64_t lddc, void* dA, int64_t ldda, void* dB,
                          int64_t lddb, std::vector<jb_gemm_desc>& out) {
  const bool tA = !(ta == 'N' || ta == 'n'), tB = !(tb == 'N' || tb == 'n');
  jb_gemm_desc d;
  memset(&d, 0, sizeof d);
  d.alpha = alpha;
  d.beta = 0.0;
  if (dA) {
    // dA (shape of A) = alpha*dC*op(B)^T ('N')  or  alpha*op(B)*dC^T ('T')
    if (!tA) {
      d.transA = 'N'; d.A = dC; d.lda = lddc;
      d.transB = tB ? 'N' : 'T'; d.B = B; d.ldb = ldb;     // op(B)^T
      d.m = m; d.n = k; d.k = n;
    } else {
      d.transA = tB ? 'T' : 'N'; d.A = B; d.lda = ldb;     // op(B)
      d.transB = 'T'; d.B = dC; d.ldb = lddc;              // dC^T
      d.m = k; d.n = m; d.k = n;
    }
    d.C = dA; d.ldc = ldda;
    out.push_back(d);
  }
  if (dB) {
    // dB (shape of B) = alpha*op(A)^T*dC ('N')  or  alpha*dC^T*op(A) ('T')
    if (!tB) {
      d.transA = tA ? 'N' : 'T'; d.A = A; d.lda = lda;     // op(A)^T
      d.transB = 'N'; d.B = dC; d.ldb = lddc;
      d.m = k; d.n = n; d.k = m;
    } else {
      d.transA = 'T'; d.A = dC; d.lda = lddc;              // dC^T
      d.transB = tA ? 'T' : 'N'; d.B = A; d.ldb = lda;     // op(A)
      d.m = n; d.n = k; d.k = m;
    }
    d.C = dB; d.ldc = lddb;
    out.push_back(d);
  }
  return JB_OK;
}

// The adjoint GEMMs of one large product fill the GPU on their own: those run
// on the single-problem kernels (128x128 tiles, stream-ordered one after the
// other); smaller ones share one grouped launch.
static bool pullback_big(const jb_gemm_desc& d) {
  return (double)d.m * (double)d.n >= 256.0 * 128.0 * 128.0 && d.k >= 256;
}

static int pullback_run_locked(int method, const std::vector<jb_gemm_desc>& d, hipStream_t s) {
  std::vector<jb_gemm_desc> small;
  for (const jb_gemm_desc& x : d) {
    if (!pullback_big(x)) {
      small.push_back(x);
      continue;
    }
    GemmArgs g;
    int rc = validate_gemm(x.transA, x.transB, x.m, x.n, x.k, x.alpha == 0.0, x.A, x.lda, x.B, x.ldb, x.C, x.ldc, &g);
    if (rc) return rc;
    g.alpha = x.alpha;
    g.beta = x.beta;
    if ((rc = single_gemm_dev_locked(method, g, s))) return rc;
  }
  if (small.empty()) return JB_OK;
  return run_grouped_once(method, (int64_t)small.size(), small.data(), Forced(), s);
}

static int pullback_validate(int method, char ta, char tb, int64_t m, int64_t n, int64_t k, const void* A, int64_t lda,
                             const void* B, int64_t ldb, const void* dC, int64_t lddc, const void* dA, int64_t ldda,
                             const void* dB, int64_t lddb) {
  if (kind_of(method) < 0) return fail(-1, "method must be JB_F64_NATIVE, JB_F32_NATIVE or JB_F32_BF16X3");
  const bool tA = !(ta == 'N' || ta == 'n'), tB = !(tb == 'N' || tb == 'n');
  if (!(ta == 'N' || ta == 'n' || ta == 'T' || ta == 't' || ta == 'C' || ta == 'c')) return fail(-2, "bad transA");
  if (!(tb == 'N' || tb == 'n' || tb == 'T' || tb == 't' || tb == 'C' || tb == 'c')) return fail(-3, "bad transB");
  if (m < 0) return fail(-4, "m < 0");
  if (n < 0) return fail(-5, "n < 0");
  if (k < 0) return fail(-6, "k < 0");
  const int64_t ar = tA ? k : m, br = tB ? n : k;
  if (dB && m > 0 && !A) return fail(-8, "A is NULL");
  if (lda < std::max<int64_t>(1, ar)) return fail(-9, "lda too small");
  if (dA && n > 0 && !B) return fail(-10, "B is NULL");
  if (ldb < std::max<int64_t>(1, br)) return fail(-11, "ldb too small");
  if (!dC && m > 0 && n > 0) return fail(-12, "dC is NULL");
  if (lddc < std::max<int64_t>(1, m)) return fail(-13, "lddc too small");
  if (dA && ldda < std::max<int64_t>(1, ar)) return fail(-15, "ldda too small");
  if (dB && lddb < std::max<int64_t>(1, br)) return fail(-17, "lddb too small");
  return JB_OK;
}

int jb_gemm_pullback_dev(int method, char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                         const void* A, int64_t lda, const void* B, int64_t ldb, const void* dC, int64_t lddc,
                         void* dA, int64_t ldda, void* dB, int64_t lddb, void* stream) {
  int rc = pullback_validate(method, transA, transB, m, n, k, A, lda, B, ldb, dC, lddc, dA, ldda, dB, lddb);
  if (rc) return rc;
  std::vector<jb_gemm_desc> d;
  pullback_descs(transA, transB, m, n, k, alpha, A, lda, B, ldb, dC, lddc, dA, ldda, dB, lddb, d);
  if (d.empty()) return JB_OK;
  if ((rc = validate_descs(method, (int64_t)d.size(), d.data()))) return rc;
  std::lock_guard<std::mutex> lk(rt().mu);
  if ((rc = ensure_init())) return rc;
  return pullback_run_locked(method, d, (hipStream_t)stream);
}

int jb_gemm_pullback(int method, char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                     const void* A, int64_t lda, const void* B, int64_t ldb, const void* dC, int64_t lddc, void* dA,
                     int64_t ldda, void* dB, int64_t lddb) {
  int rc = pullback_validate(method, transA, transB, m, n, k, A, lda, B, ldb, dC, lddc, dA, ldda, dB, lddb);
  if (rc) return rc;
  const size_t elt = (method == JB_F64_NATIVE) ? 8 : 4;
  const bool tA = !(transA == 'N' || transA == 'n'), tB = !(transB == 'N' || transB == 'n');
  const int64_t ar = tA ? k : m, ac = tA ? m : k, br = tB ? n : k, bc = tB ? k : n;
  // stage host operands in device memory with packed leading dimensions
  std::lock_guard<std::mutex> lk(rt().mu);
  if ((rc = ensure_init())) return rc;
  const size_t sA = al256((size_t)ar * ac * elt), sB = al256((size_t)br * bc * elt), sC = al256((size_t)m * n * elt);
  if ((rc = rt().staging.reserve(2 * (sA + sB) + sC + 256, "staging"))) return rc;
  char* base = static_cast<char*>(rt().staging.ptr);
  char *dAv = base, *dBv = base + sA, *dCv = base + sA + sB, *gA = dCv + sC, *gB = gA + sA;
  hipStream_t s = rt().stream;
  hipError_t e = hipSuccess;
  auto h2d = [&](void* dst, const void* src, int64_t rows, int64_t cols, int64_t ld) {
    if (e == hipSuccess && rows > 0 && cols > 0)
      e = (ld == rows || cols == 1)
              ? hipMemcpyAsync(dst, src, (size_t)rows * cols * elt, hipMemcpyHostToDevice, s)
              : hipMemcpy2DAsync(dst, rows * elt, src, ld * elt, rows * elt, cols, hipMemcpyHostToDevice, s);
  };
  auto d2h = [&](void* dst, const void* src, int64_t rows, int64_t cols, int64_t ld) {
    if (e == hipSuccess && rows > 0 && cols > 0)
      e = (ld == rows || cols == 1)
              ? hipMemcpyAsync(dst, src, (size_t)rows * cols * elt, hipMemcpyDeviceToHost, s)
              : hipMemcpy2DAsync(dst, ld * elt, src, rows * elt, rows * elt, cols, hipMemcpyDeviceToHost, s);
  };
  if (dB) h2d(dAv, A, ar, ac, lda);
  if (dA) h2d(dBv, B, br, bc, ldb);
  h2d(dCv, dC, m, n, lddc);
  if (e != hipSuccess) return hip_fail(e, "pullback: copy to device");
  std::vector<jb_gemm_desc> d;
  pullback_descs(transA, transB, m, n, k, alpha, dAv, std::max<int64_t>(1, ar), dBv, std::max<int64_t>(1, br), dCv,
                 std::max<int64_t>(1, m), dA ? gA : nullptr, std::max<int64_t>(1, ar), dB ? gB : nullptr,
                 std::max<int64_t>(1, br), d);
  if (!d.empty()) {
    for (auto& x : d)
      if ((rc = validate_descs(method, 1, &x))) return rc;
    if ((rc = pullback_run_locked(method, d, s))) return rc;
  }
  if (dA) d2h(dA, gA, ar, ac, ldda);
  if (dB) d2h(dB, gB, br, bc, lddb);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return e == hipSuccess ? JB_OK : hip_fail(e, "pullback: copy to host");
}

}  // extern "C"
