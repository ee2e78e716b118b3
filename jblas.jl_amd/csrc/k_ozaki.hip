// k_ozaki.hip -- split-BF16 emulated DGEMM and 3xBF16 SGEMM on gfx950.
//
// The reference has no emulated path (its GEMM is the FP64/FP32 register tile
// of src/gemm.jl:244-348); this file implements the method BASELINE names:
//
//   1. absmax  (HBM-bound): per row of op(A) / column of op(B), the bit pattern
//      of max|x| (an unsigned max of |x| bits orders finite < inf < NaN).
//   2. split   (HBM-bound): E = frexp exponent of that max; x = ldexp(a, -E);
//      balanced digits d_1 = rint(x*2^7), d_t = rint(r*2^8) of the running
//      residual, |d| <= 128, each exact in BF16.  Slice planes are written
//      k-contiguous and zero-padded: plane[t][row][k], rows padded to the block
//      tile, k to the LDS stage -- the GEMM kernel needs no bounds checks.
//   3. slice-pair GEMM (MFMA-bound): per 32x32 wave tile one FP32 accumulator
//      set per level l = t+u (pairs t+u <= s+1), fed by
//      v_mfma_f32_32x32x16_bf16.  Every kc k-values (kc*s <= 1024, so a level
//      sum stays below 2^24 and is exact whatever the MFMA's internal order)
//      the level sums are flushed into FP64 group accumulators,
//      G_g = fma(S_l, 2^-8j, G_g) with l = 2+4g+j.  The epilogue combines the
//      groups by TwoSum from the smallest, rescales by 2^(E_i+F_j-14) and
//      applies alpha/beta.  These are the exact operations of the oracle's
//      restatement, so the result is bit-reproducible.
//   BF16X3 mode (FP32 inputs): per-element split a = a1 + a2 + a3 (bf16 round
//   to nearest of the running residual, exact), products t+u <= 2+P, level 2
//   in one FP32 accumulator and the smaller levels in a second.
#include <algorithm>

#include "jb_internal.h"

namespace jb {
namespace {

typedef float f16v __attribute__((ext_vector_type(16)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int ZNT = 512;            // 8 waves
constexpr int ZBM = 128, ZBN = 64;  // block tile (4 x 2 waves of 32 x 32)
constexpr int ZBK = 16;             // k per LDS stage = one MFMA k-step
constexpr int ZROWS = ZBM + ZBN;    // rows of one slice-plane stage
constexpr int ZPAD_K = 64;          // k padding of the planes (split tile width)

__host__ __device__ constexpr int chunk_len(int s) {
  int kc = 1024;
  while (kc * s > 1024) kc >>= 1;
  return kc;
}

__device__ __forceinline__ int exponent_of_bits(unsigned long long bits) {
  if (bits == 0ull) return 0;
  int e;
  frexp(__longlong_as_double((long long)bits), &e);
  return e;
}

// ---------------------------------------------------------------------------
// 1. absmax: maxbits[r] = max_k bits(|X(r,k)|), X(r,k) = P[r*sr + k*sk].
//    64 x 64 tiles, coalesced along whichever index is contiguous.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void absmax_kernel(const T* __restrict__ P, int64_t sr, int64_t sk, int64_t R,
                                                     int64_t K, unsigned long long* __restrict__ maxbits) {
  __shared__ unsigned long long tile[64][65];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  const bool rows_contig = (sr == 1);
#pragma unroll 4
  for (int j = 0; j < 16; ++j) {
    const int r = rows_contig ? (t % 64) : (t / 64 + 4 * j);
    const int k = rows_contig ? (t / 64 + 4 * j) : (t % 64);
    unsigned long long b = 0ull;
    if (r0 + r < R && k0 + k < K) {
      const double v = fabs((double)P[(r0 + r) * sr + (k0 + k) * sk]);
      b = (unsigned long long)__double_as_longlong(v);
    }
    tile[r][k] = b;
  }
  __syncthreads();
  if (t < 64 && r0 + t < R) {
    unsigned long long m = 0ull;
    for (int k = 0; k < 64; ++k) m = max(m, tile[t][k]);
    atomicMax(maxbits + r0 + t, m);
  }
}

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

__device__ __forceinline__ uint16_t bf16_bits_exact(double d) {
  // d is an integer with |d| <= 256: exact in float, and the float's low 16 bits are zero.
  return (uint16_t)(__float_as_uint((float)d) >> 16);
}

// ---------------------------------------------------------------------------
// 2a. FP64 split into S balanced-digit planes (+ exponents).
//     Grid covers the padded Rpad x Kpad extent in 64 x 64 tiles.
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(256) void split_f64_kernel(const double* __restrict__ P, int64_t sr, int64_t sk, int64_t R,
                                                        int64_t K, const unsigned long long* __restrict__ maxbits,
                                                        int* __restrict__ exps, uint16_t* __restrict__ planes,
                                                        int64_t plane_stride, int64_t Kpad, int* __restrict__ flag) {
  __shared__ double tile[64][65];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  const bool rows_contig = (sr == 1);
#pragma unroll 4
  for (int j = 0; j < 16; ++j) {
    const int r = rows_contig ? (t % 64) : (t / 64 + 4 * j);
    const int k = rows_contig ? (t / 64 + 4 * j) : (t % 64);
    tile[r][k] = (r0 + r < R && k0 + k < K) ? P[(r0 + r) * sr + (k0 + k) * sk] : 0.0;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int grp = t + 256 * q;  // 512 groups of 8 consecutive k
    const int row = grp / 8, kg = 8 * (grp % 8);
    int E = 0;
    if (r0 + row < R) {
      const unsigned long long mb = maxbits[r0 + row];
      if (mb >= 0x7FF0000000000000ull) {
        *flag = 1;  // Inf or NaN in this row/column
      } else {
        E = exponent_of_bits(mb);
      }
      if (blockIdx.y == 0 && kg == 0) exps[r0 + row] = E;
    }
    double res[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) res[e] = ldexp(tile[row][kg + e], -E) * 128.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      uint32_t w[4];
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        if (s > 0) {
          res[e] *= 256.0;
          res[e + 1] *= 256.0;
        }
        const double d0 = rint(res[e]), d1 = rint(res[e + 1]);
        res[e] -= d0;
        res[e + 1] -= d1;
        w[e / 2] = (uint32_t)bf16_bits_exact(d0) | ((uint32_t)bf16_bits_exact(d1) << 16);
      }
      u4 v = {w[0], w[1], w[2], w[3]};
      *reinterpret_cast<u4*>(planes + s * plane_stride + (r0 + row) * Kpad + k0 + kg) = v;
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
                                                        int64_t Kpad) {
  __shared__ float tile[64][65];
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
  const bool rows_contig = (sr == 1);
#pragma unroll 4
  for (int j = 0; j < 16; ++j) {
    const int r = rows_contig ? (t % 64) : (t / 64 + 4 * j);
    const int k = rows_contig ? (t / 64 + 4 * j) : (t % 64);
    tile[r][k] = (r0 + r < R && k0 + k < K) ? P[(r0 + r) * sr + (k0 + k) * sk] : 0.0f;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const int grp = t + 256 * q;
    const int row = grp / 8, kg = 8 * (grp % 8);
    uint32_t w[3][4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
      uint16_t b[3][2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        float r = tile[row][kg + e + h];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          b[s][h] = bf16_rn_bits(r);
          r = r - bf16_to_f(b[s][h]);
        }
      }
#pragma unroll
      for (int s = 0; s < 3; ++s) w[s][e / 2] = (uint32_t)b[s][0] | ((uint32_t)b[s][1] << 16);
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      u4 v = {w[s][0], w[s][1], w[s][2], w[s][3]};
      *reinterpret_cast<u4*>(planes + s * plane_stride + (r0 + row) * Kpad + k0 + kg) = v;
    }
  }
}

// ---------------------------------------------------------------------------
// 3. slice-pair GEMM.  MODE 0: emulated FP64 (S slices, exact level sums, FP64
//    groups).  MODE 1: 3xBF16 FP32 (S = 3, pairs t+u <= LMAX).
// ---------------------------------------------------------------------------
struct PairArgs {
  const uint16_t* pa;  // op(A) planes
  const uint16_t* pb;  // op(B) planes
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

template <int MODE, int S, int LMAX>
__global__ __launch_bounds__(ZNT, 1) void slice_pair_gemm_kernel(PairArgs p) {
  // one LDS stage: S planes x ZROWS rows x 16 k (2 chunks of 8 bf16 per row)
  constexpr int NCH = S * ZROWS * 2;             // 16-byte chunks per stage
  constexpr int PER = (NCH + ZNT - 1) / ZNT;     // chunks per thread
  constexpr int NL = (MODE == 0) ? S : 2;        // FP32 accumulator sets
  constexpr int NG = (S + 3) / 4;                // FP64 groups (MODE 0)
  constexpr int KC = chunk_len(S);
  __shared__ __attribute__((aligned(16))) u4 lds[2][NCH];

  const int nblocks = p.tiles_m * p.tiles_n;
  const int tile = xcd_remap(blockIdx.x, nblocks, 8);
  int tm, tn;
  tile_coords(tile, p.tiles_m, p.tiles_n, 8, tm, tn);
  const int64_t m0 = (int64_t)tm * ZBM, n0 = (int64_t)tn * ZBN;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int wm0 = 32 * (wave % 4), wn0 = 32 * (wave / 4);
  const int l32 = lane % 32, hl = lane / 32;

  // per-thread chunk sources (fixed for the whole k loop, advanced by 16 k per stage)
  const uint16_t* src[PER];
  int dst[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int c = tid + ZNT * i;
    const int cc = c < NCH ? c : 0;
    const int t = cc / (ZROWS * 2), rem = cc % (ZROWS * 2), row = rem / 2, half = rem % 2;
    src[i] = (row < ZBM) ? p.pa + t * p.strideA + (m0 + row) * p.Kpad + 8 * half
                         : p.pb + t * p.strideB + (n0 + row - ZBM) * p.Kpad + 8 * half;
    dst[i] = c < NCH ? c : -1;
  }
  u4 stg[PER];
  auto load_stage = [&](int64_t k0) {
#pragma unroll
    for (int i = 0; i < PER; ++i)
      if (dst[i] >= 0) stg[i] = *reinterpret_cast<const u4*>(src[i] + k0);
  };
  auto store_stage = [&](int buf) {
#pragma unroll
    for (int i = 0; i < PER; ++i)
      if (dst[i] >= 0) lds[buf][dst[i]] = stg[i];
  };

  f16v acc[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l)
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[l][e] = 0.0f;
  double G[NG > 0 ? NG : 1][16];
#pragma unroll
  for (int g = 0; g < NG; ++g)
#pragma unroll
    for (int e = 0; e < 16; ++e) G[g][e] = 0.0;

  auto flush = [&]() {
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const double sc = (l % 4 == 0) ? 1.0 : (l % 4 == 1) ? 0x1.0p-8 : (l % 4 == 2) ? 0x1.0p-16 : 0x1.0p-24;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        G[l / 4][e] = fma((double)acc[l][e], sc, G[l / 4][e]);
        acc[l][e] = 0.0f;
      }
    }
  };

  const int nst = (int)(p.Kpad / ZBK);
  constexpr int KC_ST = KC / ZBK;
  load_stage(0);
  store_stage(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int cur = st & 1;
    if (st + 1 < nst) load_stage((int64_t)(st + 1) * ZBK);
    const u4* L = lds[cur];
    bf8 bfr[S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
      const u4 v = L[u * ZROWS * 2 + (ZBM + wn0 + l32) * 2 + hl];
      bfr[u] = __builtin_bit_cast(bf8, v);
    }
#pragma unroll
    for (int t = 0; t < S; ++t) {
      const u4 v = L[t * ZROWS * 2 + (wm0 + l32) * 2 + hl];
      const bf8 afr = __builtin_bit_cast(bf8, v);
#pragma unroll
      for (int u = 0; u < S; ++u) {
        if (t + u + 2 > LMAX) continue;
        const int l = (MODE == 0) ? (t + u) : (t + u == 0 ? 0 : 1);
        // X = op(B) plane (p -> n), Y = op(A) plane (q -> m): lane l holds column-major
        // C[m0+wm0+l%32][n0+wn0+8(r/4)+4(l/32)+r%4] in item r.
        acc[l] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(bfr[u], afr, acc[l], 0, 0, 0);
      }
    }
    if (st + 1 < nst) store_stage(cur ^ 1);
    if (MODE == 0 && (((st + 1) % KC_ST) == 0 || st + 1 == nst)) flush();
    __syncthreads();
  }

  const int64_t m = m0 + wm0 + l32;
  if (MODE == 0) {
    double* C = static_cast<double*>(p.C);
    const int ea = m < p.M ? p.eA[m] : 0;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int64_t n = n0 + wn0 + 8 * (e / 4) + 4 * hl + (e % 4);
      if (m < p.M && n < p.N) {
        double hi = 0.0, lo = 0.0;
#pragma unroll
        for (int g = NG - 1; g >= 0; --g) {
          double err;
          two_sum(hi, ldexp(G[g][e], -32 * g), hi, err);
          lo += err;
        }
        const double v = ldexp(hi + lo, ea + p.eB[n] - 14);
        double c = p.alpha * v;
        if (p.beta != 0.0) c = fma(p.beta, C[m + n * p.ldc], c);
        C[m + n * p.ldc] = c;
      }
    }
  } else {
    float* C = static_cast<float*>(p.C);
    const float alpha = (float)p.alpha, beta = (float)p.beta;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int64_t n = n0 + wn0 + 8 * (e / 4) + 4 * hl + (e % 4);
      if (m < p.M && n < p.N) {
        float c = alpha * (acc[0][e] + acc[1][e]);
        if (beta != 0.0f) c = fmaf(beta, C[m + n * p.ldc], c);
        C[m + n * p.ldc] = c;
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
  size_t offA, offB, offMA, offMB, offEA, offEB, total;
};

Layout make_layout(int64_t m, int64_t n, int64_t k, int s) {
  Layout L;
  L.Mpad = rup(std::max<int64_t>(m, 1), 128);
  L.Npad = rup(std::max<int64_t>(n, 1), 128);
  L.Kpad = rup(std::max<int64_t>(k, 1), ZPAD_K);
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t off = 0;
  L.offA = off;
  off = al(off + (size_t)s * L.Mpad * L.Kpad * 2);
  L.offB = off;
  off = al(off + (size_t)s * L.Npad * L.Kpad * 2);
  L.offMA = off;
  off = al(off + (size_t)L.Mpad * 8);
  L.offMB = off;
  off = al(off + (size_t)L.Npad * 8);
  L.offEA = off;
  off = al(off + (size_t)L.Mpad * 4);
  L.offEB = off;
  off = al(off + (size_t)L.Npad * 4);
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
hipError_t launch_split_f64(const double* P, int64_t sr, int64_t sk, int64_t R, int64_t K, int64_t Rpad, int64_t Kpad,
                            unsigned long long* mb, int* ex, uint16_t* planes, int* flag, hipStream_t st) {
  const dim3 g((unsigned)(Rpad / 64), (unsigned)(Kpad / 64));
  hipLaunchKernelGGL(zero_u64_kernel, dim3(std::min<int64_t>((Rpad + 255) / 256, 1024)), dim3(256), 0, st, mb, Rpad);
  const dim3 gr((unsigned)((R + 63) / 64), (unsigned)((K + 63) / 64));
  hipLaunchKernelGGL(absmax_kernel<double>, gr, dim3(256), 0, st, P, sr, sk, R, K, mb);
  hipLaunchKernelGGL(split_f64_kernel<S>, g, dim3(256), 0, st, P, sr, sk, R, K, mb, ex, planes, Rpad * Kpad, Kpad, flag);
  return hipGetLastError();
}

template <int S>
hipError_t run_ozaki(const GemmArgs& g, const Layout& L, char* ws, int* flag, hipStream_t st) {
  int64_t sr, sk;
  auto* mbA = reinterpret_cast<unsigned long long*>(ws + L.offMA);
  auto* mbB = reinterpret_cast<unsigned long long*>(ws + L.offMB);
  auto* eA = reinterpret_cast<int*>(ws + L.offEA);
  auto* eB = reinterpret_cast<int*>(ws + L.offEB);
  auto* pA = reinterpret_cast<uint16_t*>(ws + L.offA);
  auto* pB = reinterpret_cast<uint16_t*>(ws + L.offB);
  strides_a(g.ta, g.lda, sr, sk);
  hipError_t e = launch_split_f64<S>(static_cast<const double*>(g.A), sr, sk, g.m, g.k, L.Mpad, L.Kpad, mbA, eA, pA, flag, st);
  if (e != hipSuccess) return e;
  strides_b(g.tb, g.ldb, sr, sk);
  e = launch_split_f64<S>(static_cast<const double*>(g.B), sr, sk, g.n, g.k, L.Npad, L.Kpad, mbB, eB, pB, flag, st);
  if (e != hipSuccess) return e;
  PairArgs p;
  p.pa = pA;
  p.pb = pB;
  p.Kpad = L.Kpad;
  p.strideA = L.Mpad * L.Kpad;
  p.strideB = L.Npad * L.Kpad;
  p.eA = eA;
  p.eB = eB;
  p.M = g.m;
  p.N = g.n;
  p.alpha = g.alpha;
  p.beta = g.beta;
  p.C = g.C;
  p.ldc = g.ldc;
  p.tiles_m = (int)((g.m + ZBM - 1) / ZBM);
  p.tiles_n = (int)((g.n + ZBN - 1) / ZBN);
  hipLaunchKernelGGL((slice_pair_gemm_kernel<0, S, S + 1>), dim3(p.tiles_m * p.tiles_n), dim3(ZNT), 0, st, p);
  return hipGetLastError();
}

}  // namespace

size_t ozaki_workspace_bytes(int64_t m, int64_t n, int64_t k, int s) { return make_layout(m, n, k, s).total; }

hipError_t launch_dgemm_ozaki(const GemmArgs& g, int s, void* ws, int* flag, hipStream_t st) {
  const Layout L = make_layout(g.m, g.n, g.k, s);
  char* w = static_cast<char*>(ws);
  switch (s) {
    case 1: return run_ozaki<1>(g, L, w, flag, st);
    case 2: return run_ozaki<2>(g, L, w, flag, st);
    case 3: return run_ozaki<3>(g, L, w, flag, st);
    case 4: return run_ozaki<4>(g, L, w, flag, st);
    case 5: return run_ozaki<5>(g, L, w, flag, st);
    case 6: return run_ozaki<6>(g, L, w, flag, st);
    case 7: return run_ozaki<7>(g, L, w, flag, st);
    case 8: return run_ozaki<8>(g, L, w, flag, st);
    default: return hipErrorInvalidValue;
  }
}

size_t bf16x3_workspace_bytes(int64_t m, int64_t n, int64_t k) { return make_layout(m, n, k, 3).total; }

hipError_t launch_sgemm_bf16x3(const GemmArgs& g, int nprod, void* ws, hipStream_t st) {
  const Layout L = make_layout(g.m, g.n, g.k, 3);
  char* w = static_cast<char*>(ws);
  auto* pA = reinterpret_cast<uint16_t*>(w + L.offA);
  auto* pB = reinterpret_cast<uint16_t*>(w + L.offB);
  int64_t sr, sk;
  strides_a(g.ta, g.lda, sr, sk);
  hipLaunchKernelGGL(split_f32_kernel, dim3((unsigned)(L.Mpad / 64), (unsigned)(L.Kpad / 64)), dim3(256), 0, st,
                     static_cast<const float*>(g.A), sr, sk, g.m, g.k, pA, L.Mpad * L.Kpad, L.Kpad);
  strides_b(g.tb, g.ldb, sr, sk);
  hipLaunchKernelGGL(split_f32_kernel, dim3((unsigned)(L.Npad / 64), (unsigned)(L.Kpad / 64)), dim3(256), 0, st,
                     static_cast<const float*>(g.B), sr, sk, g.n, g.k, pB, L.Npad * L.Kpad, L.Kpad);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  PairArgs p;
  p.pa = pA;
  p.pb = pB;
  p.Kpad = L.Kpad;
  p.strideA = L.Mpad * L.Kpad;
  p.strideB = L.Npad * L.Kpad;
  p.eA = nullptr;
  p.eB = nullptr;
  p.M = g.m;
  p.N = g.n;
  p.alpha = g.alpha;
  p.beta = g.beta;
  p.C = g.C;
  p.ldc = g.ldc;
  p.tiles_m = (int)((g.m + ZBM - 1) / ZBM);
  p.tiles_n = (int)((g.n + ZBN - 1) / ZBN);
  const dim3 grid(p.tiles_m * p.tiles_n);
  if (nprod == 3) hipLaunchKernelGGL((slice_pair_gemm_kernel<1, 3, 3>), grid, dim3(ZNT), 0, st, p);
  else if (nprod == 6) hipLaunchKernelGGL((slice_pair_gemm_kernel<1, 3, 4>), grid, dim3(ZNT), 0, st, p);
  else hipLaunchKernelGGL((slice_pair_gemm_kernel<1, 3, 6>), grid, dim3(ZNT), 0, st, p);
  return hipGetLastError();
}

}  // namespace jb
