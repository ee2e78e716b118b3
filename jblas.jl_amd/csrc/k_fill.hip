// k_fill.hip -- seeded device-side input generator.
//
// Replaces mrandn/rand! (src/randmat.jl:5-14, unseeded normals drawn on the
// host) with a counter-based SplitMix64 stream so that 8192^2 operands are
// produced directly in HBM, bit-identical to the host generator: output i is
// mix(seed + (i+1)*0x9e3779b97f4a7c15), and the conversions below are exact.
#include "jb_internal.h"

namespace jb {
namespace {

__device__ __forceinline__ uint64_t sm_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t sm_at(uint64_t seed, uint64_t i) {
  return sm_mix(seed + (i + 1) * 0x9e3779b97f4a7c15ULL);
}
__device__ __forceinline__ double unit_uniform(uint64_t x) {  // [-1, 1)
  return 2.0 * ((double)(x >> 11) * 0x1.0p-53) - 1.0;
}

__global__ void fill_uniform_kernel(uint64_t seed, int64_t n, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = unit_uniform(sm_at(seed, (uint64_t)i));
}
__global__ void fill_uniform_f32_kernel(uint64_t seed, int64_t n, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)unit_uniform(sm_at(seed, (uint64_t)i));
}
__global__ void fill_wide_kernel(uint64_t seed, int64_t n, int emin, int emax, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double u = unit_uniform(sm_at(seed, 2 * (uint64_t)i));
    const double v = (double)(sm_at(seed, 2 * (uint64_t)i + 1) >> 11) * 0x1.0p-53;
    const int e = emin + (int)floor((double)(emax - emin + 1) * v);
    out[i] = ldexp(u, e);
  }
}

__global__ void fill_wide_f32_kernel(uint64_t seed, int64_t n, int emin, int emax, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double u = unit_uniform(sm_at(seed, 2 * (uint64_t)i));
    const double v = (double)(sm_at(seed, 2 * (uint64_t)i + 1) >> 11) * 0x1.0p-53;
    const int e = emin + (int)floor((double)(emax - emin + 1) * v);
    out[i] = (float)ldexp(u, e);  // the double of the FP64 stream, rounded to nearest
  }
}
// Normals by Box-Muller on the same stream (the draw of mrandn, src/randmat.jl:5-14):
// z = sqrt(-2 ln u1) cos(2 pi u2), u1 = ((x_2i >> 11) + 1) 2^-53 in (0, 1],
// u2 = (x_2i+1 >> 11) 2^-53 in [0, 1); out = scale * z.  log and cos are the
// device's, so values agree with the host generator to a few ulp, not bitwise.
__device__ __forceinline__ double box_muller(uint64_t seed, uint64_t i) {
  const double u1 = ((double)(sm_at(seed, 2 * i) >> 11) + 1.0) * 0x1.0p-53;
  const double u2 = (double)(sm_at(seed, 2 * i + 1) >> 11) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);  // the double nearest 2 pi, times u2
}
__global__ void fill_normal_kernel(uint64_t seed, int64_t n, double scale, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = scale * box_muller(seed, (uint64_t)i);
}
__global__ void fill_normal_f32_kernel(uint64_t seed, int64_t n, double scale, float* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (float)(scale * box_muller(seed, (uint64_t)i));
}

int grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (int)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace

hipError_t launch_fill_uniform(uint64_t seed, int64_t n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(fill_uniform_kernel, dim3(grid_for(n)), dim3(256), 0, s, seed, n, out);
  return hipGetLastError();
}
hipError_t launch_fill_uniform_f32(uint64_t seed, int64_t n, float* out, hipStream_t s) {
  hipLaunchKernelGGL(fill_uniform_f32_kernel, dim3(grid_for(n)), dim3(256), 0, s, seed, n, out);
  return hipGetLastError();
}
hipError_t launch_fill_wide(uint64_t seed, int64_t n, int emin, int emax, double* out, hipStream_t s) {
  hipLaunchKernelGGL(fill_wide_kernel, dim3(grid_for(n)), dim3(256), 0, s, seed, n, emin, emax, out);
  return hipGetLastError();
}

hipError_t launch_fill_wide_f32(uint64_t seed, int64_t n, int emin, int emax, float* out, hipStream_t s) {
  hipLaunchKernelGGL(fill_wide_f32_kernel, dim3(grid_for(n)), dim3(256), 0, s, seed, n, emin, emax, out);
  return hipGetLastError();
}
hipError_t launch_fill_normal(uint64_t seed, int64_t n, double scale, double* out, hipStream_t s) {
  hipLaunchKernelGGL(fill_normal_kernel, dim3(grid_for(n)), dim3(256), 0, s, seed, n, scale, out);
  return hipGetLastError();
}
hipError_t launch_fill_normal_f32(uint64_t seed, int64_t n, double scale, float* out, hipStream_t s) {
  hipLaunchKernelGGL(fill_normal_f32_kernel, dim3(grid_for(n)), dim3(256), 0, s, seed, n, scale, out);
  return hipGetLastError();
}

}  // namespace jb
