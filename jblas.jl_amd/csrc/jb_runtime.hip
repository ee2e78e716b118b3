// jb_runtime.hip -- host runtime behind include/jblas_hip.h.
//
// Responsibilities: BLAS-style argument validation (before any HIP call),
// library state and device-property query (the GPU replacement of the
// build-time CPU probe deps/build.jl:1-38 -> src/cpu_info.jl), a stream-ordered
// device workspace pool, host<->device staging for the host-pointer entry
// points, and dispatch to the kernels' launchers.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/jblas_hip.h"
#include "jb_internal.h"
#include "jb_runtime.h"

namespace jb {

// ---- scratch of the allocation-free emulated entry point -------------------
// One array per device, allocated by the code-object loader and held until the
// process ends.  Its first JB_SCRATCH_HEADER bytes hold the sticky non-finite
// word of the allocation-free calls; the emulated workspace follows.  Host
// code never uses the array's name as an address: each device's address comes
// from hipGetSymbolAddress at init (g_scratch_ptr).
constexpr size_t JB_SCRATCH_HEADER = 256;
constexpr int JB_MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[JB_SCRATCH_HEADER + JB_NOALLOC_WORKSPACE_BYTES];
static void* g_scratch_ptr[JB_MAX_DEVICES];  // per device, set once under the runtime lock
static thread_local bool tl_no_timing = false;  // allocation-free calls create no timing events

Runtime& rt() {
  static Runtime r;
  return r;
}

static thread_local std::string tl_error;

// ---- JB_LOG ------------------------------------------------------------------
// Quiet unless JB_LOG is set: "1" or "stderr" writes to stderr, any other
// non-empty value other than "0" names a file that lines are appended to.
// One line per event ("jblas: " prefix, flushed): errors returned by any entry
// point, device init, device buffer growth, grouped plans, tuner decisions.
static std::mutex g_log_mu;
static FILE* log_stream() {
  static FILE* f = []() -> FILE* {
    const char* v = getenv("JB_LOG");
    if (!v || !*v || strcmp(v, "0") == 0) return nullptr;
    if (strcmp(v, "1") == 0 || strcmp(v, "stderr") == 0) return stderr;
    FILE* fp = fopen(v, "a");
    return fp ? fp : stderr;
  }();
  return f;
}
bool log_enabled() { return log_stream() != nullptr; }
void log_line(const char* fmt, ...) {
  FILE* f = log_stream();
  if (!f) return;
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  std::lock_guard<std::mutex> lk(g_log_mu);
  fprintf(f, "jblas: %s\n", buf);
  fflush(f);
}

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  tl_error = buf;
  log_line("error %d: %s", code, buf);
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(JB_ERR_HIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}

static bool valid_trans(char t) { return t == 'N' || t == 'n' || t == 'T' || t == 't' || t == 'C' || t == 'c'; }
static bool is_trans(char t) { return !(t == 'N' || t == 'n'); }

int validate_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, bool alpha_zero, const void* A, int64_t lda,
                  const void* B, int64_t ldb, const void* C, int64_t ldc, GemmArgs* out) {
  tl_error.clear();
  if (!valid_trans(ta)) return fail(-1, "transA must be 'N', 'T' or 'C' (got %d)", (int)ta);
  if (!valid_trans(tb)) return fail(-2, "transB must be 'N', 'T' or 'C' (got %d)", (int)tb);
  if (m < 0) return fail(-3, "m < 0");
  if (n < 0) return fail(-4, "n < 0");
  if (k < 0) return fail(-5, "k < 0");
  const int64_t arows = is_trans(ta) ? k : m, brows = is_trans(tb) ? n : k;
  const bool uses_ab = m > 0 && n > 0 && k > 0 && !alpha_zero;
  if (uses_ab && A == nullptr) return fail(-7, "A is NULL");
  if (lda < (arows > 1 ? arows : 1)) return fail(-8, "lda < max(1, %lld)", (long long)arows);
  if (uses_ab && B == nullptr) return fail(-9, "B is NULL");
  if (ldb < (brows > 1 ? brows : 1)) return fail(-10, "ldb < max(1, %lld)", (long long)brows);
  if (m > 0 && n > 0 && C == nullptr) return fail(-12, "C is NULL");
  if (ldc < (m > 1 ? m : 1)) return fail(-13, "ldc < max(1, %lld)", (long long)m);
  // Tile-grid sizes are int32 inside the kernels, and a launch grid holds fewer
  // than 2^32 work-items: the output may have at most 2^23 tiles of 64 x 64
  // (2^35 elements, more than one GPU's memory holds in FP64).
  if (m > (int64_t)1 << 30 || n > (int64_t)1 << 30) return fail(-3, "m or n above 2^30 is not supported");
  if (((m + 63) / 64) * ((n + 63) / 64) >= ((int64_t)1 << 23))
    return fail(JB_ERR_UNSUPPORTED, "C of %lld x %lld is too large for one launch", (long long)m, (long long)n);
  out->ta = is_trans(ta);
  out->tb = is_trans(tb);
  out->m = m;
  out->n = n;
  out->k = k;
  out->A = A;
  out->lda = lda;
  out->B = B;
  out->ldb = ldb;
  out->C = const_cast<void*>(C);
  out->ldc = ldc;
  return JB_OK;
}

int ensure_init() {
  if (rt().inited) return JB_OK;
  return init_locked(0);
}

int init_locked(int device) {
  Runtime& r = rt();
  if (r.inited && r.device == device) return JB_OK;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(JB_ERR_HIP, "no HIP device available");
  if (device < 0 || device >= count) return fail(-1, "device %d out of range [0,%d)", device, count);
  if ((e = hipSetDevice(device)) != hipSuccess) return hip_fail(e, "hipSetDevice");
  if (device < JB_MAX_DEVICES && g_scratch_ptr[device] == nullptr) {
    void* sp = nullptr;
    if ((e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch))) != hipSuccess)
      return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = sp;
  }
  hipDeviceProp_t p;
  if ((e = hipGetDeviceProperties(&p, device)) != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
  jb_device_info& d = r.info;
  memset(&d, 0, sizeof d);
  snprintf(d.name, sizeof d.name, "%s", p.name);
  snprintf(d.arch, sizeof d.arch, "%s", p.gcnArchName);
  d.compute_units = p.multiProcessorCount;
  d.clock_khz = p.clockRate;
  d.wavefront_size = p.warpSize;
  d.lds_bytes_per_cu = (int32_t)p.maxSharedMemoryPerMultiProcessor;
  d.l2_bytes = p.l2CacheSize;
  d.hbm_bytes = (int64_t)p.totalGlobalMem;
  const double cu_ghz = (double)p.multiProcessorCount * (double)p.clockRate * 1e-6;  // CU * GHz
  d.peak_fp64_tflops = cu_ghz * 128.0 * 1e-3;
  d.peak_fp32_tflops = cu_ghz * 256.0 * 1e-3;
  d.peak_bf16_tflops = cu_ghz * 4096.0 * 1e-3;
  if (r.stream == nullptr) {
    if ((e = hipStreamCreateWithFlags(&r.stream, hipStreamNonBlocking)) != hipSuccess)
      return hip_fail(e, "hipStreamCreate");
  }
  if (r.ws_event == nullptr) {
    if ((e = hipEventCreateWithFlags(&r.ws_event, hipEventDisableTiming)) != hipSuccess)
      return hip_fail(e, "hipEventCreate");
  }
  if (r.sticky == nullptr) {
    if ((e = hipMalloc(&r.sticky, 256)) != hipSuccess) {
      r.sticky = nullptr;
      return fail(JB_ERR_ALLOC, "non-finite flag allocation failed: %s", hipGetErrorString(e));
    }
    if ((e = hipMemset(r.sticky, 0, 256)) != hipSuccess) return hip_fail(e, "clear the non-finite flag");
  }
  r.device = device;
  r.inited = true;
  log_line("init device %d: %s %s, %d CUs, %.2f GHz, LDS %d B/CU, L2 %lld B, HBM %lld B", device, d.name, d.arch,
           d.compute_units, d.clock_khz * 1e-6, d.lds_bytes_per_cu, (long long)d.l2_bytes, (long long)d.hbm_bytes);
  const char* tf = getenv("JB_TUNE_FILE");
  if (tf && *tf) tune_load_locked(tf);  // a missing or corrupt table leaves the static defaults
  return JB_OK;
}

// ---- device buffers --------------------------------------------------------
int Buffer::reserve(size_t bytes, const char* what) {
  if (bytes <= size) return JB_OK;
  const char* fault = getenv("JB_FAULT_ALLOC");
  if (fault && *fault == '1') return fail(JB_ERR_ALLOC, "%s: allocation of %zu bytes refused (JB_FAULT_ALLOC)", what, bytes);
  if (ptr) {
    hipDeviceSynchronize();  // earlier work may still use the old buffer
    hipFree(ptr);
    ptr = nullptr;
    size = 0;
  }
  size_t want = bytes + (bytes >> 3);  // grow with headroom
  hipError_t e = hipMalloc(&ptr, want);
  if (e != hipSuccess) {
    ptr = nullptr;
    hipGetLastError();
    return fail(JB_ERR_ALLOC, "%s: hipMalloc(%zu) failed: %s", what, want, hipGetErrorString(e));
  }
  size = want;
  log_line("buffer %s grown to %zu bytes", what, want);
  return JB_OK;
}

void Buffer::release() {
  if (ptr) hipFree(ptr);
  ptr = nullptr;
  size = 0;
}

static bool capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// Stream-ordered use of the shared workspace: a call waits for the previous
// user of the workspace (on any stream) before touching it.  While `s` is being
// captured into a graph the workspace must already be large enough (a graph
// keeps its address; growing it would free memory that the graph uses), and
// the captured work is not ordered against eager calls: the caller orders the
// graph's replays with other calls of the library.
int ws_acquire(size_t bytes, hipStream_t s, void** out) {
  Runtime& r = rt();
  if (capturing(s)) {
    if (bytes > r.ws.size)
      return fail(JB_ERR_STATE,
                  "the workspace cannot grow during graph capture: make one call of this size (%zu bytes of "
                  "workspace) outside the capture first",
                  bytes);
    r.ws_in_graph = true;
    *out = r.ws.ptr;
    return JB_OK;
  }
  if (bytes > r.ws.size && r.ws_in_graph) {
    // a captured graph refers to the current buffer: keep it until jb_finalize
    r.retired.push_back(r.ws.ptr);
    r.ws.ptr = nullptr;
    r.ws.size = 0;
    r.ws_in_graph = false;
  }
  int rc = r.ws.reserve(bytes, "workspace");
  if (rc) return rc;
  if (r.ws_pending) {
    hipError_t e = hipStreamWaitEvent(s, r.ws_event, 0);
    if (e != hipSuccess) return hip_fail(e, "hipStreamWaitEvent");
  }
  *out = r.ws.ptr;
  return JB_OK;
}

int ws_release(hipStream_t s) {
  Runtime& r = rt();
  if (capturing(s)) return JB_OK;
  hipError_t e = hipEventRecord(r.ws_event, s);
  if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
  r.ws_pending = true;
  return JB_OK;
}

// ---- host staging ----------------------------------------------------------
// Copies a rows x cols column-major host matrix (leading dim ld) to device
// memory with the same leading dimension.
// A dense matrix (ld == rows) is one contiguous block: one 1-D copy (a 2-D copy
// of pageable memory is carried out row by row).
static hipError_t h2d(void* dst, const void* src, int64_t rows, int64_t cols, int64_t ld, size_t elt, hipStream_t s) {
  if (rows == 0 || cols == 0) return hipSuccess;
  if (ld == rows || cols == 1)
    return hipMemcpyAsync(dst, src, (size_t)(ld * (cols - 1) + rows) * elt, hipMemcpyHostToDevice, s);
  return hipMemcpy2DAsync(dst, ld * elt, src, ld * elt, rows * elt, cols, hipMemcpyHostToDevice, s);
}
static hipError_t d2h(void* dst, const void* src, int64_t rows, int64_t cols, int64_t ld, size_t elt, hipStream_t s) {
  if (rows == 0 || cols == 0) return hipSuccess;
  if (ld == rows || cols == 1)
    return hipMemcpyAsync(dst, src, (size_t)(ld * (cols - 1) + rows) * elt, hipMemcpyDeviceToHost, s);
  return hipMemcpy2DAsync(dst, ld * elt, src, ld * elt, rows * elt, cols, hipMemcpyDeviceToHost, s);
}
static size_t span_bytes(int64_t rows, int64_t cols, int64_t ld, size_t elt) {
  if (rows == 0 || cols == 0) return 0;
  return (size_t)(ld * (cols - 1) + rows) * elt;
}
static size_t round_up(size_t x) { return (x + 255) & ~(size_t)255; }

// Runs `dev` (a device-pointer GEMM) on staged copies of host operands.  With
// alpha == 0 or k == 0 the device GEMM reads neither A nor B (C = beta*C), so
// they are not staged: the caller may pass NULL or a short buffer then.
int run_staged(const GemmArgs& g, size_t elt, bool read_c, const std::function<int(const GemmArgs&, hipStream_t)>& dev) {
  Runtime& r = rt();
  const bool uses_ab = g.alpha != 0.0 && g.k > 0;
  const int64_t arows = g.ta ? g.k : g.m, acols = g.ta ? g.m : g.k;
  const int64_t brows = g.tb ? g.n : g.k, bcols = g.tb ? g.k : g.n;
  const size_t sa = uses_ab ? round_up(span_bytes(arows, acols, g.lda, elt)) : 0;
  const size_t sb = uses_ab ? round_up(span_bytes(brows, bcols, g.ldb, elt)) : 0;
  const size_t sc = round_up(span_bytes(g.m, g.n, g.ldc, elt));
  int rc = r.staging.reserve(sa + sb + sc + 256, "staging");
  if (rc) return rc;
  char* base = static_cast<char*>(r.staging.ptr);
  GemmArgs d = g;
  d.A = base;
  d.B = base + sa;
  d.C = base + sa + sb;
  hipStream_t s = r.stream;
  hipError_t e;
  if (uses_ab && (e = h2d(base, g.A, arows, acols, g.lda, elt, s)) != hipSuccess) return hip_fail(e, "copy A to device");
  if (uses_ab && (e = h2d(base + sa, g.B, brows, bcols, g.ldb, elt, s)) != hipSuccess)
    return hip_fail(e, "copy B to device");
  if (read_c && (e = h2d(d.C, g.C, g.m, g.n, g.ldc, elt, s)) != hipSuccess) return hip_fail(e, "copy C to device");
  rc = dev(d, s);
  if (rc) {
    hipStreamSynchronize(s);
    return rc;
  }
  if ((e = d2h(g.C, d.C, g.m, g.n, g.ldc, elt, s)) != hipSuccess) return hip_fail(e, "copy C to host");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
  return JB_OK;
}

void finalize_locked() {
  Runtime& r = rt();
  if (!r.inited) return;
  hipDeviceSynchronize();
  grouped_release_all();
  r.ws.release();
  for (void* p : r.retired) hipFree(p);
  r.retired.clear();
  r.ws_in_graph = false;
  r.staging.release();
  if (r.sticky) {
    hipFree(r.sticky);
    r.sticky = nullptr;
  }
  if (r.ws_event) hipEventDestroy(r.ws_event);
  r.ws_event = nullptr;
  r.ws_pending = false;
  if (r.stream) hipStreamDestroy(r.stream);
  r.stream = nullptr;
  r.inited = false;
}

// ---- per-kernel timing -------------------------------------------------------
struct TimingRec {
  std::string name;
  hipEvent_t a = nullptr, b = nullptr;
};
static bool g_timing = false;
static std::vector<TimingRec> g_pending;
static std::map<std::string, std::pair<double, int64_t>> g_totals;

int runtime_compute_units() { return rt().info.compute_units; }

void timing_begin(hipStream_t s, const char* kernel) {
  if (!g_timing || tl_no_timing || capturing(s)) return;
  TimingRec r;
  r.name = kernel;
  if (hipEventCreate(&r.a) != hipSuccess) return;
  hipEventRecord(r.a, s);
  g_pending.push_back(r);
}

void timing_end(hipStream_t s, const char* kernel) {
  if (!g_timing || tl_no_timing || capturing(s)) return;
  for (auto it = g_pending.rbegin(); it != g_pending.rend(); ++it) {
    if (it->b == nullptr && it->name == kernel) {
      if (hipEventCreate(&it->b) == hipSuccess) hipEventRecord(it->b, s);
      return;
    }
  }
}

static void timing_collect() {
  for (auto& r : g_pending) {
    if (r.a && r.b && hipEventSynchronize(r.b) == hipSuccess) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        auto& t = g_totals[r.name];
        t.first += ms;
        t.second += 1;
      }
    }
    if (r.a) hipEventDestroy(r.a);
    if (r.b) hipEventDestroy(r.b);
  }
  g_pending.clear();
}


}  // namespace jb

using namespace jb;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int jb_version(void) { return 10000; }

const char* jb_last_error(void) { return tl_error.c_str(); }

int jb_init(int device) {
  std::lock_guard<std::mutex> lk(rt().mu);
  tl_error.clear();
  if (rt().inited && rt().device != device) finalize_locked();
  return init_locked(device);
}

int jb_finalize(void) {
  std::lock_guard<std::mutex> lk(rt().mu);
  finalize_locked();
  return JB_OK;
}

int jb_kernel_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(rt().mu);
  tl_error.clear();
  g_timing = on != 0;
  return JB_OK;
}

int jb_kernel_timing_read(const char* kernel, double* total_ms, int64_t* launches) {
  if (!kernel) return fail(-1, "kernel is NULL");
  if (!total_ms) return fail(-2, "total_ms is NULL");
  if (!launches) return fail(-3, "launches is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  tl_error.clear();
  timing_collect();
  auto it = g_totals.find(kernel);
  *total_ms = it == g_totals.end() ? 0.0 : it->second.first;
  *launches = it == g_totals.end() ? 0 : it->second.second;
  return JB_OK;
}

int jb_kernel_timing_reset(void) {
  std::lock_guard<std::mutex> lk(rt().mu);
  timing_collect();
  g_totals.clear();
  return JB_OK;
}

int jb_get_device_info(jb_device_info* out) {
  if (!out) return fail(-1, "out is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  int rc = ensure_init();
  if (rc) return rc;
  *out = rt().info;
  return JB_OK;
}

// ---- FP64 native / recurrence ---------------------------------------------
static int dgemm_dev_locked(const GemmArgs& g, bool recurrence, hipStream_t s) {
  if (g.m == 0 || g.n == 0) return JB_OK;
  hipError_t e;
  if (g.alpha == 0.0 || g.k == 0) {
    e = launch_scale_f64(g.m, g.n, g.beta, static_cast<double*>(g.C), g.ldc, s);
  } else {
    e = recurrence ? launch_dgemm_recurrence(g, s) : launch_dgemm_mfma(g, s);
  }
  return e == hipSuccess ? JB_OK : hip_fail(e, recurrence ? "dgemm_recurrence launch" : "dgemm_mfma launch");
}

#define JB_GEMM_PROLOGUE(ALPHA, BETA)                                                                   \
  GemmArgs g;                                                                                            \
  int rc = validate_gemm(transA, transB, m, n, k, (ALPHA) == 0, A, lda, B, ldb, C, ldc, &g);            \
  if (rc) return rc;                                                                                     \
  g.alpha = (double)(ALPHA);                                                                             \
  g.beta = (double)(BETA);                                                                               \
  std::lock_guard<std::mutex> lk(rt().mu);                                                              \
  if (g.m == 0 || g.n == 0) return JB_OK;                                                                \
  rc = ensure_init();                                                                                    \
  if (rc) return rc;

int jb_dgemm_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                 int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, void* stream) {
  JB_GEMM_PROLOGUE(alpha, beta)
  return dgemm_dev_locked(g, false, (hipStream_t)stream);
}

int jb_dgemm(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
             const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  JB_GEMM_PROLOGUE(alpha, beta)
  return run_staged(g, sizeof(double), beta != 0.0, [](const GemmArgs& d, hipStream_t s) {
    return dgemm_dev_locked(d, false, s);
  });
}

int jb_dgemm_recurrence_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                            const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                            int64_t ldc, void* stream) {
  JB_GEMM_PROLOGUE(alpha, beta)
  return dgemm_dev_locked(g, true, (hipStream_t)stream);
}

int jb_dgemm_recurrence(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                        int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc) {
  JB_GEMM_PROLOGUE(alpha, beta)
  return run_staged(g, sizeof(double), beta != 0.0, [](const GemmArgs& d, hipStream_t s) {
    return dgemm_dev_locked(d, true, s);
  });
}

static int jmul_validate(int64_t M, int64_t N, int64_t P, const void* D, const void* A, const void* X) {
  tl_error.clear();
  if (M < 0) return fail(-1, "M < 0");
  if (N < 1) return fail(-2, "N < 1 (jmul! needs at least one k-step, src/gemm.jl:86)");
  if (P < 0) return fail(-3, "P < 0");
  if (M > 0 && P > 0 && D == nullptr) return fail(-4, "D is NULL");
  if (M > 0 && A == nullptr) return fail(-5, "A is NULL");
  if (P > 0 && X == nullptr) return fail(-6, "X is NULL");
  return JB_OK;
}

// jmul!(D, A, X) == gemm('N', 'N', 1, A, X, 0, D) on the FP64 MFMA kernel: its
// per-element chains are the reference recurrence bit for bit (plain first
// product, fma k-steps in increasing k, signed zeros included); N >= 1, so
// the chain is never empty.
int jb_jmul_dev(int64_t M, int64_t N, int64_t P, double* D, const double* A, const double* X, void* stream) {
  int rc = jmul_validate(M, N, P, D, A, X);
  if (rc) return rc;
  return jb_dgemm_dev('N', 'N', M, P, N, 1.0, A, M > 1 ? M : 1, X, N, 0.0, D, M > 1 ? M : 1, stream);
}

int jb_jmul(int64_t M, int64_t N, int64_t P, double* D, const double* A, const double* X) {
  int rc = jmul_validate(M, N, P, D, A, X);
  if (rc) return rc;
  return jb_dgemm('N', 'N', M, P, N, 1.0, A, M > 1 ? M : 1, X, N, 0.0, D, M > 1 ? M : 1);
}

// jmul! in Float32 (the reference's jmul! is generic in T, src/gemm.jl:244):
// the same recurrence in binary32 (plain first product, then fmaf in
// increasing k) on the sequential-fma kernel, bit-identical.
extern "C++" {
static int jmul_f32_locked(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X, void* stream,
                           bool host) {
  GemmArgs g{};
  g.m = M;
  g.n = P;
  g.k = N;
  g.alpha = 1.0;
  g.beta = 0.0;
  g.A = A;
  g.lda = M > 1 ? M : 1;
  g.B = X;
  g.ldb = N;
  g.C = D;
  g.ldc = M > 1 ? M : 1;
  auto run = [](const GemmArgs& d, hipStream_t s) {
    const hipError_t e = launch_sgemm_recurrence(d, s);
    return e == hipSuccess ? JB_OK : hip_fail(e, "sgemm_recurrence launch");
  };
  if (!host) return run(g, (hipStream_t)stream);
  return run_staged(g, sizeof(float), false, run);
}

static int jmul_f32_entry(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X, void* stream,
                          bool host) {
  int rc = jmul_validate(M, N, P, D, A, X);
  if (rc) return rc;
  if (M > (int64_t)1 << 30 || P > (int64_t)1 << 30) return fail(-1, "M or P above 2^30 is not supported");
  if (((M + 63) / 64) * ((P + 63) / 64) >= ((int64_t)1 << 23))
    return fail(JB_ERR_UNSUPPORTED, "D of %lld x %lld is too large for one launch", (long long)M, (long long)P);
  std::lock_guard<std::mutex> lk(rt().mu);
  if (M == 0 || P == 0) return JB_OK;
  if ((rc = ensure_init())) return rc;
  return jmul_f32_locked(M, N, P, D, A, X, stream, host);
}
}  // extern "C++"

int jb_jmul_f32(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X) {
  return jmul_f32_entry(M, N, P, D, A, X, nullptr, true);
}

int jb_jmul_f32_dev(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X, void* stream) {
  return jmul_f32_entry(M, N, P, D, A, X, stream, false);
}

// ---- kernel! / initkernel! (strided microkernels) --------------------------
static int kernel_validate(bool init, int64_t M, int64_t N, int64_t P, const void* D, int64_t ldd, const void* A,
                           int64_t lda, const void* X, int64_t ldx, GemmArgs* g) {
  tl_error.clear();
  if (M < 0) return fail(-1, "M < 0");
  if (init ? N < 1 : N < 0) return fail(-2, init ? "N < 1 (initkernel! starts with a product)" : "N < 0");
  if (P < 0) return fail(-3, "P < 0");
  if (M > 0 && P > 0 && D == nullptr) return fail(-4, "D is NULL");
  if (ldd < (M > 1 ? M : 1)) return fail(-5, "ldd < max(1, M)");
  if (M > 0 && N > 0 && A == nullptr) return fail(-6, "A is NULL");
  if (lda < (M > 1 ? M : 1)) return fail(-7, "lda < max(1, M)");
  if (P > 0 && N > 0 && X == nullptr) return fail(-8, "X is NULL");
  if (ldx < (N > 1 ? N : 1)) return fail(-9, "ldx < max(1, N)");
  if (M > (int64_t)1 << 30 || P > (int64_t)1 << 30) return fail(-1, "M or P above 2^30 is not supported");
  if (((M + 63) / 64) * ((P + 63) / 64) >= ((int64_t)1 << 23))
    return fail(JB_ERR_UNSUPPORTED, "D of %lld x %lld is too large for one launch", (long long)M, (long long)P);
  *g = GemmArgs{};
  g->m = M;
  g->n = P;
  g->k = N;
  g->alpha = 1.0;
  g->beta = init ? 0.0 : 1.0;
  g->A = A;
  g->lda = lda;
  g->B = X;
  g->ldb = ldx;
  g->C = const_cast<void*>(D);
  g->ldc = ldd;
  return JB_OK;
}

// kernel! / initkernel! on the FP64 MFMA kernel, whose chains are the
// reference's bit for bit: initkernel! is D = A*X (chains from -0.0, i.e. a
// plain first product), kernel! starts each chain from D.
static int kernel_dev_locked(const GemmArgs& g, bool init, hipStream_t s) {
  if (g.m == 0 || g.n == 0 || (!init && g.k == 0)) return JB_OK;
  GemmArgs h = g;
  h.alpha = 1.0;
  h.beta = 0.0;
  const hipError_t e = init ? launch_dgemm_mfma(h, s) : launch_dgemm_mfma_acc(h, s);
  return e == hipSuccess ? JB_OK : hip_fail(e, init ? "initkernel launch" : "kernel launch");
}

extern "C++" {
static int kernel_entry(bool init, int64_t M, int64_t N, int64_t P, double* D, int64_t ldd, const double* A,
                        int64_t lda, const double* X, int64_t ldx, void* stream, bool host) {
  GemmArgs g;
  int rc = kernel_validate(init, M, N, P, D, ldd, A, lda, X, ldx, &g);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(rt().mu);
  if (M == 0 || P == 0 || (!init && N == 0)) return JB_OK;
  if ((rc = ensure_init())) return rc;
  if (!host) return kernel_dev_locked(g, init, (hipStream_t)stream);
  return run_staged(g, sizeof(double), !init, [init](const GemmArgs& d, hipStream_t s) {
    return kernel_dev_locked(d, init, s);
  });
}
}  // extern "C++"

int jb_kernel(int64_t M, int64_t N, int64_t P, double* D, int64_t ldd, const double* A, int64_t lda, const double* X,
              int64_t ldx) {
  return kernel_entry(false, M, N, P, D, ldd, A, lda, X, ldx, nullptr, true);
}
int jb_kernel_dev(int64_t M, int64_t N, int64_t P, double* D, int64_t ldd, const double* A, int64_t lda,
                  const double* X, int64_t ldx, void* stream) {
  return kernel_entry(false, M, N, P, D, ldd, A, lda, X, ldx, stream, false);
}
int jb_initkernel(int64_t M, int64_t N, int64_t P, double* D, int64_t ldd, const double* A, int64_t lda,
                  const double* X, int64_t ldx) {
  return kernel_entry(true, M, N, P, D, ldd, A, lda, X, ldx, nullptr, true);
}
int jb_initkernel_dev(int64_t M, int64_t N, int64_t P, double* D, int64_t ldd, const double* A, int64_t lda,
                      const double* X, int64_t ldx, void* stream) {
  return kernel_entry(true, M, N, P, D, ldd, A, lda, X, ldx, stream, false);
}

// ---- fastmul! (whole-matrix small GEMM) -----------------------------------
extern "C++" {
template <typename T>
static int fastmul_dev_locked(int64_t M, int64_t N, int64_t P, T* D, const T* A, const T* X, hipStream_t s) {
  hipError_t e = sizeof(T) == 8
                     ? launch_fastmul_f64(M, N, P, (double*)D, (const double*)A, (const double*)X, s)
                     : launch_fastmul_f32(M, N, P, (float*)D, (const float*)A, (const float*)X, s);
  return e == hipSuccess ? JB_OK : hip_fail(e, "fastmul launch");
}

template <typename T>
static int fastmul_entry(int64_t M, int64_t N, int64_t P, T* D, const T* A, const T* X, void* stream, bool host) {
  int rc = jmul_validate(M, N, P, D, A, X);
  if (rc) return rc;
  // grid: x = row blocks of 64 (<= 2^24), y = column strips of 16 (<= 65535)
  if (M > (int64_t(1) << 30) || P > int64_t(16) * 65535)
    return fail(JB_ERR_UNSUPPORTED, "fastmul is a small-matrix path: M <= 2^30 and P <= 1048560");
  std::lock_guard<std::mutex> lk(rt().mu);
  if (M == 0 || P == 0) return JB_OK;
  rc = ensure_init();
  if (rc) return rc;
  if (!host) return fastmul_dev_locked<T>(M, N, P, D, A, X, (hipStream_t)stream);
  GemmArgs g{};
  g.m = M;
  g.n = P;
  g.k = N;
  g.alpha = 1.0;
  g.A = A;
  g.lda = M;
  g.B = X;
  g.ldb = N;
  g.C = D;
  g.ldc = M;
  return run_staged(g, sizeof(T), false, [](const GemmArgs& d, hipStream_t s) {
    return fastmul_dev_locked<T>(d.m, d.k, d.n, (T*)d.C, (const T*)d.A, (const T*)d.B, s);
  });
}
}  // extern "C++"

int jb_fastmul(int64_t M, int64_t N, int64_t P, double* D, const double* A, const double* X) {
  return fastmul_entry<double>(M, N, P, D, A, X, nullptr, true);
}
int jb_fastmul_dev(int64_t M, int64_t N, int64_t P, double* D, const double* A, const double* X, void* stream) {
  return fastmul_entry<double>(M, N, P, D, A, X, stream, false);
}
int jb_fastmul_f32(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X) {
  return fastmul_entry<float>(M, N, P, D, A, X, nullptr, true);
}
int jb_fastmul_f32_dev(int64_t M, int64_t N, int64_t P, float* D, const float* A, const float* X, void* stream) {
  return fastmul_entry<float>(M, N, P, D, A, X, stream, false);
}

// ---- FP64 emulated (split-BF16 / split-INT8) -------------------------------
// Enqueues the emulated GEMM on `st` without synchronising.  A non-finite
// operand makes the C elements it feeds NaN and raises the sticky flag
// (jb_nonfinite_check); with `host_flag` the call's own flag is also copied to
// *host_flag in stream order (the host-pointer entry points, which synchronise).
static int ozaki_dev_locked(const GemmArgs& g, int s, bool i8, hipStream_t st, int* host_flag = nullptr) {
  if (g.m == 0 || g.n == 0) return JB_OK;
  if (g.alpha == 0.0 || g.k == 0) {
    hipError_t e = launch_scale_f64(g.m, g.n, g.beta, static_cast<double*>(g.C), g.ldc, st);
    return e == hipSuccess ? JB_OK : hip_fail(e, "scale launch");
  }
  void* ws = nullptr;
  int rc = ws_acquire(i8 ? ozaki_i8_workspace_bytes(g.m, g.n, g.k, s) : ozaki_workspace_bytes(g.m, g.n, g.k, s), st, &ws);
  if (rc) return rc;
  hipError_t e = i8 ? launch_dgemm_ozaki_i8(g, s, ws, rt().sticky, st) : launch_dgemm_ozaki(g, s, ws, rt().sticky, st);
  if (e == hipSuccess && host_flag)
    e = hipMemcpyAsync(host_flag, static_cast<char*>(ws) + emulated_flag_offset(g.m, g.n, g.k, s, i8), sizeof(int),
                       hipMemcpyDeviceToHost, st);
  int relrc = ws_release(st);
  if (e != hipSuccess) return hip_fail(e, "dgemm_ozaki launch");
  return relrc;
}

static int nonfinite_host(int flag, const char* what) {
  if (!flag) return JB_OK;
  return fail(JB_ERR_NONFINITE, "an operand of the %s holds Inf or NaN: the C elements it feeds are NaN", what);
}

int jb_dgemm_ozaki_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                       int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int nslices,
                       void* stream) {
  if (nslices < 1 || nslices > 8) return fail(-14, "nslices must be in 1..8 (got %d)", nslices);
  JB_GEMM_PROLOGUE(alpha, beta)
  return ozaki_dev_locked(g, nslices, false, (hipStream_t)stream);
}

// jb_dgemm_ozaki_dev without any allocation: the workspace is this device's
// `scratch` array (address from init), the kernels are the same.
int jb_dgemm_ozaki_noalloc_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha,
                               const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                               int64_t ldc, int nslices, void* stream) {
  if (nslices < 1 || nslices > 8) return fail(-14, "nslices must be in 1..8 (got %d)", nslices);
  GemmArgs g;
  int rc = validate_gemm(transA, transB, m, n, k, alpha == 0, A, lda, B, ldb, C, ldc, &g);
  if (rc) return rc;
  g.alpha = alpha;
  g.beta = beta;
  const size_t need = ozaki_workspace_bytes(m, n, k, nslices);
  if (need > JB_NOALLOC_WORKSPACE_BYTES)
    return fail(JB_ERR_UNSUPPORTED,
                "the allocation-free emulated DGEMM needs %zu bytes of workspace for %lld x %lld x %lld with %d "
                "slices; its scratch holds %llu",
                need, (long long)m, (long long)n, (long long)k, nslices,
                (unsigned long long)JB_NOALLOC_WORKSPACE_BYTES);
  std::lock_guard<std::mutex> lk(rt().mu);
  Runtime& r = rt();
  if (!r.inited) return fail(JB_ERR_STATE, "call jb_init first: the allocation-free entry point does not initialise");
  if (m == 0 || n == 0) return JB_OK;
  char* sp = r.device < JB_MAX_DEVICES ? static_cast<char*>(g_scratch_ptr[r.device]) : nullptr;
  if (!sp) return fail(JB_ERR_STATE, "no scratch address for device %d", r.device);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (alpha == 0.0 || k == 0) {
    e = launch_scale_f64(m, n, beta, C, ldc, st);
  } else {
    tl_no_timing = true;
    e = launch_dgemm_ozaki(g, nslices, sp + JB_SCRATCH_HEADER, reinterpret_cast<int*>(sp), st);
    tl_no_timing = false;
  }
  return e == hipSuccess ? JB_OK : hip_fail(e, "dgemm_ozaki (allocation-free) launch");
}

int jb_dgemm_ozaki(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                   int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int nslices) {
  if (nslices < 1 || nslices > 8) return fail(-14, "nslices must be in 1..8 (got %d)", nslices);
  JB_GEMM_PROLOGUE(alpha, beta)
  int bad = 0;
  rc = run_staged(g, sizeof(double), beta != 0.0, [nslices, &bad](const GemmArgs& d, hipStream_t s) {
    return ozaki_dev_locked(d, nslices, false, s, &bad);
  });
  return rc ? rc : nonfinite_host(bad, "emulated DGEMM");
}

int jb_dgemm_ozaki_i8_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                          int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int nslices,
                          void* stream) {
  if (nslices < 1 || nslices > 8) return fail(-14, "nslices must be in 1..8 (got %d)", nslices);
  JB_GEMM_PROLOGUE(alpha, beta)
  return ozaki_dev_locked(g, nslices, true, (hipStream_t)stream);
}

int jb_dgemm_ozaki_i8(char transA, char transB, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                      int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int nslices) {
  if (nslices < 1 || nslices > 8) return fail(-14, "nslices must be in 1..8 (got %d)", nslices);
  JB_GEMM_PROLOGUE(alpha, beta)
  int bad = 0;
  rc = run_staged(g, sizeof(double), beta != 0.0, [nslices, &bad](const GemmArgs& d, hipStream_t s) {
    return ozaki_dev_locked(d, nslices, true, s, &bad);
  });
  return rc ? rc : nonfinite_host(bad, "emulated DGEMM");
}

// ---- FP32 -------------------------------------------------------------------
static int sgemm_dev_locked(const GemmArgs& g, int method, int nprod, hipStream_t s, int* host_flag = nullptr) {
  if (g.m == 0 || g.n == 0) return JB_OK;
  hipError_t e;
  if (g.alpha == 0.0 || g.k == 0) {
    e = launch_scale_f32(g.m, g.n, (float)g.beta, static_cast<float*>(g.C), g.ldc, s);
    return e == hipSuccess ? JB_OK : hip_fail(e, "scale launch");
  }
  if (method == JB_F32_NATIVE) {
    e = launch_sgemm_mfma(g, s);
    return e == hipSuccess ? JB_OK : hip_fail(e, "sgemm_mfma launch");
  }
  void* ws = nullptr;
  int rc = ws_acquire(bf16x3_workspace_bytes(g.m, g.n, g.k), s, &ws);
  if (rc) return rc;
  e = launch_sgemm_bf16x3(g, nprod, ws, rt().sticky, s);
  if (e == hipSuccess && host_flag)
    e = hipMemcpyAsync(host_flag, static_cast<char*>(ws) + bf16x3_flag_offset(g.m, g.n, g.k), sizeof(int),
                       hipMemcpyDeviceToHost, s);
  rc = ws_release(s);
  if (e != hipSuccess) return hip_fail(e, "sgemm_bf16x3 launch");
  return rc;
}

// One problem on the single-problem kernels of a grouped method (the caller
// holds the lock and has validated `g`).
extern "C++" int jb::single_gemm_dev_locked(int method, const GemmArgs& g, hipStream_t s) {
  if (method == JB_F64_NATIVE) return dgemm_dev_locked(g, false, s);
  if (method == JB_F32_NATIVE) return sgemm_dev_locked(g, JB_F32_NATIVE, 0, s);
  if (method == JB_F32_BF16X3) return sgemm_dev_locked(g, JB_F32_BF16X3, 6, s);
  return fail(-1, "unknown method %d", method);
}

int jb_sgemm_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                 const float* B, int64_t ldb, float beta, float* C, int64_t ldc, void* stream) {
  JB_GEMM_PROLOGUE(alpha, beta)
  return sgemm_dev_locked(g, JB_F32_NATIVE, 0, (hipStream_t)stream);
}

int jb_sgemm(char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
             const float* B, int64_t ldb, float beta, float* C, int64_t ldc) {
  JB_GEMM_PROLOGUE(alpha, beta)
  return run_staged(g, sizeof(float), beta != 0.0f, [](const GemmArgs& d, hipStream_t s) {
    return sgemm_dev_locked(d, JB_F32_NATIVE, 0, s);
  });
}

int jb_sgemm_bf16x3_dev(char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha, const float* A,
                        int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int nprod,
                        void* stream) {
  if (nprod != 3 && nprod != 6 && nprod != 9) return fail(-14, "nprod must be 3, 6 or 9 (got %d)", nprod);
  JB_GEMM_PROLOGUE(alpha, beta)
  return sgemm_dev_locked(g, JB_F32_BF16X3, nprod, (hipStream_t)stream);
}

int jb_sgemm_bf16x3(char transA, char transB, int64_t m, int64_t n, int64_t k, float alpha, const float* A,
                    int64_t lda, const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int nprod) {
  if (nprod != 3 && nprod != 6 && nprod != 9) return fail(-14, "nprod must be 3, 6 or 9 (got %d)", nprod);
  JB_GEMM_PROLOGUE(alpha, beta)
  int bad = 0;
  rc = run_staged(g, sizeof(float), beta != 0.0f, [nprod, &bad](const GemmArgs& d, hipStream_t s) {
    return sgemm_dev_locked(d, JB_F32_BF16X3, nprod, s, &bad);
  });
  return rc ? rc : nonfinite_host(bad, "3xBF16 SGEMM");
}

int jb_nonfinite_check(void* stream) {
  std::lock_guard<std::mutex> lk(rt().mu);
  tl_error.clear();
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (capturing(s)) return fail(JB_ERR_STATE, "jb_nonfinite_check synchronises: not allowed during graph capture");
  // the library's sticky word and the allocation-free calls' one (scratch header)
  int* words[2] = {rt().sticky, nullptr};
  const int dev = rt().device;
  if (dev < JB_MAX_DEVICES && g_scratch_ptr[dev]) words[1] = static_cast<int*>(g_scratch_ptr[dev]);
  int seen = 0;
  for (int* w : words) {
    if (!w) continue;
    int h = 0;
    hipError_t e = hipMemcpyAsync(&h, w, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && h) e = hipMemsetAsync(w, 0, sizeof(int), s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "non-finite flag readback");
    seen |= h;
  }
  return seen ? fail(JB_ERR_NONFINITE, "an emulated call since the last check met Inf or NaN in an operand") : JB_OK;
}

// ---- generator --------------------------------------------------------------
int jb_fill_uniform_dev(uint64_t seed, int64_t n, double* out, void* stream) {
  tl_error.clear();
  if (n < 0) return fail(-2, "n < 0");
  if (n > 0 && !out) return fail(-3, "out is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  if (n == 0) return JB_OK;
  int rc = ensure_init();
  if (rc) return rc;
  hipError_t e = launch_fill_uniform(seed, n, out, (hipStream_t)stream);
  return e == hipSuccess ? JB_OK : hip_fail(e, "fill launch");
}

int jb_fill_uniform_f32_dev(uint64_t seed, int64_t n, float* out, void* stream) {
  tl_error.clear();
  if (n < 0) return fail(-2, "n < 0");
  if (n > 0 && !out) return fail(-3, "out is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  if (n == 0) return JB_OK;
  int rc = ensure_init();
  if (rc) return rc;
  hipError_t e = launch_fill_uniform_f32(seed, n, out, (hipStream_t)stream);
  return e == hipSuccess ? JB_OK : hip_fail(e, "fill launch");
}

extern "C++" {
template <typename T>
static int fill_wide_entry(uint64_t seed, int64_t n, int emin, int emax, T* out, void* stream) {
  tl_error.clear();
  if (n < 0) return fail(-2, "n < 0");
  if (emin > emax) return fail(-4, "emax < emin");
  if (emin < -1000 || emax > 1000) return fail(-3, "exponent range outside [-1000, 1000]");
  if (n > 0 && !out) return fail(-5, "out is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  if (n == 0) return JB_OK;
  int rc = ensure_init();
  if (rc) return rc;
  hipError_t e = sizeof(T) == 8 ? launch_fill_wide(seed, n, emin, emax, (double*)out, (hipStream_t)stream)
                                : launch_fill_wide_f32(seed, n, emin, emax, (float*)out, (hipStream_t)stream);
  return e == hipSuccess ? JB_OK : hip_fail(e, "fill launch");
}

template <typename T>
static int fill_normal_entry(uint64_t seed, int64_t n, double scale, T* out, void* stream) {
  tl_error.clear();
  if (n < 0) return fail(-2, "n < 0");
  if (!(scale == scale) || scale - scale != 0.0) return fail(-3, "scale is not finite");
  if (n > 0 && !out) return fail(-4, "out is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  if (n == 0) return JB_OK;
  int rc = ensure_init();
  if (rc) return rc;
  hipError_t e = sizeof(T) == 8 ? launch_fill_normal(seed, n, scale, (double*)out, (hipStream_t)stream)
                                : launch_fill_normal_f32(seed, n, scale, (float*)out, (hipStream_t)stream);
  return e == hipSuccess ? JB_OK : hip_fail(e, "fill launch");
}
}  // extern "C++"

int jb_fill_wide_f32_dev(uint64_t seed, int64_t n, int emin, int emax, float* out, void* stream) {
  return fill_wide_entry<float>(seed, n, emin, emax, out, stream);
}

int jb_fill_normal_dev(uint64_t seed, int64_t n, double scale, double* out, void* stream) {
  return fill_normal_entry<double>(seed, n, scale, out, stream);
}

int jb_fill_normal_f32_dev(uint64_t seed, int64_t n, double scale, float* out, void* stream) {
  return fill_normal_entry<float>(seed, n, scale, out, stream);
}

int jb_fill_wide_dev(uint64_t seed, int64_t n, int emin, int emax, double* out, void* stream) {
  tl_error.clear();
  if (n < 0) return fail(-2, "n < 0");
  if (emin > emax) return fail(-4, "emax < emin");
  if (emin < -1000 || emax > 1000) return fail(-3, "exponent range outside [-1000, 1000]");
  if (n > 0 && !out) return fail(-5, "out is NULL");
  std::lock_guard<std::mutex> lk(rt().mu);
  if (n == 0) return JB_OK;
  int rc = ensure_init();
  if (rc) return rc;
  hipError_t e = launch_fill_wide(seed, n, emin, emax, out, (hipStream_t)stream);
  return e == hipSuccess ? JB_OK : hip_fail(e, "fill launch");
}

}  // extern "C"
