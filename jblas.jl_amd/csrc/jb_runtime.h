// jb_runtime.h -- library state shared by the runtime translation units.
#pragma once
#include <functional>
#include <mutex>
#include <vector>

#include "../../include/jblas_hip.h"
#include "jb_internal.h"

namespace jb {

// Grow-only device buffer.
struct Buffer {
  void* ptr = nullptr;
  size_t size = 0;
  int reserve(size_t bytes, const char* what);
  void release();
};

struct Runtime {
  std::mutex mu;  // serialises every entry point
  bool inited = false;
  int device = -1;
  jb_device_info info{};
  hipStream_t stream = nullptr;  // used by the host-pointer entry points
  Buffer ws;                     // emulation slices / split-K partials
  Buffer staging;                // host-pointer entry points: A, B, C copies
  hipEvent_t ws_event = nullptr; // last use of `ws`
  bool ws_pending = false;
  bool ws_in_graph = false;      // a graph captured a call that uses `ws`
  std::vector<void*> retired;    // workspaces that captured graphs may use (freed by jb_finalize)
  int* sticky = nullptr;         // set by any emulated call that met Inf/NaN; cleared by jb_nonfinite_check
};

Runtime& rt();
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
// JB_LOG diagnostics (no-ops unless the variable is set)
bool log_enabled();
void log_line(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
int hip_fail(hipError_t e, const char* what);
int validate_gemm(char ta, char tb, int64_t m, int64_t n, int64_t k, bool alpha_zero, const void* A, int64_t lda,
                  const void* B, int64_t ldb, const void* C, int64_t ldc, GemmArgs* out);
int ensure_init();
int init_locked(int device);
void finalize_locked();
int ws_acquire(size_t bytes, hipStream_t s, void** out);
int ws_release(hipStream_t s);
int run_staged(const GemmArgs& g, size_t elt, bool read_c, const std::function<int(const GemmArgs&, hipStream_t)>& dev);

// One GEMM of a grouped method (JB_F64_NATIVE, JB_F32_NATIVE, JB_F32_BF16X3
// with six products) on the single-problem kernels.
int single_gemm_dev_locked(int method, const GemmArgs& g, hipStream_t s);

// grouped / autotuner (jb_grouped.hip)
void grouped_release_all();
int tune_load_locked(const char* path);

}  // namespace jb
