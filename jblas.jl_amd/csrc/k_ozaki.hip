// k_ozaki.hip -- split-BF16 emulated DGEMM (placeholder until the kernels land).
#include "jb_internal.h"
namespace jb {
size_t ozaki_workspace_bytes(int64_t, int64_t, int64_t, int) { return 256; }
hipError_t launch_dgemm_ozaki(const GemmArgs&, int, void*, int*, hipStream_t) { return hipErrorNotSupported; }
size_t bf16x3_workspace_bytes(int64_t, int64_t, int64_t) { return 256; }
hipError_t launch_sgemm_bf16x3(const GemmArgs&, int, void*, hipStream_t) { return hipErrorNotSupported; }
}  // namespace jb
