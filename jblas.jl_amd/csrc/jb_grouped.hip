// jb_grouped.hip -- grouped GEMM, hipGraph plans, autotuner, pullback.
// (Entry points present; implementation lands with the grouped kernel.)
#include "jb_runtime.h"
namespace jb {
void grouped_release_all() {}
int tune_load_locked(const char*) { return JB_OK; }
}  // namespace jb

using namespace jb;
extern "C" {
int jb_gemm_grouped_dev(int, int64_t, const jb_gemm_desc*, void*) { return fail(JB_ERR_UNSUPPORTED, "grouped: not built"); }
int jb_grouped_plan_create(int, int64_t, const jb_gemm_desc*, void*, jb_grouped_plan**) { return fail(JB_ERR_UNSUPPORTED, "grouped: not built"); }
int jb_grouped_plan_launch(jb_grouped_plan*, void*) { return fail(JB_ERR_UNSUPPORTED, "grouped: not built"); }
int jb_grouped_plan_destroy(jb_grouped_plan*) { return JB_OK; }
int jb_autotune_grouped(int, int64_t, const jb_gemm_desc*, int, void*) { return fail(JB_ERR_UNSUPPORTED, "autotune: not built"); }
int jb_autotune_save(const char*) { return fail(JB_ERR_UNSUPPORTED, "autotune: not built"); }
int jb_autotune_load(const char*) { return fail(JB_ERR_UNSUPPORTED, "autotune: not built"); }
int jb_autotune_lookup(int, int64_t, int64_t, int64_t, int*, int*) { return fail(JB_ERR_UNSUPPORTED, "autotune: not built"); }
int jb_autotune_reset(void) { return JB_OK; }
int jb_gemm_pullback_dev(int, char, char, int64_t, int64_t, int64_t, double, const void*, int64_t, const void*, int64_t,
                         const void*, int64_t, void*, int64_t, void*, int64_t, void*) {
  return fail(JB_ERR_UNSUPPORTED, "pullback: not built");
}
int jb_gemm_pullback(int, char, char, int64_t, int64_t, int64_t, double, const void*, int64_t, const void*, int64_t,
                     const void*, int64_t, void*, int64_t, void*, int64_t) {
  return fail(JB_ERR_UNSUPPORTED, "pullback: not built");
}
}
