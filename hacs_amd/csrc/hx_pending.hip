// hx_pending.hip -- entry points declared in the header whose kernels are not built yet:
// each fails loudly with HX_E_RUNTIME.
#include "hx_common.h"
using namespace hx;
#define PENDING(name) return fail(HX_E_RUNTIME, #name ": not implemented in this build")
extern "C" {
HX_API int hx_gemm_batch_create(const hx_modctx*, const hx_gemm_job*, uint32_t, int, hx_gemm_batch*) { PENDING(hx_gemm_batch_create); }
HX_API int hx_gemm_batch_run(hx_gemm_batch, hx_stream) { PENDING(hx_gemm_batch_run); }
HX_API int hx_gemm_batch_destroy(hx_gemm_batch) { PENDING(hx_gemm_batch_destroy); }
}
