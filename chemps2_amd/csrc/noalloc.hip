// The Heff product for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): chm_heff_apply_noalloc runs the kernels of
// chm_heff_apply on a prepared site with their partial-sum scratch taken from a
// file-scope device array instead of the context's buffers. It calls no hipMalloc or
// hipFree and no allocating library routine.
//
// The array is one per GPU, for the life of the process; every call on a GPU shares it
// and it keeps the last call's contents (the kernels overwrite every partial sum they
// read, so nothing needs clearing). Calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/chemps2_amd.h"
#include "context.hpp"

namespace {

// 32 Mi doubles (256 MiB) of partial sums
constexpr int64_t kScratchDoubles = int64_t(1) << 25;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per device (each
// GPU holds its own copy of the array at its own address).
double* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw chm::Error{CHM_ERR_ARG, "device index out of range"};
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw chm::Error{CHM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
    g_ptr[device] = (double*)p;
  }
  return g_ptr[device];
}

}  // namespace

extern "C" {

int chm_heff_apply_noalloc(chm_ctx* ctx, const double* x, double* y) {
  if (!ctx) return CHM_ERR_ARG;
  chm::Ctx& c = ctx->c;
  try {
    if (!x || !y) throw chm::Error{CHM_ERR_ARG, "null pointer"};
    if (c.site < 0) throw chm::Error{CHM_ERR_STATE, "no prepared site"};
    chm::hip_check(hipSetDevice(c.device), "hipSetDevice");
    const int64_t nb = c.bpart_need, nc = c.cpart_need;
    const int64_t nlp = c.n_lower > 0 ? (int64_t)chm::dot_parts(c.V) * c.n_lower : 0, nld = c.n_lower;
    // each part starts on a 256-byte boundary
    auto up = [](int64_t v) { return (v + 31) / 32 * 32; };
    const int64_t need = up(nb) + up(nc) + up(nlp) + up(nld);
    if (need > kScratchDoubles)
      throw chm::Error{CHM_ERR_SIZE, "the prepared site needs " + std::to_string(need) + " doubles of partial sums; " +
                                         "chm_heff_apply_noalloc holds at most " + std::to_string(kScratchDoubles)};
    double* p = scratch_on(c.device);
    double* bp = p;
    double* cp = bp + up(nb);
    double* lp = cp + up(nc);
    double* ld = lp + up(nlp);
    c.mv_low = false;
    c.apply_with(x, y, bp, cp, lp, ld);
    chm::hip_check(hipStreamSynchronize(c.stream), "apply");
    c.err.clear();
    return CHM_OK;
  } catch (const chm::Error& e) {
    c.err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    c.err = e.what();
    return CHM_ERR_STATE;
  }
}

}  // extern "C"
