// Fused evolution without device allocation, for callers that cannot
// allocate (tight serving loops, harnesses that forbid hipMalloc).
//
// bsim_evolve (fused mode) keeps the plan's device tables -- passes,
// sweeps and tile ops -- in a workspace of the state handle that grows with
// hipMalloc.  bsim_evolve_noalloc puts them in `scratch`, a file-scope
// device array of fixed size that every GPU holds for the life of the
// process, and launches the same tile kernels on them; it takes only plans
// made entirely of tile passes (the plans that need no other staging
// memory).  Each device's copy of the array has its own address, taken once
// per device with hipGetSymbolAddress while that device is current; host
// code uses only that address, never the array's name.
//
// All calls on one GPU share that GPU's array: calls that overlap in time
// on one GPU (two threads, or two handles on the same device) race on it.
// The call waits for its own kernels before it returns.
#include <mutex>

#include "fused.h"

namespace bsim {
namespace {

constexpr size_t kScratchBytes = (size_t)4 << 20;  // 4 MiB: about 10,000 tile ops
constexpr int kMaxDevices = 64;

__device__ __align__(256) char scratch[kScratchBytes];

std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// The address of `scratch` on `device`, taken once with the device current.
int scratch_address(int device, char** out) {
    if (device < 0 || device >= kMaxDevices) return fail(BSIM_ENODEV, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = 0;
        BSIM_HIP(hipGetDevice(&cur));
        if (cur != device) BSIM_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(BSIM_EHIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[device] = (char*)p;
    }
    *out = g_scratch_ptr[device];
    return BSIM_OK;
}

}  // namespace
}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_evolve_noalloc(bsim_state* h, const bsim_op* ops, int n_ops) {
    if (int rc = check_state(h)) return rc;
    if (n_ops < 0 || (n_ops > 0 && !ops)) return fail(BSIM_EINVAL, "bad op list");
    char* base = nullptr;
    if (int rc = scratch_address(h->device, &base)) return rc;
    std::vector<EOp> eops;
    if (int rc = lower_ops(h->nq, h->kind, ops, n_ops, eops, false)) return rc;
    Plan plan;
    if (int rc = make_plan(h->vq, eops, tile_config(), plan)) return rc;
    for (const Plan::Step& st : plan.steps)
        if (!st.tile)
            return fail(BSIM_ENOTSUP,
                        "evolve without allocation takes fusable ops only (dense <= 2 targets, diagonal <= 4)");
    const size_t bytes = plan_table_bytes(plan);
    if (bytes > kScratchBytes)
        return fail(BSIM_ENOTSUP, "evolve without allocation: the plan's tables (" + std::to_string(bytes) +
                                      " bytes) exceed the 4 MiB scratch; split the op list");
    if (int rc = prepare_tile_kernel()) return rc;
    h->alias_ready = false;
    if (int rc = upload_plan(plan, base, h->stream)) return rc;
    if (int rc = launch_plan(h, h->d, eops, plan, base, false)) return rc;
    // the tables were copied from host vectors that end with this call, and
    // the next call on this GPU rewrites the scratch
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

}  // extern "C"
