// State handles, error reporting and device bookkeeping of libbsim.
// Mirrors the storage side of StateVectorSimulator / DensityMatrixSimulator
// (src/sv_simulator.jl:6-127, src/dm_simulator.jl:6-128).
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <string>

#include "common.h"

namespace bsim {

static thread_local std::string g_last_error;
static std::atomic<uint64_t> g_state_gen{ 0 };

void set_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error_text() { return g_last_error; }

int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    std::string m = std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ") in " + what;
    set_error(m);
    return e == hipErrorOutOfMemory ? BSIM_ENOMEM : BSIM_EHIP;
}

int ensure_ws(Workspace& w, size_t bytes) {
    if (w.bytes >= bytes && w.ptr) return BSIM_OK;
    if (w.ptr) {
        hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
    }
    if (bytes == 0) return BSIM_OK;
    hipError_t e = hipMalloc(&w.ptr, bytes);
    if (e != hipSuccess) {
        w.ptr = nullptr;
        return hip_fail(e, "hipMalloc(workspace)");
    }
    w.bytes = bytes;
    return BSIM_OK;
}

void free_ws(Workspace& w) {
    if (w.ptr) hipFree(w.ptr);
    w.ptr = nullptr;
    w.bytes = 0;
}

int check_state(const bsim_state* h) {
    if (!h || !h->d) return fail(BSIM_EINVAL, "null or destroyed state handle");
    return BSIM_OK;
}

int device_cu_count() {
    static int cus = 0;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        hipGetDevice(&dev);
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess) cus = p.multiProcessorCount;
        if (cus <= 0) cus = 256;
    });
    return cus;
}

__global__ void k_reset(double2* __restrict__ d, int64_t len, int64_t one_at) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride)
        d[i] = make_double2(i == one_at ? 1.0 : 0.0, 0.0);
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_version(void) { return 10100; }

int bsim_last_error(char* buf, int64_t len) {
    const std::string& e = g_last_error;
    if (buf && len > 0) {
        size_t n = e.size() < (size_t)(len - 1) ? e.size() : (size_t)(len - 1);
        memcpy(buf, e.data(), n);
        buf[n] = 0;
    }
    return (int)e.size();
}

int bsim_device_info(char* name, int64_t name_len, int* cu_count, int64_t* hbm_bytes) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(BSIM_ENODEV, "no HIP device visible");
    int dev = 0;
    BSIM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    BSIM_HIP(hipGetDeviceProperties(&p, dev));
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s", p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
    return BSIM_OK;
}

int bsim_state_create(int qubit_count, int kind, bsim_state** out) {
    if (!out) return fail(BSIM_EINVAL, "out is null");
    *out = nullptr;
    if (kind != BSIM_SV && kind != BSIM_DM) return fail(BSIM_EINVAL, "kind must be BSIM_SV or BSIM_DM");
    int vq = kind == BSIM_SV ? qubit_count : 2 * qubit_count;
    if (qubit_count < 1 || vq > 36) return fail(BSIM_EINVAL, "qubit count out of range (SV 1..36, DM 1..18)");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(BSIM_ENODEV, "no HIP device visible");
    bsim_state* h = new bsim_state();
    h->gen = ++g_state_gen;
    h->nq = qubit_count;
    h->kind = kind;
    h->vq = vq;
    h->len = (int64_t)1 << vq;
    hipGetDevice(&h->device);
    hipError_t e = hipMalloc((void**)&h->d, sizeof(double2) * (size_t)h->len);
    if (e != hipSuccess) {
        delete h;
        return hip_fail(e, "hipMalloc(state)");
    }
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        hipFree(h->d);
        delete h;
        return hip_fail(e, "hipStreamCreate");
    }
    e = hipMalloc((void**)&h->tile_ctr, 256);
    if (e == hipSuccess) e = hipMemsetAsync(h->tile_ctr, 0, 256, h->stream);
    if (e != hipSuccess) {
        bsim_state_destroy(h);
        return hip_fail(e, "hipMalloc(tile counter)");
    }
    int rc = bsim_state_reset(h);
    if (rc != BSIM_OK) {
        bsim_state_destroy(h);
        return rc;
    }
    *out = h;
    return BSIM_OK;
}

int bsim_state_destroy(bsim_state* h) {
    if (!h) return BSIM_OK;
    if (h->stream) hipStreamSynchronize(h->stream);
    detach_graphs(h);
    if (h->d) hipFree(h->d);
    if (h->tile_ctr) hipFree(h->tile_ctr);
    free_ws(h->red);
    free_ws(h->res);
    free_ws(h->scratch);
    free_ws(h->ap);
    free_ws(h->alias);
    free_ws(h->alias_work);
    free_ws(h->ops);
    free_ws(h->lam);
    if (h->stream) hipStreamDestroy(h->stream);
    h->d = nullptr;
    delete h;
    return BSIM_OK;
}

int bsim_state_reset(bsim_state* h) {
    if (int rc = check_state(h)) return rc;
    int64_t grid = grid_for(h->len, 256, 8 * 1024);
    hipLaunchKernelGGL(k_reset, dim3((unsigned)grid), dim3(256), 0, h->stream, h->d, h->len, (int64_t)0);
    BSIM_HIP(hipGetLastError());
    h->alias_ready = false;
    return BSIM_OK;
}

int bsim_state_qubits(const bsim_state* h, int* qubit_count, int* kind) {
    if (int rc = check_state(h)) return rc;
    if (qubit_count) *qubit_count = h->nq;
    if (kind) *kind = h->kind;
    return BSIM_OK;
}

int bsim_state_upload(bsim_state* h, const double* host, int64_t n_complex) {
    if (int rc = check_state(h)) return rc;
    if (!host || n_complex != h->len) return fail(BSIM_EINVAL, "upload: length must equal the state size");
    BSIM_HIP(hipMemcpyAsync(h->d, host, sizeof(double2) * (size_t)n_complex, hipMemcpyHostToDevice, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    h->alias_ready = false;
    return BSIM_OK;
}

int bsim_state_download(bsim_state* h, double* host, int64_t n_complex) {
    if (int rc = check_state(h)) return rc;
    if (!host || n_complex != h->len) return fail(BSIM_EINVAL, "download: length must equal the state size");
    BSIM_HIP(hipMemcpyAsync(host, h->d, sizeof(double2) * (size_t)n_complex, hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_state_copy(bsim_state* dst, const bsim_state* src) {
    if (int rc = check_state(dst)) return rc;
    if (int rc = check_state(src)) return rc;
    if (dst->len != src->len || dst->kind != src->kind) return fail(BSIM_EINVAL, "copy: shapes differ");
    BSIM_HIP(hipStreamSynchronize(src->stream));
    BSIM_HIP(hipMemcpyAsync(dst->d, src->d, sizeof(double2) * (size_t)src->len, hipMemcpyDeviceToDevice, dst->stream));
    dst->alias_ready = false;
    return BSIM_OK;
}

int bsim_synchronize(bsim_state* h) {
    if (int rc = check_state(h)) return rc;
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_state_device_ptr(bsim_state* h, void** dptr) {
    if (int rc = check_state(h)) return rc;
    if (!dptr) return fail(BSIM_EINVAL, "dptr is null");
    *dptr = h->d;
    return BSIM_OK;
}

}  // extern "C"
