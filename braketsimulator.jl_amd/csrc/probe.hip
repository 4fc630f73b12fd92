// HBM stream probes: the bandwidth ceiling the fused passes and the
// per-gate kernels are judged against.  A plain streaming kernel of known
// traffic, in three forms:
//   copy  -- read n x 16 B, write n x 16 B (the per-gate kernels' pattern:
//            every amplitude read and written once);
//   read  -- read n x 16 B, reduce to one partial per workgroup (the
//            expectation reductions' pattern);
//   write -- write n x 16 B (a fill).
// Each thread moves PU x 16 B per trip (global_load_dwordx4, non-temporal
// stores; PU = 4, 8 or 16, the fastest is reported), grid-stride over a
// buffer whose length is a multiple of the largest trip, so every trip is
// whole.  Measurement only: no reference counterpart.
#include "common.h"

namespace bsim {
namespace {

constexpr int PB = 256;                 // threads per workgroup
constexpr int64_t PTRIP_MAX = PB * 16;  // the largest trip (16-byte elements)

template <int PU>
__global__ __launch_bounds__(PB) void k_probe_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                                   int64_t n) {
    constexpr int64_t PTRIP = PB * PU;
    const int64_t stride = (int64_t)gridDim.x * PTRIP;
    for (int64_t b = (int64_t)blockIdx.x * PTRIP; b < n; b += stride) {
        double2 v[PU];
#pragma unroll
        for (int u = 0; u < PU; ++u) v[u] = src[b + u * PB + threadIdx.x];
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            double* p = (double*)(dst + b + u * PB + threadIdx.x);
            __builtin_nontemporal_store(v[u].x, p);
            __builtin_nontemporal_store(v[u].y, p + 1);
        }
    }
}

template <int PU>
__global__ __launch_bounds__(PB) void k_probe_read(const double2* __restrict__ src, double* __restrict__ partial,
                                                   int64_t n) {
    constexpr int64_t PTRIP = PB * PU;
    const int64_t stride = (int64_t)gridDim.x * PTRIP;
    double acc = 0.0;
    for (int64_t b = (int64_t)blockIdx.x * PTRIP; b < n; b += stride) {
        double2 v[PU];
#pragma unroll
        for (int u = 0; u < PU; ++u) v[u] = src[b + u * PB + threadIdx.x];
#pragma unroll
        for (int u = 0; u < PU; ++u) acc += v[u].x + v[u].y;
    }
    __shared__ double red[PB];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = PB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

template <int PU>
__global__ __launch_bounds__(PB) void k_probe_write(double2* __restrict__ dst, int64_t n, double value) {
    constexpr int64_t PTRIP = PB * PU;
    const int64_t stride = (int64_t)gridDim.x * PTRIP;
    for (int64_t b = (int64_t)blockIdx.x * PTRIP; b < n; b += stride) {
#pragma unroll
        for (int u = 0; u < PU; ++u) {
            double* p = (double*)(dst + b + u * PB + threadIdx.x);
            __builtin_nontemporal_store(value, p);
            __builtin_nontemporal_store(value, p + 1);
        }
    }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace
}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_stream_probe(int mode, int64_t bytes, int blocks, int reps, double* best_ms) {
    if (mode < 0 || mode > 2) return fail(BSIM_EINVAL, "mode must be 0 (copy), 1 (read) or 2 (write)");
    if (bytes <= 0 || bytes % (PTRIP_MAX * 16) != 0)
        return fail(BSIM_EINVAL, "bytes must be a positive multiple of " + std::to_string(PTRIP_MAX * 16));
    if (reps < 1 || !best_ms) return fail(BSIM_EINVAL, "reps must be >= 1 and best_ms non-null");
    const int64_t n = bytes / 16;
    const int64_t trips = n / PTRIP_MAX;  // whole trips at every unroll depth
    if (blocks <= 0) blocks = 8 * device_cu_count();
    const unsigned grid = (unsigned)std::min<int64_t>(blocks, trips);
    DevBuf a, b, part;
    BSIM_HIP(hipMalloc(&a.p, (size_t)bytes));
    if (mode == 0) BSIM_HIP(hipMalloc(&b.p, (size_t)bytes));
    BSIM_HIP(hipMalloc(&part.p, sizeof(double) * grid));
    hipStream_t s = nullptr;
    BSIM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{ s };
    hipEvent_t e0, e1;
    BSIM_HIP(hipEventCreate(&e0));
    BSIM_HIP(hipEventCreate(&e1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{ e0, e1 };
    // defined contents for the reads
    hipLaunchKernelGGL(k_probe_write<4>, dim3(grid), dim3(PB), 0, s, (double2*)a.p, n, 0.5);
    BSIM_HIP(hipGetLastError());
    double best = 1e300;
    for (int pu : { 4, 8, 16 })
        for (int r = 0; r <= reps; ++r) {  // r = 0: warm-up
            BSIM_HIP(hipEventRecord(e0, s));
            const double2* src = (const double2*)a.p;
            if (mode == 0) {
                double2* dst = (double2*)b.p;
                if (pu == 4) hipLaunchKernelGGL(k_probe_copy<4>, dim3(grid), dim3(PB), 0, s, src, dst, n);
                else if (pu == 8) hipLaunchKernelGGL(k_probe_copy<8>, dim3(grid), dim3(PB), 0, s, src, dst, n);
                else hipLaunchKernelGGL(k_probe_copy<16>, dim3(grid), dim3(PB), 0, s, src, dst, n);
            } else if (mode == 1) {
                double* pp = (double*)part.p;
                if (pu == 4) hipLaunchKernelGGL(k_probe_read<4>, dim3(grid), dim3(PB), 0, s, src, pp, n);
                else if (pu == 8) hipLaunchKernelGGL(k_probe_read<8>, dim3(grid), dim3(PB), 0, s, src, pp, n);
                else hipLaunchKernelGGL(k_probe_read<16>, dim3(grid), dim3(PB), 0, s, src, pp, n);
            } else {
                double2* dst = (double2*)a.p;
                if (pu == 4) hipLaunchKernelGGL(k_probe_write<4>, dim3(grid), dim3(PB), 0, s, dst, n, 0.25);
                else if (pu == 8) hipLaunchKernelGGL(k_probe_write<8>, dim3(grid), dim3(PB), 0, s, dst, n, 0.25);
                else hipLaunchKernelGGL(k_probe_write<16>, dim3(grid), dim3(PB), 0, s, dst, n, 0.25);
            }
            BSIM_HIP(hipGetLastError());
            BSIM_HIP(hipEventRecord(e1, s));
            BSIM_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            BSIM_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (r > 0 && ms < best) best = ms;
        }
    *best_ms = best;
    return BSIM_OK;
}

}  // extern "C"
