// Shared definitions of libbsim: the state handle, error reporting and the
// complex128 device arithmetic every kernel uses.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "bsim.h"

namespace bsim {

// ---- error handling --------------------------------------------------

void set_error(const std::string& msg);
const std::string& last_error_text();  // this thread's last message
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define BSIM_HIP(call)                                  \
    do {                                                \
        hipError_t _e = (call);                         \
        if (_e != hipSuccess) return ::bsim::hip_fail(_e, #call); \
    } while (0)

// ---- state handle ----------------------------------------------------

struct Workspace {
    void* ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace bsim

struct bsim_state {
    int nq = 0;          // reference qubit count
    int kind = BSIM_SV;  // BSIM_SV / BSIM_DM
    int vq = 0;          // vector qubits: nq (SV) or 2 nq (DM)
    int64_t len = 0;     // 2^vq complex elements
    double2* d = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    // reduction partials, gate scratch copy, alias table, planner buffers
    bsim::Workspace red, res, scratch, ap, alias, alias_work, ops, lam;
    bool alias_ready = false;
    // ticket counter of the streaming tile pass: zero between passes (each
    // pass's last ticket resets it), so it is set once, at creation
    unsigned long long* tile_ctr = nullptr;
    uint64_t gen = 0;                   // unique per created state (process-wide counter)
    std::vector<bsim_graph*> graphs;    // captured evolutions on this state
};

namespace bsim {

int ensure_ws(Workspace& w, size_t bytes);
void free_ws(Workspace& w);
int check_state(const bsim_state* h);
void detach_graphs(bsim_state* h);  // graph.hip
int device_cu_count();

// ---- device helpers ----------------------------------------------------

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {  // a*b + c
    return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }

// |a|^2 with a fixed rounding (re*re rounded into the fma with im*im): every
// kernel that turns amplitudes into probabilities uses this one expression,
// so probability vectors and the alias tables built from them agree bit for
// bit.
__device__ __forceinline__ double sv_prob(double2 a) { return __fma_rn(a.x, a.x, __dmul_rn(a.y, a.y)); }

// Probability sources of the alias build: a probability vector, or the
// amplitudes of a state vector (no probability vector is written).
struct ProbVec {
    const double* p;
    __device__ __forceinline__ double operator()(int64_t i) const { return p[i]; }
};
struct ProbAmp {
    const double2* sv;
    __device__ __forceinline__ double operator()(int64_t i) const { return sv_prob(sv[i]); }
};

// Insert a zero bit at `bit` (src/gate_kernels.jl:34-38).
__host__ __device__ __forceinline__ uint64_t pad_bit(uint64_t ix, int bit) {
    uint64_t lo = ix & ((1ull << bit) - 1);
    return ((ix - lo) << 1) | lo;
}

// Philox-4x32-10 (Salmon et al., SC'11): ten rounds on ctr with key (k0, k1).
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t ctr[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)M0 * ctr[0];
        uint64_t p1 = (uint64_t)M1 * ctr[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ ctr[1] ^ k0;
        uint32_t n2 = hi0 ^ ctr[3] ^ k1;
        ctr[0] = n0;
        ctr[1] = lo1;
        ctr[2] = n2;
        ctr[3] = lo0;
        k0 += W0;
        k1 += W1;
    }
}

// Host-side helpers ---------------------------------------------------------
inline int64_t grid_for(int64_t work, int block, int64_t cap) {
    int64_t g = (work + block - 1) / block;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return g;
}

}  // namespace bsim
