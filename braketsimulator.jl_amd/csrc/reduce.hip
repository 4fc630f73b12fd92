// Reductions: probabilities, Pauli / matrix expectations, marginals
// (src/sv_simulator.jl:240-271, src/dm_simulator.jl:181-262,
// src/result_types.jl:96-167).
//
// Every reduction is deterministic: a fixed grid writes one partial per
// block (wavefront shuffle tree -> LDS across wavefronts), then one small
// kernel sums the partials in block order.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.h"
#include "gates.h"
#include "reduce.h"

namespace bsim {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Sum NV values per thread over the block; thread 0 holds the result in out[].
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* NV * 16 */) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = wave_sum(v[i]);
        if (lane == 0) lds[i * 16 + wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double s = 0;
            for (int w = 0; w < nw; ++w) s += lds[i * 16 + w];
            v[i] = s;
        }
    }
    __syncthreads();
}

// partials[b * nv + i] -> out[i]; launched with one 256-thread block per
// value i: fixed strided partial sums, then a fixed tree (deterministic).
__global__ __launch_bounds__(256) void k_finish(const double* __restrict__ partials, int nblocks, int nv,
                                                double* __restrict__ out) {
    __shared__ double sh[256];
    const int i = blockIdx.x;
    double s = 0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partials[(size_t)b * nv + i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = sh[0];
}

constexpr int RBLOCK = 256;
static int red_grid(int64_t work) {
    // power of two, so that the bits below log2(grid*RBLOCK) are thread constants
    int64_t want = (work + RBLOCK - 1) / RBLOCK;
    int64_t cap = 4096;
    int64_t g = 1;
    while (g < want && g < cap) g <<= 1;
    return (int)g;
}

// ---- probabilities ----------------------------------------------------

__global__ void k_prob_sv(const double2* __restrict__ sv, int64_t len, double* __restrict__ p) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        p[i] = sv_prob(sv[i]);
    }
}

__global__ void k_prob_dm(const double2* __restrict__ dm, int64_t n, double* __restrict__ p) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = dm[i + i * n].x;
}

int probabilities_device(bsim_state* h, double* dp) {
    int64_t n = (int64_t)1 << h->nq;
    int64_t grid = grid_for(n, 256, (int64_t)device_cu_count() * 64);
    if (h->kind == BSIM_SV)
        hipLaunchKernelGGL(k_prob_sv, dim3((unsigned)grid), dim3(256), 0, h->stream, h->d, n, dp);
    else
        hipLaunchKernelGGL(k_prob_dm, dim3((unsigned)grid), dim3(256), 0, h->stream, h->d, n, dp);
    BSIM_HIP(hipGetLastError());
    return BSIM_OK;
}

// ---- all <Z_q> in one pass (SV) ----------------------------------------
// s[b] = sum of p_i over i with bit b set, tot = sum p_i.  For a power-of-two
// grid stride S, bits below log2(S) are constant per thread: only the upper
// bits are accumulated per amplitude.
__global__ __launch_bounds__(RBLOCK) void k_zsums_sv(const double2* __restrict__ sv, int vq, int lstride,
                                                     double* __restrict__ partials) {
    __shared__ double lds[37 * 16];
    double v[37];
#pragma unroll
    for (int i = 0; i < 37; ++i) v[i] = 0;
    const int64_t len = (int64_t)1 << vq;
    const int64_t stride = (int64_t)1 << lstride;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // Layout of v: [0..16) sums of bits lstride+b (the bits of the iteration
    // count k), [16..36) sums of bits 0..19 (thread constants), [36] total.
    const int hi_bits = vq > lstride ? vq - lstride : 0;
    if (hi_bits >= 3) {
        // eight iterations per trip, their loads issued together: bits 0-2
        // of the iteration count vary inside a trip (fixed pattern), the
        // higher bits are the trip's, so those sums take the trip total
        const int64_t iters = (int64_t)1 << hi_bits;
        const double2* __restrict__ base = sv + tid;
        for (int64_t k = 0; k < iters; k += 8) {
            double2 a[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) a[u] = base[(k + u) << lstride];
            double p[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) p[u] = a[u].x * a[u].x + a[u].y * a[u].y;
            const double s01 = p[0] + p[1], s23 = p[2] + p[3], s45 = p[4] + p[5], s67 = p[6] + p[7];
            const double tot = (s01 + s23) + (s45 + s67);
            v[36] += tot;
            v[0] += (p[1] + p[3]) + (p[5] + p[7]);
            v[1] += (p[2] + p[3]) + (p[6] + p[7]);
            v[2] += s45 + s67;
#pragma unroll
            for (int b = 3; b < 16; ++b)
                if (b < hi_bits && ((k >> b) & 1)) v[b] += tot;
        }
    } else {
        int64_t k = 0;
        for (int64_t i = tid; i < len; i += stride, ++k) {
            double2 a = sv[i];
            double p = a.x * a.x + a.y * a.y;
            v[36] += p;
#pragma unroll
            for (int b = 0; b < 16; ++b)
                if (b < hi_bits && ((k >> b) & 1)) v[b] += p;
        }
    }
#pragma unroll
    for (int b = 0; b < 20; ++b)
        if (b < lstride && ((tid >> b) & 1)) v[16 + b] = v[36];
    block_sum<37>(v, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < 37; ++i) partials[(size_t)blockIdx.x * 37 + i] = v[i];
}

// DM: diagonal only (n entries, stride n+1)
__global__ __launch_bounds__(RBLOCK) void k_zsums_dm(const double2* __restrict__ dm, int nq,
                                                     double* __restrict__ partials) {
    __shared__ double lds[37 * 16];
    double v[37];
#pragma unroll
    for (int i = 0; i < 37; ++i) v[i] = 0;
    const int64_t n = (int64_t)1 << nq;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        double p = dm[i + i * n].x;
        v[36] += p;
#pragma unroll
        for (int b = 0; b < 36; ++b)
            if (b < nq && ((i >> b) & 1)) v[b] += p;
    }
    block_sum<37>(v, lds);
    if (threadIdx.x == 0)
        for (int i = 0; i < 37; ++i) partials[(size_t)blockIdx.x * 37 + i] = v[i];
}

int zsums_device(bsim_state* h, double* dout /* 37 doubles */, int* layout) {
    int vq = h->kind == BSIM_SV ? h->vq : h->nq;
    int64_t len = (int64_t)1 << vq;
    int grid = red_grid(len);
    int lstride = 0;
    while ((1ll << lstride) < (int64_t)grid * RBLOCK) ++lstride;
    size_t need = sizeof(double) * ((size_t)grid * 37 + 64);
    if (int rc = ensure_ws(h->red, need)) return rc;
    double* partials = (double*)h->red.ptr;
    if (h->kind == BSIM_SV)
        hipLaunchKernelGGL(k_zsums_sv, dim3(grid), dim3(RBLOCK), 0, h->stream, h->d, vq, lstride, partials);
    else
        hipLaunchKernelGGL(k_zsums_dm, dim3(grid), dim3(RBLOCK), 0, h->stream, h->d, vq, partials);
    BSIM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finish, dim3(37), dim3(256), 0, h->stream, partials, grid, 37, dout);
    BSIM_HIP(hipGetLastError());
    if (layout) *layout = h->kind == BSIM_SV ? lstride : -1;
    return BSIM_OK;
}

// ---- Pauli words ------------------------------------------------------
// (P psi)_i = phase(i) psi_{i ^ f}, phase(i) = i^ny (-1)^{popc(~i & y)} (-1)^{popc(i & z)}
// with f = x | y.  <P> = sum_i conj(psi_i) (P psi)_i.  Real part returned.
__device__ __forceinline__ double2 pauli_phase(uint64_t i, uint64_t ymask, uint64_t zmask, int ny) {
    int neg = (__popcll(~i & ymask) + __popcll(i & zmask)) & 1;
    double s = neg ? -1.0 : 1.0;
    switch (ny & 3) {
        case 0: return make_double2(s, 0);
        case 1: return make_double2(0, s);
        case 2: return make_double2(-s, 0);
        default: return make_double2(0, -s);
    }
}

__global__ __launch_bounds__(RBLOCK) void k_pauli_sv(const double2* __restrict__ sv, int64_t len, uint64_t fmask,
                                                     uint64_t ymask, uint64_t zmask, int ny,
                                                     double* __restrict__ partials) {
    __shared__ double lds[2 * 16];
    double v[2] = { 0, 0 };
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        double2 a = sv[i];
        double2 b = sv[(uint64_t)i ^ fmask];
        double2 t = cmul(pauli_phase((uint64_t)i, ymask, zmask, ny), b);
        v[0] += a.x * t.x + a.y * t.y;  // Re(conj(a) t)
        v[1] += a.x * t.y - a.y * t.x;
    }
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 2] = v[0];
        partials[blockIdx.x * 2 + 1] = v[1];
    }
}

// DM: <P> = sum_r phase(r) S[r, r ^ f]   (S column-major, stored rho^T)
__global__ __launch_bounds__(RBLOCK) void k_pauli_dm(const double2* __restrict__ dm, int64_t n, uint64_t fmask,
                                                     uint64_t ymask, uint64_t zmask, int ny,
                                                     double* __restrict__ partials) {
    __shared__ double lds[2 * 16];
    double v[2] = { 0, 0 };
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += stride) {
        double2 s = dm[r + (int64_t)((uint64_t)r ^ fmask) * n];
        double2 t = cmul(pauli_phase((uint64_t)r, ymask, zmask, ny), s);
        v[0] += t.x;
        v[1] += t.y;
    }
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 2] = v[0];
        partials[blockIdx.x * 2 + 1] = v[1];
    }
}

int parse_pauli(int nq, const char* word, PauliMask& pm) {
    pm = PauliMask();
    for (int q = 0; q < nq; ++q) {
        uint64_t bit = 1ull << (nq - 1 - q);
        switch (word[q]) {
            case 'I': case 'i': break;
            case 'X': case 'x': pm.x |= bit; break;
            case 'Y': case 'y': pm.y |= bit; pm.ny++; break;
            case 'Z': case 'z': pm.z |= bit; break;
            default: return fail(BSIM_EINVAL, "Pauli words use the letters I, X, Y, Z");
        }
    }
    return BSIM_OK;
}

int pauli_expect_device(bsim_state* h, const PauliMask& pm, double* dout2) {
    int64_t n = (int64_t)1 << h->nq;
    int grid = red_grid(n);
    if (int rc = ensure_ws(h->red, sizeof(double) * ((size_t)grid * 2 + 64))) return rc;
    double* partials = (double*)h->red.ptr;
    uint64_t f = pm.x | pm.y;
    if (h->kind == BSIM_SV)
        hipLaunchKernelGGL(k_pauli_sv, dim3(grid), dim3(RBLOCK), 0, h->stream, h->d, n, f, pm.y, pm.z, pm.ny,
                           partials);
    else
        hipLaunchKernelGGL(k_pauli_dm, dim3(grid), dim3(RBLOCK), 0, h->stream, h->d, n, f, pm.y, pm.z, pm.ny,
                           partials);
    BSIM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finish, dim3(2), dim3(256), 0, h->stream, partials, grid, 2, dout2);
    BSIM_HIP(hipGetLastError());
    return BSIM_OK;
}

// ---- dense Hermitian observable, SV, K = 1, 2 without a copy ------------
template <int K>
__global__ __launch_bounds__(RBLOCK) void k_expect_dense(const double2* __restrict__ sv, int64_t ngroups, int4 ins,
                                                         int t0, int t1, const double2* __restrict__ mg,
                                                         double* __restrict__ partials) {
    constexpr int D = 1 << K;
    __shared__ double lds[2 * 16];
    double2 M[D * D];
#pragma unroll
    for (int i = 0; i < D * D; ++i) M[i] = mg[i];
    double v[2] = { 0, 0 };
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
        uint64_t base = (uint64_t)g;
        base = pad_bit(base, ins.x);
        if (K > 1) base = pad_bit(base, ins.y);
        double2 a[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            uint64_t off = base | ((j & 1) ? (1ull << t0) : 0) | ((j & 2) ? (1ull << t1) : 0);
            a[j] = sv[off];
        }
#pragma unroll
        for (int r = 0; r < D; ++r) {
            double2 acc = cmul(M[r], a[0]);
#pragma unroll
            for (int c = 1; c < D; ++c) acc = cfma(M[r + D * c], a[c], acc);
            v[0] += a[r].x * acc.x + a[r].y * acc.y;
            v[1] += a[r].x * acc.y - a[r].y * acc.x;
        }
    }
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 2] = v[0];
        partials[blockIdx.x * 2 + 1] = v[1];
    }
}

// conj(a) . b over len amplitudes
__global__ __launch_bounds__(RBLOCK) void k_dot(const double2* __restrict__ a, const double2* __restrict__ b,
                                                int64_t len, double* __restrict__ partials) {
    __shared__ double lds[2 * 16];
    double v[2] = { 0, 0 };
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        double2 x = a[i], y = b[i];
        v[0] += x.x * y.x + x.y * y.y;
        v[1] += x.x * y.y - x.y * y.x;
    }
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 2] = v[0];
        partials[blockIdx.x * 2 + 1] = v[1];
    }
}

int dot_device(const double2* a, const double2* b, int64_t len, Workspace& red, hipStream_t s, double* dout2) {
    int grid = red_grid(len);
    if (int rc = ensure_ws(red, sizeof(double) * ((size_t)grid * 2 + 64))) return rc;
    double* partials = (double*)red.ptr;
    hipLaunchKernelGGL(k_dot, dim3(grid), dim3(RBLOCK), 0, s, a, b, len, partials);
    BSIM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finish, dim3(2), dim3(256), 0, s, partials, grid, 2, dout2);
    BSIM_HIP(hipGetLastError());
    return BSIM_OK;
}

// DM Hermitian: <O> = sum_rest sum_{j,j'} O[j', j] S[rest|j', rest|j]
// (the trace of the reference's O^T-on-rows product, src/dm_simulator.jl:190-224)
__global__ __launch_bounds__(RBLOCK) void k_expect_dm_matrix(const double2* __restrict__ dm, int nq, int K,
                                                             const int* __restrict__ ord, const double2* __restrict__ mg,
                                                             double* __restrict__ partials) {
    __shared__ double lds[2 * 16];
    const int D = 1 << K;
    const int64_t n = (int64_t)1 << nq;
    const int64_t nrest = n >> K;
    double v[2] = { 0, 0 };
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nrest * D * D; t += stride) {
        int64_t rest = t / (D * D);
        int jj = (int)(t % (D * D));
        int j = jj / D, jp = jj % D;
        uint64_t base = (uint64_t)rest;
        for (int b = 0; b < K; ++b) base = pad_bit(base, ord[b]);
        uint64_t r = base, c = base;
        for (int b = 0; b < K; ++b) {
            if ((jp >> b) & 1) r |= 1ull << ord[b];
            if ((j >> b) & 1) c |= 1ull << ord[b];
        }
        double2 s = dm[r + c * n];
        double2 o = mg[jp + D * j];
        double2 p = cmul(o, s);
        v[0] += p.x;
        v[1] += p.y;
    }
    block_sum<2>(v, lds);
    if (threadIdx.x == 0) {
        partials[blockIdx.x * 2] = v[0];
        partials[blockIdx.x * 2 + 1] = v[1];
    }
}

// ---- marginals ----------------------------------------------------------
// out[j] = sum over rest of p[scatter(j, tbits) | scatter(rest, other bits)]
__global__ __launch_bounds__(RBLOCK) void k_marginal(const double* __restrict__ p, int nq, int nt,
                                                     const int* __restrict__ tsorted, int64_t nchunks, int chunk_bits,
                                                     double* __restrict__ partials) {
    int64_t nout = (int64_t)1 << nt;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nout * nchunks) return;
    int64_t j = t / nchunks, chunk = t % nchunks;
    double s = 0;
    int64_t per = (int64_t)1 << chunk_bits;
    for (int64_t r = 0; r < per; ++r) {
        uint64_t rest = (uint64_t)(chunk * per + r);
        uint64_t idx = rest;
        for (int b = 0; b < nt; ++b) idx = pad_bit(idx, tsorted[b]);
        for (int b = 0; b < nt; ++b)
            if ((j >> b) & 1) idx |= 1ull << tsorted[b];
        s += p[idx];
    }
    partials[j * nchunks + chunk] = s;
}

__global__ void k_marginal_finish(const double* __restrict__ partials, int64_t nout, int64_t nchunks,
                                  double* __restrict__ out) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nout) return;
    double s = 0;
    for (int64_t c = 0; c < nchunks; ++c) s += partials[j * nchunks + c];
    out[j] = s;
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_probabilities(bsim_state* h, double* out) {
    if (int rc = check_state(h)) return rc;
    if (!out) return fail(BSIM_EINVAL, "out is null");
    size_t n = (size_t)1 << h->nq;
    if (int rc = ensure_ws(h->scratch, sizeof(double) * n)) return rc;
    if (int rc = probabilities_device(h, (double*)h->scratch.ptr)) return rc;
    BSIM_HIP(hipMemcpyAsync(out, h->scratch.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_marginal_probabilities(bsim_state* h, int nt, const int* targets, double* out) {
    if (int rc = check_state(h)) return rc;
    if (nt < 1 || nt > h->nq || !targets || !out) return fail(BSIM_EINVAL, "bad marginal targets");
    if (nt > 24) return fail(BSIM_ENOTSUP, "marginals over more than 24 qubits");
    std::vector<int> ts(nt);
    uint64_t used = 0;
    for (int i = 0; i < nt; ++i) {
        if (targets[i] < 0 || targets[i] >= h->nq) return fail(BSIM_EINVAL, "target qubit out of range");
        ts[i] = h->nq - 1 - targets[i];
        if (used & (1ull << ts[i])) return fail(BSIM_EINVAL, "repeated target qubit");
        used |= 1ull << ts[i];
    }
    std::sort(ts.begin(), ts.end());
    int64_t n = (int64_t)1 << h->nq, nout = (int64_t)1 << nt;
    int rest_bits = h->nq - nt;
    int chunk_bits = rest_bits > 10 ? rest_bits - 10 : 0;
    while (chunk_bits > 0 && (nout << (rest_bits - chunk_bits)) < (1 << 20)) --chunk_bits;
    int64_t nchunks = (int64_t)1 << (rest_bits - chunk_bits);
    size_t bytes = sizeof(double) * (size_t)(n + nout * nchunks + nout) + 64 * sizeof(int);
    if (int rc = ensure_ws(h->scratch, bytes)) return rc;
    double* p = (double*)h->scratch.ptr;
    double* partials = p + n;
    double* dout = partials + nout * nchunks;
    int* dts = (int*)(dout + nout);
    if (int rc = probabilities_device(h, p)) return rc;
    BSIM_HIP(hipMemcpyAsync(dts, ts.data(), sizeof(int) * nt, hipMemcpyHostToDevice, h->stream));
    int64_t nthreads = nout * nchunks;
    hipLaunchKernelGGL(k_marginal, dim3((unsigned)((nthreads + RBLOCK - 1) / RBLOCK)), dim3(RBLOCK), 0, h->stream, p,
                       h->nq, nt, dts, nchunks, chunk_bits, partials);
    BSIM_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_marginal_finish, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, h->stream, partials,
                       nout, nchunks, dout);
    BSIM_HIP(hipGetLastError());
    BSIM_HIP(hipMemcpyAsync(out, dout, sizeof(double) * nout, hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_expect_z_all(bsim_state* h, double* out) {
    if (int rc = check_state(h)) return rc;
    if (!out) return fail(BSIM_EINVAL, "out is null");
    if (int rc = ensure_ws(h->res, 4096)) return rc;
    double* dres = (double*)h->res.ptr;
    int layout = -1;
    if (int rc = zsums_device(h, dres, &layout)) return rc;
    double host[37];
    BSIM_HIP(hipMemcpyAsync(host, dres, sizeof(host), hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    for (int q = 0; q < h->nq; ++q) {
        int b = h->nq - 1 - q;
        // SV layout: bits below `layout` at 16 + b, bits above at b - layout
        double s = layout < 0 ? host[b] : (b < layout ? host[16 + b] : host[b - layout]);
        out[q] = host[36] - 2.0 * s;
    }
    return BSIM_OK;
}

int bsim_expect_pauli(bsim_state* h, int n_words, const char* paulis, double* out) {
    if (int rc = check_state(h)) return rc;
    if (n_words < 0 || (n_words > 0 && (!paulis || !out))) return fail(BSIM_EINVAL, "bad Pauli words");
    if (n_words == 0) return BSIM_OK;
    if (int rc = ensure_ws(h->res, sizeof(double) * (2 * (size_t)n_words + 64))) return rc;
    double* dres = (double*)h->res.ptr;
    for (int w = 0; w < n_words; ++w) {
        PauliMask pm;
        if (int rc = parse_pauli(h->nq, paulis + (size_t)w * h->nq, pm)) return rc;
        if (int rc = pauli_expect_device(h, pm, dres + 2 * w)) return rc;
    }
    std::vector<double> host(2 * (size_t)n_words);
    BSIM_HIP(hipMemcpyAsync(host.data(), dres, sizeof(double) * host.size(), hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    for (int w = 0; w < n_words; ++w) out[w] = host[2 * w];
    return BSIM_OK;
}

int bsim_expect_pauli_sum(bsim_state* h, int n_words, const char* paulis, const double* coeffs, double* out) {
    if (!coeffs || !out) return fail(BSIM_EINVAL, "coeffs/out is null");
    std::vector<double> v((size_t)(n_words > 0 ? n_words : 0));
    if (int rc = bsim_expect_pauli(h, n_words, paulis, v.data())) return rc;
    double s = 0;
    for (int w = 0; w < n_words; ++w) s += coeffs[w] * v[w];
    *out = s;
    return BSIM_OK;
}

int bsim_expect_matrix(bsim_state* h, int k, const int* targets, const double* m, double* out) {
    if (int rc = check_state(h)) return rc;
    if (k < 1 || k > 6 || !targets || !m || !out) return fail(BSIM_EINVAL, "bad observable arguments");
    if (k > h->nq) return fail(BSIM_EINVAL, "more targets than qubits");
    double res[2];
    if (h->kind == BSIM_SV) {
        // reference conventions: 1q standard; 2q t1 = MSB (src/sv_simulator.jl:185-195);
        // k > 2 sorted endian targets (src/sv_simulator.jl:207-218)
        GateSpec g;
        if (k == 2) {
            int t[2] = { targets[1], targets[0] };
            if (int rc = make_spec(h->vq, 2, t, 0, nullptr, nullptr, m, false, false, false, g)) return rc;
        } else {
            if (int rc = make_spec(h->vq, k, targets, 0, nullptr, nullptr, m, false, false, k > 2, g)) return rc;
        }
        int64_t n = h->len;
        int grid = red_grid(n >> k);
        if (int rc = ensure_ws(h->red, sizeof(double) * ((size_t)grid * 2 + 64))) return rc;
        if (int rc = ensure_ws(h->res, 4096)) return rc;
        double* partials = (double*)h->red.ptr;
        double* dres = (double*)h->res.ptr;
        if (k <= 2) {
            int ins[2] = { g.tb[0], k > 1 ? g.tb[1] : 0 };
            if (k == 2 && ins[0] > ins[1]) std::swap(ins[0], ins[1]);
            size_t D = (size_t)1 << k;
            if (int rc = ensure_ws(h->scratch, sizeof(double2) * D * D)) return rc;
            BSIM_HIP(hipMemcpyAsync(h->scratch.ptr, g.m.data(), sizeof(double2) * D * D, hipMemcpyHostToDevice,
                                    h->stream));
            if (k == 1)
                hipLaunchKernelGGL(k_expect_dense<1>, dim3(grid), dim3(RBLOCK), 0, h->stream, h->d, n >> 1,
                                   make_int4(ins[0], 0, 0, 0), g.tb[0], 0, (const double2*)h->scratch.ptr, partials);
            else
                hipLaunchKernelGGL(k_expect_dense<2>, dim3(grid), dim3(RBLOCK), 0, h->stream, h->d, n >> 2,
                                   make_int4(ins[0], ins[1], 0, 0), g.tb[0], g.tb[1], (const double2*)h->scratch.ptr,
                                   partials);
            BSIM_HIP(hipGetLastError());
            hipLaunchKernelGGL(k_finish, dim3(2), dim3(256), 0, h->stream, partials, grid, 2, dres);
            BSIM_HIP(hipGetLastError());
        } else {
            // copy, apply, dot (the reference's apply_observable on a deepcopy)
            if (int rc = ensure_ws(h->lam, sizeof(double2) * (size_t)n)) return rc;
            BSIM_HIP(hipMemcpyAsync(h->lam.ptr, h->d, sizeof(double2) * (size_t)n, hipMemcpyDeviceToDevice, h->stream));
            if (int rc = launch_gate((double2*)h->lam.ptr, h->vq, g, h->scratch, h->stream)) return rc;
            if (int rc = dot_device(h->d, (const double2*)h->lam.ptr, n, h->red, h->stream, dres)) return rc;
        }
        BSIM_HIP(hipMemcpyAsync(res, dres, sizeof(res), hipMemcpyDeviceToHost, h->stream));
        BSIM_HIP(hipStreamSynchronize(h->stream));
    } else {
        std::vector<int> ord(k);
        uint64_t used = 0;
        for (int i = 0; i < k; ++i) {
            if (targets[i] < 0 || targets[i] >= h->nq) return fail(BSIM_EINVAL, "target qubit out of range");
            ord[i] = h->nq - 1 - targets[i];
            if (used & (1ull << ord[i])) return fail(BSIM_EINVAL, "repeated target qubit");
            used |= 1ull << ord[i];
        }
        std::sort(ord.begin(), ord.end());
        size_t D = (size_t)1 << k;
        int64_t work = ((int64_t)1 << h->nq) * (int64_t)D;
        int grid = red_grid(work);
        if (int rc = ensure_ws(h->red, sizeof(double) * ((size_t)grid * 2 + 64))) return rc;
        if (int rc = ensure_ws(h->res, 4096)) return rc;
        double* partials = (double*)h->red.ptr;
        double* dres = (double*)h->res.ptr;
        if (int rc = ensure_ws(h->scratch, sizeof(double2) * D * D + 64)) return rc;
        BSIM_HIP(hipMemcpyAsync(h->scratch.ptr, m, sizeof(double2) * D * D, hipMemcpyHostToDevice, h->stream));
        int* dord = (int*)((double2*)h->scratch.ptr + D * D);
        BSIM_HIP(hipMemcpyAsync(dord, ord.data(), sizeof(int) * k, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_expect_dm_matrix, dim3(grid), dim3(RBLOCK), 0, h->stream, h->d, h->nq, k, dord,
                           (const double2*)h->scratch.ptr, partials);
        BSIM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_finish, dim3(2), dim3(256), 0, h->stream, partials, grid, 2, dres);
        BSIM_HIP(hipGetLastError());
        BSIM_HIP(hipMemcpyAsync(res, dres, sizeof(res), hipMemcpyDeviceToHost, h->stream));
        BSIM_HIP(hipStreamSynchronize(h->stream));
    }
    *out = res[0];
    return BSIM_OK;
}

}  // extern "C"
