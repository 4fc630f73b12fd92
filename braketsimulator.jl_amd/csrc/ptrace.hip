// Reduced density matrices (partial_trace, src/sv_simulator.jl:272-316 and
// src/dm_simulator.jl:280-323; calculate(::DensityMatrix),
// src/result_types.jl:240-246).
//
// Output index a (0 <= a < 2^k) holds target b in bit k-1-b (the first
// target is the most significant bit); x(a, r) is the state index whose
// target bits hold a and whose other bits hold r (in increasing order).
//   state vector:   out[i, j] = sum_r psi[x(i, r)] conj(psi[x(j, r)])
//   density matrix: out[i, j] = sum_r stored[x(i, r), x(j, r)]
// (the reference traces the stored matrix as it is).
//
// State vector: a Gram product M M^dagger of the 2^k x 2^(n-k) matrix
// M[a, r] = psi[x(a, r)].  A workgroup owns one 16x16 tile of the output
// (or the whole output below 16 x 16) and a range of columns; 64-column
// slabs of the tile's rows are staged in LDS and every thread accumulates
// one output.  Per-workgroup partials are summed in a fixed order.
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.h"

namespace bsim {

constexpr int PT_BLOCK = 256;
static size_t pt_align(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr int PT_COLS = 64;

struct PtTables {
    uint64_t rest[4][256];  // deposit of byte q of r into the non-target bits
    uint64_t tgt[256];      // deposit of a into the target bits
};

__device__ __forceinline__ uint64_t pt_index(const PtTables& t, uint64_t a, uint64_t r) {
    return t.tgt[a] | t.rest[0][r & 255] | t.rest[1][(r >> 8) & 255] | t.rest[2][(r >> 16) & 255] |
           t.rest[3][(r >> 24) & 255];
}

__global__ __launch_bounds__(PT_BLOCK) void k_ptrace_sv(const double2* __restrict__ sv, const PtTables* __restrict__ tab,
                                                         int tile, int ntj, uint64_t ncols, uint64_t cols_per_block,
                                                         double2* __restrict__ partials) {
    __shared__ double2 A[16][PT_COLS];
    __shared__ double2 B[16][PT_COLS];
    const int ti = blockIdx.y / ntj, tj = blockIdx.y % ntj;
    const int t = threadIdx.x;
    const int ii = t / tile, jj = t % tile;
    const bool owner = t < tile * tile;
    const uint64_t c0 = (uint64_t)blockIdx.x * cols_per_block;
    uint64_t c1 = c0 + cols_per_block;
    if (c1 > ncols) c1 = ncols;
    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t cs = c0; cs < c1; cs += PT_COLS) {
        for (int e = t; e < 2 * tile * PT_COLS; e += PT_BLOCK) {
            const int which = e / (tile * PT_COLS), rem = e % (tile * PT_COLS);
            const int row = rem / PT_COLS, col = rem % PT_COLS;
            const uint64_t r = cs + col;
            const uint64_t a = (uint64_t)((which ? tj : ti) * tile + row);
            const double2 v = r < c1 ? sv[pt_index(*tab, a, r)] : make_double2(0.0, 0.0);
            if (which) B[row][col] = v;
            else A[row][col] = v;
        }
        __syncthreads();
        if (owner) {
#pragma unroll 8
            for (int c = 0; c < PT_COLS; ++c) {
                const double2 x = A[ii][c], y = B[jj][c];  // x * conj(y)
                acc.x = fma(x.x, y.x, fma(x.y, y.y, acc.x));
                acc.y = fma(x.y, y.x, fma(-x.x, y.y, acc.y));
            }
        }
        __syncthreads();
    }
    if (owner) partials[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * (tile * tile) + t] = acc;
}

// Density matrix: one thread per (output, column chunk), a gather of the
// elements whose row and column share the same non-target bits.
__global__ __launch_bounds__(PT_BLOCK) void k_ptrace_dm(const double2* __restrict__ dm, int nq, const PtTables* __restrict__ tab,
                                                         int nt, uint64_t ncols, uint64_t cols_per_chunk,
                                                         int64_t nchunks, double2* __restrict__ partials) {
    const int64_t D = (int64_t)1 << nt, nout = D * D;
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nout * nchunks) return;
    const int64_t o = idx % nout, ch = idx / nout;
    const uint64_t i = (uint64_t)(o % D), j = (uint64_t)(o / D);  // out[i, j], column-major
    const uint64_t c0 = (uint64_t)ch * cols_per_chunk;
    uint64_t c1 = c0 + cols_per_chunk;
    if (c1 > ncols) c1 = ncols;
    double2 acc = make_double2(0.0, 0.0);
    for (uint64_t r = c0; r < c1; ++r) {
        const uint64_t row = pt_index(*tab, i, r), col = pt_index(*tab, j, r);
        const double2 v = dm[row + (col << nq)];
        acc.x += v.x;
        acc.y += v.y;
    }
    partials[(size_t)ch * nout + o] = acc;
}

__global__ void k_ptrace_finish(const double2* __restrict__ partials, int64_t nout, int64_t nparts,
                                double2* __restrict__ out) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nout) return;
    double2 s = make_double2(0.0, 0.0);
    for (int64_t p = 0; p < nparts; ++p) {
        const double2 v = partials[(size_t)p * nout + o];
        s.x += v.x;
        s.y += v.y;
    }
    out[o] = s;
}

// SV partials are laid out per tile: gather them into the column-major output.
__global__ void k_ptrace_sv_finish(const double2* __restrict__ partials, int tile, int ntj, int64_t nx, int64_t D,
                                   double2* __restrict__ out) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= D * D) return;
    const int64_t i = o % D, j = o / D;
    const int64_t ti = i / tile, tj = j / tile, ii = i % tile, jj = j % tile;
    const int64_t ty = ti * ntj + tj, t = ii * tile + jj;
    double2 s = make_double2(0.0, 0.0);
    for (int64_t x = 0; x < nx; ++x) {
        const double2 v = partials[((size_t)ty * nx + x) * (tile * tile) + t];
        s.x += v.x;
        s.y += v.y;
    }
    out[o] = s;
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_partial_trace(bsim_state* h, int nt, const int* targets, double* out) {
    if (int rc = check_state(h)) return rc;
    const int nq = h->nq;
    if (nt < 0 || nt > nq || (nt > 0 && !targets) || !out) return fail(BSIM_EINVAL, "bad partial-trace arguments");
    if (nt > 8) return fail(BSIM_ENOTSUP, "reduced density matrices of at most 8 qubits");
    if (nq - nt > 32) return fail(BSIM_ENOTSUP, "more than 32 traced qubits");
    PtTables tab;
    memset(&tab, 0, sizeof(tab));
    uint64_t used = 0;
    for (int b = 0; b < nt; ++b) {
        if (targets[b] < 0 || targets[b] >= nq) return fail(BSIM_EINVAL, "target qubit out of range");
        const int pos = nq - 1 - targets[b];
        if (used & (1ull << pos)) return fail(BSIM_EINVAL, "repeated target qubit");
        used |= 1ull << pos;
    }
    const int64_t D = (int64_t)1 << nt;
    for (int64_t a = 0; a < D; ++a) {
        uint64_t x = 0;
        for (int b = 0; b < nt; ++b)
            if ((a >> (nt - 1 - b)) & 1) x |= 1ull << (nq - 1 - targets[b]);
        tab.tgt[a] = x;
    }
    int rest[64], nr = 0;
    for (int p = 0; p < nq; ++p)
        if (!(used & (1ull << p))) rest[nr++] = p;
    for (int q = 0; q < 4; ++q)
        for (int v = 0; v < 256; ++v) {
            uint64_t x = 0;
            for (int k = 0; k < 8; ++k)
                if (((v >> k) & 1) && 8 * q + k < nr) x |= 1ull << rest[8 * q + k];
            tab.rest[q][v] = x;
        }
    const uint64_t ncols = 1ull << nr;
    const int64_t nout = D * D;
    if (h->kind == BSIM_SV) {
        const int tile = (int)std::min<int64_t>(D, 16);
        const int ntj = (int)(D / tile), nty = ntj * ntj;
        // about 2048 workgroups in all, at least one 64-column slab each
        int64_t nx = std::max<int64_t>(1, 2048 / nty);
        uint64_t per = (ncols + nx - 1) / nx;
        per = ((per + PT_COLS - 1) / PT_COLS) * PT_COLS;
        nx = (int64_t)((ncols + per - 1) / per);
        const size_t part_bytes = sizeof(double2) * (size_t)nty * nx * tile * tile;
        const size_t bytes = pt_align(sizeof(PtTables)) + pt_align(part_bytes) + sizeof(double2) * nout;
        if (int rc = ensure_ws(h->scratch, bytes)) return rc;
        char* base = (char*)h->scratch.ptr;
        PtTables* dtab = (PtTables*)base;
        double2* partials = (double2*)(base + pt_align(sizeof(PtTables)));
        double2* dout = (double2*)(base + pt_align(sizeof(PtTables)) + pt_align(part_bytes));
        BSIM_HIP(hipMemcpyAsync(dtab, &tab, sizeof(tab), hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_ptrace_sv, dim3((unsigned)nx, (unsigned)nty), dim3(PT_BLOCK), 0, h->stream, h->d, dtab, tile,
                           ntj, ncols, per, partials);
        BSIM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ptrace_sv_finish, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, h->stream, partials,
                           tile, ntj, nx, D, dout);
        BSIM_HIP(hipGetLastError());
        BSIM_HIP(hipMemcpyAsync(out, dout, sizeof(double2) * nout, hipMemcpyDeviceToHost, h->stream));
    } else {
        // chunks of columns so that about 2^20 threads gather
        int64_t nch = std::max<int64_t>(1, std::min<int64_t>((int64_t)ncols, ((int64_t)1 << 20) / nout));
        const uint64_t per = (ncols + nch - 1) / nch;
        nch = (int64_t)((ncols + per - 1) / per);
        const size_t part_bytes = sizeof(double2) * (size_t)(nout * nch);
        const size_t bytes = pt_align(sizeof(PtTables)) + pt_align(part_bytes) + sizeof(double2) * nout;
        if (int rc = ensure_ws(h->scratch, bytes)) return rc;
        char* base = (char*)h->scratch.ptr;
        PtTables* dtab = (PtTables*)base;
        double2* partials = (double2*)(base + pt_align(sizeof(PtTables)));
        double2* dout = (double2*)(base + pt_align(sizeof(PtTables)) + pt_align(part_bytes));
        BSIM_HIP(hipMemcpyAsync(dtab, &tab, sizeof(tab), hipMemcpyHostToDevice, h->stream));
        const int64_t nthreads = nout * nch;
        hipLaunchKernelGGL(k_ptrace_dm, dim3((unsigned)((nthreads + PT_BLOCK - 1) / PT_BLOCK)), dim3(PT_BLOCK), 0,
                           h->stream, h->d, nq, dtab, nt, ncols, per, nch, partials);
        BSIM_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ptrace_finish, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, h->stream, partials,
                           nout, nch, dout);
        BSIM_HIP(hipGetLastError());
        BSIM_HIP(hipMemcpyAsync(out, dout, sizeof(double2) * nout, hipMemcpyDeviceToHost, h->stream));
    }
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

}  // extern "C"
