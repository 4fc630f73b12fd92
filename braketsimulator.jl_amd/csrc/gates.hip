// Per-gate kernels: one HBM pass per gate, the drop-in for the reference's
// apply_gate! / apply_controlled_gate! loops (src/gate_kernels.jl:489-775,
// src/custom_gates.jl:364-416).
//
// Kernel families:
//  - k_1q_lane:   single-target gates on amplitude bits 0..5.  Every lane
//                 loads one consecutive amplitude (fully coalesced 1 KiB per
//                 wavefront load) and receives its partner from lane
//                 (lane ^ 2^e) by DPP (e = 0, 1), ds_swizzle (e = 2..4) or
//                 ds_bpermute (e = 5).
//  - k_dense<K>:  K = 1, 2 targets with the matrix in kernel arguments;
//                 threads own one amplitude group (strided pairs/quads whose
//                 low halves are consecutive across lanes).
//  - k_dense_mfma: K = 4, 5 targets on the FP64 matrix cores (the real
//                 form of the gate as a GEMM over amplitude groups);
//                 k_dense_mfma6 the same for K = 6, rows split over waves.
//  - k_dense_lds: K = 3 targets; matrix and amplitude groups staged in
//                 LDS, one output amplitude per thread.
//  - k_diag:      diagonal gates, one coalesced pass over the control
//                 subspace.
#include <string.h>

#include <algorithm>
#include <atomic>
#include <vector>

#include "gates.h"

namespace bsim {

// ---------------------------------------------------------------------------
// spec construction
// ---------------------------------------------------------------------------

int make_spec(int vq, int k, const int* targets, int nc, const int* controls, const int* cvals, const double* m,
              bool diag, bool conj, bool sorted, GateSpec& g) {
    if (k < 1 || k > (diag ? 8 : 12)) return fail(BSIM_ENOTSUP, "gate target count must be 1..12 (diagonal 1..8)");
    if (nc < 0 || nc > BSIM_MAX_CONTROLS) return fail(BSIM_ENOTSUP, "at most 8 controls");
    if (k + nc > 16) return fail(BSIM_ENOTSUP, "at most 16 targets and controls together");
    if (!m) return fail(BSIM_EINVAL, "matrix pointer is null");
    uint64_t used = 0;
    g = GateSpec();
    g.k = k;
    g.diag = diag;
    for (int b = 0; b < k; ++b) {
        int q = targets[b];
        if (q < 0 || q >= vq) return fail(BSIM_EINVAL, "target qubit out of range");
        int e = vq - 1 - q;
        if (used & (1ull << e)) return fail(BSIM_EINVAL, "repeated target/control qubit");
        used |= 1ull << e;
        g.tb[b] = e;
    }
    if (sorted) std::sort(g.tb, g.tb + k);
    for (int c = 0; c < nc; ++c) {
        int q = controls[c];
        if (q < 0 || q >= vq) return fail(BSIM_EINVAL, "control qubit out of range");
        int e = vq - 1 - q;
        if (used & (1ull << e)) return fail(BSIM_EINVAL, "repeated target/control qubit");
        used |= 1ull << e;
        g.cmask |= 1ull << e;
        if (cvals == nullptr || cvals[c]) g.cval |= 1ull << e;
    }
    size_t n = diag ? ((size_t)1 << k) : ((size_t)1 << (2 * k));
    g.m.resize(n);
    for (size_t i = 0; i < n; ++i) g.m[i] = make_double2(m[2 * i], conj ? -m[2 * i + 1] : m[2 * i + 1]);
    return BSIM_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

struct InsArgs {
    int nins;       // number of bits inserted (targets + controls), ascending
    int ins[16];
    uint64_t cval;  // control values OR-ed into every group base
    int64_t ngroups;
};

__host__ __device__ __forceinline__ uint64_t group_base(uint64_t g, const InsArgs& a) {
    for (int i = 0; i < a.nins; ++i) g = pad_bit(g, a.ins[i]);
    return g | a.cval;
}

template <int K>
struct SmallMat {
    double2 m[1 << (2 * K)];
};

template <int K>
__global__ __launch_bounds__(256) void k_dense(double2* __restrict__ sv, InsArgs a, SmallMat<K> M, int t0, int t1) {
    constexpr int D = 1 << K;
    const uint64_t b0 = 1ull << t0;
    const uint64_t b1 = K > 1 ? (1ull << t1) : 0;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.ngroups; g += stride) {
        uint64_t base = group_base((uint64_t)g, a);
        uint64_t off[D];
        double2 v[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            off[j] = base | ((j & 1) ? b0 : 0) | ((j & 2) ? b1 : 0);
            v[j] = sv[off[j]];
        }
        double2 r[D];
#pragma unroll
        for (int row = 0; row < D; ++row) {
            double2 acc = cmul(M.m[row], v[0]);
#pragma unroll
            for (int col = 1; col < D; ++col) acc = cfma(M.m[row + D * col], v[col], acc);
            r[row] = acc;
        }
#pragma unroll
        for (int j = 0; j < D; ++j) sv[off[j]] = r[j];
    }
}

// Lane exchange of a double2 with lane ^ (1 << E).
template <int E>
__device__ __forceinline__ double2 xor_lane(double2 v) {
    int w[4];
    memcpy(w, &v, 16);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if constexpr (E == 0) {
            w[i] = __builtin_amdgcn_update_dpp(0, w[i], 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
        } else if constexpr (E == 1) {
            w[i] = __builtin_amdgcn_update_dpp(0, w[i], 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
        } else if constexpr (E <= 4) {
            w[i] = __builtin_amdgcn_ds_swizzle(w[i], ((1 << E) << 10) | 0x1F);    // bitmask mode: xor
        } else {
            w[i] = __shfl_xor(w[i], 32, 64);
        }
    }
    double2 r;
    memcpy(&r, w, 16);
    return r;
}

template <int E>
__global__ __launch_bounds__(256) void k_1q_lane(double2* __restrict__ sv, int64_t len, uint64_t cmask, uint64_t cval,
                                                 double2 m00, double2 m10, double2 m01, double2 m11) {
    // len and the grid stride are multiples of 64: whole wavefronts stay
    // active together, so the lane exchange always has its partner.
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        double2 a = sv[i];
        double2 b = xor_lane<E>(a);
        if (((uint64_t)i & cmask) == cval) {
            double2 out = ((i >> E) & 1) ? cfma(m10, b, cmul(m11, a)) : cfma(m01, b, cmul(m00, a));
            sv[i] = out;
        }
    }
}

// K = 3: matrix (D x D) and a block's amplitude groups in LDS.
__global__ __launch_bounds__(256) void k_dense_lds(double2* __restrict__ sv, InsArgs a, const double2* __restrict__ mg,
                                                   int K, int4 tlo, int2 thi) {
    extern __shared__ double2 smem[];
    const int D = 1 << K;
    double2* M = smem;            // D*D
    double2* A = smem + D * D;    // 256 amplitudes (256/D groups)
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) M[i] = mg[i];
    int tb[6] = { tlo.x, tlo.y, tlo.z, tlo.w, thi.x, thi.y };
    const int gpb = 256 / D;  // groups per block
    int local_g = threadIdx.x / D, r = threadIdx.x % D;
    int64_t nblk = (a.ngroups + gpb - 1) / gpb;
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        int64_t g = blk * gpb + local_g;
        bool valid = g < a.ngroups;
        uint64_t off = 0;
        if (valid) {
            uint64_t base = group_base((uint64_t)g, a);
            off = base;
            for (int b = 0; b < K; ++b)
                if ((r >> b) & 1) off |= 1ull << tb[b];
        }
        __syncthreads();
        if (valid) A[threadIdx.x] = sv[off];
        __syncthreads();
        if (valid) {
            const double2* col = A + local_g * D;
            double2 acc = cmul(M[r], col[0]);
            for (int c = 1; c < D; ++c) acc = cfma(M[r + D * c], col[c], acc);
            sv[off] = acc;
        }
    }
}

// K = 7..12 targets (the k-qubit fallback beyond the matrix-core kernels).
// A workgroup holds GB groups (GB D = 2048 amplitudes, or one group of D =
// 4096) in LDS and computes NOUT = GB D / 256 outputs per thread: for D <=
// 256 thread t owns row t % D of NOUT groups, so each matrix element it
// reads (column-major, consecutive threads on consecutive rows) serves NOUT
// outputs; for D > 256 it owns rows t, t + 256, ... of every group.  The
// matrix (4^K x 16 B) streams through L2 once per workgroup trip.
template <int K>
__global__ __launch_bounds__(256) void k_dense_big(double2* __restrict__ sv, InsArgs a,
                                                   const double2* __restrict__ mg, int4 tlo, int4 tmid, int4 thi) {
    constexpr int D = 1 << K;
    constexpr int GB = D >= 2048 ? 1 : 2048 / D;
    constexpr int NA = GB * D;  // amplitudes staged per trip
    constexpr int NOUT = NA / 256;
    extern __shared__ double2 A[];  // NA
    const int tb[12] = { tlo.x, tlo.y, tlo.z, tlo.w, tmid.x, tmid.y, tmid.z, tmid.w, thi.x, thi.y, thi.z, thi.w };
    auto offset = [&](int j) {  // target-bit part of local amplitude j of a group
        uint64_t o = 0;
#pragma unroll
        for (int b = 0; b < K; ++b)
            if ((j >> b) & 1) o |= 1ull << tb[b];
        return o;
    };
    const int64_t trips = (a.ngroups + GB - 1) / GB;
    for (int64_t tr = blockIdx.x; tr < trips; tr += gridDim.x) {
        const int64_t g0 = tr * GB;
        __syncthreads();  // the previous trip's reads of A are done
        for (int q = threadIdx.x; q < NA; q += 256) {
            const int64_t g = g0 + q / D;
            if (g < a.ngroups) A[q] = sv[group_base((uint64_t)g, a) | offset(q % D)];
        }
        __syncthreads();
        double2 acc[NOUT];
#pragma unroll
        for (int i = 0; i < NOUT; ++i) acc[i] = make_double2(0.0, 0.0);
        if constexpr (D <= 256) {
            const int r = threadIdx.x % D, gi0 = threadIdx.x / D;
            for (int c = 0; c < D; ++c) {
                const double2 m = mg[r + (size_t)D * c];
#pragma unroll
                for (int i = 0; i < NOUT; ++i) acc[i] = cfma(m, A[(gi0 + i * (256 / D)) * D + c], acc[i]);
            }
#pragma unroll
            for (int i = 0; i < NOUT; ++i) {
                const int64_t g = g0 + gi0 + i * (256 / D);
                if (g < a.ngroups) sv[group_base((uint64_t)g, a) | offset(r)] = acc[i];
            }
        } else {
            constexpr int RPT = D / 256;  // rows per thread, per group
            for (int c = 0; c < D; ++c) {
#pragma unroll
                for (int j = 0; j < RPT; ++j) {
                    const double2 m = mg[threadIdx.x + 256 * j + (size_t)D * c];
#pragma unroll
                    for (int gi = 0; gi < GB; ++gi) acc[gi * RPT + j] = cfma(m, A[gi * D + c], acc[gi * RPT + j]);
                }
            }
#pragma unroll
            for (int gi = 0; gi < GB; ++gi) {
                const int64_t g = g0 + gi;
                if (g < a.ngroups) {
                    const uint64_t base = group_base((uint64_t)g, a);
#pragma unroll
                    for (int j = 0; j < RPT; ++j) sv[base | offset(threadIdx.x + 256 * j)] = acc[gi * RPT + j];
                }
            }
        }
    }
}

// K = 4, 5 on the matrix cores.  A dense K-qubit gate is the real GEMM
//   [Re y; Im y] = [[Re M, -Im M], [Im M, Re M]] [Re x; Im x]
// (2D x 2D times 2D x groups, D = 2^K) done with v_mfma_f64_16x16x4f64:
// A = 16 x 4 blocks of the real matrix (lane l: row l % 16, column l / 16),
// B = 4 x 16 blocks of the amplitudes (lane l: row l / 16, group l % 16),
// D = 16 x 16 accumulator tiles (lane l, register r: row 4 r + l / 16,
// group l % 16; measured on gfx950 with a one-hot probe).
// A workgroup stages 64 groups (2^(6+K) amplitudes) in LDS; element e of the
// block takes its bits, in ascending order, at the block's local positions
// (the 6 lowest non-target bits and the K target bits, sorted), so that
// consecutive threads load and store consecutive addresses wherever the
// targets are.  Each wave multiplies 16 groups with the matrix blocks
// (`areal`, prepared by the host in lane order) in registers (K = 4) or in
// LDS (K = 5).
struct MfmaLayout {
    int nloc;          // 6 + K local bits
    int16_t code[12];  // local bit i: target b (code b) or group bit t (code 16 + t)
};

template <int K>
__global__ __launch_bounds__(256) void k_dense_mfma(double2* __restrict__ sv, InsArgs a,
                                                    const double* __restrict__ areal, int4 tlo, int thi,
                                                    MfmaLayout lay) {
    constexpr int D = 1 << K, RT = 2 * D / 16, KS = 2 * D / 4, GB = 64, PAD = D + 1, NA = RT * KS;
    constexpr bool A_IN_LDS = K >= 5;
    __shared__ double2 X[GB * PAD];
    __shared__ double AL[A_IN_LDS ? NA * 64 : 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double av[A_IN_LDS ? 1 : NA];
    if constexpr (A_IN_LDS) {
        for (int i = threadIdx.x; i < NA * 64; i += 256) AL[i] = areal[i];
    } else {
#pragma unroll
        for (int i = 0; i < NA; ++i) av[i] = areal[(size_t)i * 64 + lane];
    }
    const int tb[5] = { tlo.x, tlo.y, tlo.z, tlo.w, thi };
    // this thread's elements e = threadIdx.x + 256 i: LDS slot (group, amplitude)
    // and address offset from the block's base (the deposit of the group's 6
    // low bits at the 6 lowest free positions, the amplitude's at the targets)
    int slot[D / 4];
    uint64_t loc[D / 4];
#pragma unroll
    for (int i = 0; i < D / 4; ++i) {
        const int e = threadIdx.x + 256 * i;
        int gc = 0, j = 0;
        for (int b = 0; b < lay.nloc; ++b) {
            const int bit = (e >> b) & 1, cd = lay.code[b];
            if (cd >= 16) gc |= bit << (cd - 16);
            else j |= bit << cd;
        }
        slot[i] = gc * PAD + j;
        uint64_t off = group_base((uint64_t)gc, a) & ~a.cval;
#pragma unroll
        for (int b = 0; b < K; ++b)
            if ((j >> b) & 1) off |= 1ull << tb[b];
        loc[i] = off | ((uint64_t)gc << 58);  // group index (6 bits) kept in the top bits for the range test
    }
    const int64_t nblk = (a.ngroups + GB - 1) / GB;
    const int c = 16 * wave + (lane & 15), kq = lane >> 4;
    // software pipeline: the next block's amplitudes are loaded into registers
    // while this block's products run on the matrix cores
    double2 pre[D / 4];
    auto fetch = [&](int64_t blk) {
        const int64_t g0 = blk * GB;
        const uint64_t base = group_base((uint64_t)g0, a);  // g0 is a multiple of 64
#pragma unroll
        for (int i = 0; i < D / 4; ++i) {
            const int64_t g = g0 + (int64_t)(loc[i] >> 58);
            pre[i] = g < a.ngroups ? sv[base | (loc[i] & ((1ull << 58) - 1))] : make_double2(0.0, 0.0);
        }
    };
    if (blockIdx.x < nblk) fetch(blockIdx.x);
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int64_t g0 = blk * GB;
        const uint64_t base = group_base((uint64_t)g0, a);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < D / 4; ++i) X[slot[i]] = pre[i];
        __syncthreads();
        if (blk + gridDim.x < nblk) fetch(blk + gridDim.x);
        double bv[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int q = 4 * ks + kq;
            bv[ks] = q < D ? X[c * PAD + q].x : X[c * PAD + q - D].y;
        }
        // the RT accumulator tiles are independent: interleave them so that
        // consecutive MFMAs do not wait for each other
        typedef double d4 __attribute__((ext_vector_type(4)));
        d4 acc[RT];
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) acc[rt] = d4{ 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                const double am = A_IN_LDS ? AL[(rt * KS + ks) * 64 + lane] : av[A_IN_LDS ? 0 : rt * KS + ks];
                acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(am, bv[ks], acc[rt], 0, 0, 0);
            }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * rt + 4 * r + kq;
                if (row < D) X[c * PAD + row].x = acc[rt][r];
                else X[c * PAD + row - D].y = acc[rt][r];
            }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < D / 4; ++i) {
            const int64_t g = g0 + (int64_t)(loc[i] >> 58);
            if (g < a.ngroups) sv[base | (loc[i] & ((1ull << 58) - 1))] = X[slot[i]];
        }
    }
}

// K = 6: the real form is 128 x 128 (128 KiB), too large for LDS beside the
// groups, so the rows are split over the waves instead: each of 8 waves
// computes one 16-row tile for all 64 groups and keeps those 16 rows of the
// matrix (32 operands per lane) in registers.
struct Mfma6Hi {  // the parts of elements e = t + 512 i from i's bits (host-computed)
    int gc[8], sl[8];
    uint64_t off[8];
};

__global__ __launch_bounds__(512) void k_dense_mfma6(double2* __restrict__ sv, InsArgs a,
                                                      const double* __restrict__ areal, int4 tlo, int2 thi,
                                                      MfmaLayout lay, Mfma6Hi hi) {
    constexpr int K = 6, D = 64, KS = 2 * D / 4, GB = 64, PAD = D + 1, NE = D / 8;
    __shared__ double2 X[GB * PAD];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double av[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) av[ks] = areal[((size_t)wave * KS + ks) * 64 + lane];
    const int tb[6] = { tlo.x, tlo.y, tlo.z, tlo.w, thi.x, thi.y };
    // element e = threadIdx.x + 512 i: its group in the block, LDS slot and
    // address offset from the block's base.  Each is the OR (sum) of a part
    // from e's low 9 bits (the thread's, computed here) and a part from i's
    // bits (uniform, in the kernel arguments).
    int gct = 0, slt;
    uint64_t offt;
    {
        int j = 0;
        for (int b = 0; b < 9; ++b) {
            const int bit = (threadIdx.x >> b) & 1, cd = lay.code[b];
            if (cd >= 16) gct |= bit << (cd - 16);
            else j |= bit << cd;
        }
        slt = gct * PAD + j;
        offt = group_base((uint64_t)gct, a) & ~a.cval;
#pragma unroll
        for (int b = 0; b < K; ++b)
            if ((j >> b) & 1) offt |= 1ull << tb[b];
    }
    const int64_t nblk = (a.ngroups + GB - 1) / GB;
    const int cl = lane & 15, kq = lane >> 4;
    // software pipeline: the next block's amplitudes are loaded into registers
    // while this block's products run (one 8-wave workgroup fits a CU either
    // way, so the registers cost no occupancy)
    double2 pre[NE];
    auto fetch = [&](int64_t blk) {
        const int64_t g0 = blk * GB;
        const uint64_t base = group_base((uint64_t)g0, a);  // g0 is a multiple of 64
#pragma unroll
        for (int i = 0; i < NE; ++i)
            pre[i] = g0 + (gct | hi.gc[i]) < a.ngroups ? sv[base | offt | hi.off[i]] : make_double2(0.0, 0.0);
    };
    if (blockIdx.x < nblk) fetch(blockIdx.x);
    for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
        const int64_t g0 = blk * GB;
        const uint64_t base = group_base((uint64_t)g0, a);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NE; ++i) X[slt + hi.sl[i]] = pre[i];
        __syncthreads();
        if (blk + gridDim.x < nblk) fetch(blk + gridDim.x);
        typedef double d4 __attribute__((ext_vector_type(4)));
        d4 acc[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[ct] = d4{ 0.0, 0.0, 0.0, 0.0 };
        // B operands one k-step ahead; the scheduling barrier keeps the
        // compiler from hoisting every LDS read of the unrolled loop (which
        // costs a register per read)
        auto bload = [&](int ks, double* b) {
            const int q = 4 * ks + kq;
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const double2& z = X[(16 * ct + cl) * PAD + (q & (D - 1))];
                b[ct] = q < D ? z.x : z.y;
            }
        };
        double bn[4];
        bload(0, bn);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            double bc[4] = { bn[0], bn[1], bn[2], bn[3] };
            if (ks + 1 < KS) bload(ks + 1, bn);
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
                acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[ks], bc[ct], acc[ct], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
        __syncthreads();  // every wave has read all of X
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * wave + 4 * r + kq, c = 16 * ct + cl;
                if (row < D) X[c * PAD + row].x = acc[ct][r];
                else X[c * PAD + row - D].y = acc[ct][r];
            }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NE; ++i) {
            if (g0 + (gct | hi.gc[i]) < a.ngroups) sv[base | offt | hi.off[i]] = X[slt + hi.sl[i]];
        }
    }
}

// A blocks of the real form of the D x D complex matrix m (column-major), in
// the order k_dense_mfma reads them: block (rt, ks), lane l holds
// R[16 rt + l % 16][4 ks + l / 16].
static void mfma_operands(const std::vector<double2>& m, int K, std::vector<double>& out) {
    const int D = 1 << K, RT = 2 * D / 16, KS = 2 * D / 4;
    out.assign((size_t)RT * KS * 64, 0.0);
    auto R = [&](int r, int k) {
        const double2 z = m[(size_t)(r % D) + (size_t)D * (k % D)];
        if (r < D) return k < D ? z.x : -z.y;
        return k < D ? z.y : z.x;
    };
    for (int rt = 0; rt < RT; ++rt)
        for (int ks = 0; ks < KS; ++ks)
            for (int l = 0; l < 64; ++l) out[((size_t)rt * KS + ks) * 64 + l] = R(16 * rt + l % 16, 4 * ks + l / 16);
}

// Diagonal: amplitude idx (in the control subspace) times d[j(idx)].
__global__ __launch_bounds__(256) void k_diag(double2* __restrict__ sv, InsArgs a, const double2* __restrict__ dg,
                                              int K, int4 tlo, int4 thi) {
    int tb[8] = { tlo.x, tlo.y, tlo.z, tlo.w, thi.x, thi.y, thi.z, thi.w };
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.ngroups; g += stride) {
        uint64_t idx = group_base((uint64_t)g, a);
        int j = 0;
        for (int b = 0; b < K; ++b) j |= (int)((idx >> tb[b]) & 1) << b;
        sv[idx] = cmul(dg[j], sv[idx]);
    }
}

// src/custom_gates.jl:390-416
__global__ __launch_bounds__(256) void k_double_exc(double2* __restrict__ sv, int64_t ngroups, int4 ord, int4 e,
                                                    double c, double s) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ngroups; g += stride) {
        uint64_t p = pad_bit(pad_bit(pad_bit(pad_bit((uint64_t)g, ord.x), ord.y), ord.z), ord.w);
        uint64_t i0011 = p ^ (1ull << e.z) ^ (1ull << e.w);
        uint64_t i1100 = p ^ (1ull << e.x) ^ (1ull << e.y);
        double2 a = sv[i0011], b = sv[i1100];
        sv[i0011] = make_double2(c * a.x - s * b.x, c * a.y - s * b.y);
        sv[i1100] = make_double2(s * a.x + c * b.x, s * a.y + c * b.y);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

static InsArgs ins_args(int vq, const GateSpec& g, bool include_targets) {
    InsArgs a;
    memset(&a, 0, sizeof(a));
    int n = 0;
    uint64_t mask = g.cmask;
    if (include_targets)
        for (int b = 0; b < g.k; ++b) mask |= 1ull << g.tb[b];
    for (int bit = 0; bit < vq; ++bit)
        if (mask & (1ull << bit)) a.ins[n++] = bit;
    a.nins = n;
    a.cval = g.cval;
    a.ngroups = (int64_t)1 << (vq - n);
    return a;
}

// Grid cap of the per-gate kernels (bsim_set_max_blocks; 0 = 64 per CU).
static std::atomic<int> g_max_blocks{ 0 };
static int64_t cap_blocks() {
    const int m = g_max_blocks.load(std::memory_order_relaxed);
    return m > 0 ? (int64_t)m : (int64_t)device_cu_count() * 64;
}

// Matrices and operands go to the handle's scratch workspace with
// hipMemcpyAsync from host vectors that end with this call: a copy from
// pageable host memory has consumed its source when the call returns (it is
// staged), and the stream orders it after the kernels that still read the
// previous contents, so no gate waits for the device.
int launch_gate(double2* d, int vq, const GateSpec& g, Workspace& scratch, hipStream_t s) {
    if (g.diag) {
        if (g.k > 8) return fail(BSIM_ENOTSUP, "diagonal gate on more than 8 targets");
        InsArgs a = ins_args(vq, g, false);
        size_t n = (size_t)1 << g.k;
        if (int rc = ensure_ws(scratch, sizeof(double2) * n)) return rc;
        BSIM_HIP(hipMemcpyAsync(scratch.ptr, g.m.data(), sizeof(double2) * n, hipMemcpyHostToDevice, s));
        int tb[8] = {0};
        for (int b = 0; b < g.k; ++b) tb[b] = g.tb[b];
        int64_t grid = grid_for(a.ngroups, 256, cap_blocks());
        hipLaunchKernelGGL(k_diag, dim3((unsigned)grid), dim3(256), 0, s, d, a, (const double2*)scratch.ptr, g.k,
                           make_int4(tb[0], tb[1], tb[2], tb[3]), make_int4(tb[4], tb[5], tb[6], tb[7]));
        BSIM_HIP(hipGetLastError());
        return BSIM_OK;
    }
    if (g.k > vq) return fail(BSIM_EINVAL, "more targets than qubits");
    if (g.k == 1 && g.tb[0] < 6 && vq >= 6) {
        int64_t len = (int64_t)1 << vq;
        int64_t grid = grid_for(len, 256, cap_blocks());
        uint64_t cm = g.cmask, cv = g.cval;
        const double2* m = g.m.data();
        switch (g.tb[0]) {
#define LANE_CASE(E)                                                                                            \
    case E:                                                                                                     \
        hipLaunchKernelGGL(k_1q_lane<E>, dim3((unsigned)grid), dim3(256), 0, s, d, len, cm, cv, m[0], m[1], m[2], \
                           m[3]);                                                                               \
        break;
            LANE_CASE(0)
            LANE_CASE(1)
            LANE_CASE(2)
            LANE_CASE(3)
            LANE_CASE(4)
            LANE_CASE(5)
#undef LANE_CASE
        }
        BSIM_HIP(hipGetLastError());
        return BSIM_OK;
    }
    InsArgs a = ins_args(vq, g, true);
    int64_t grid = grid_for(a.ngroups, 256, cap_blocks());
    if (g.k == 1) {
        SmallMat<1> M;
        for (int i = 0; i < 4; ++i) M.m[i] = g.m[i];
        hipLaunchKernelGGL(k_dense<1>, dim3((unsigned)grid), dim3(256), 0, s, d, a, M, g.tb[0], 0);
    } else if (g.k == 2) {
        SmallMat<2> M;
        for (int i = 0; i < 16; ++i) M.m[i] = g.m[i];
        hipLaunchKernelGGL(k_dense<2>, dim3((unsigned)grid), dim3(256), 0, s, d, a, M, g.tb[0], g.tb[1]);
    } else if (g.k >= 4 && g.k <= 6) {
        std::vector<double> av;
        mfma_operands(g.m, g.k, av);
        if (int rc = ensure_ws(scratch, sizeof(double) * av.size())) return rc;
        BSIM_HIP(hipMemcpyAsync(scratch.ptr, av.data(), sizeof(double) * av.size(), hipMemcpyHostToDevice, s));
        const int64_t nblk = grid_for((a.ngroups + 63) / 64, 1, cap_blocks());
        const int4 tlo = make_int4(g.tb[0], g.tb[1], g.tb[2], g.tb[3]);
        const int thi = g.k > 4 ? g.tb[4] : 0;
        const int2 thi6 = make_int2(thi, g.k > 5 ? g.tb[5] : 0);
        // local bits of a block of 64 groups: the 6 lowest bits that are neither
        // targets nor controls, and the targets, in ascending position order
        MfmaLayout lay;
        memset(&lay, 0, sizeof(lay));
        {
            uint64_t tmask = 0, used = 0;
            for (int b = 0; b < g.k; ++b) tmask |= 1ull << g.tb[b];
            for (int i = 0; i < a.nins; ++i) used |= 1ull << a.ins[i];
            int ng = 0, n = 0;
            for (int bit = 0; bit < vq && n < 6 + g.k; ++bit) {
                if (tmask & (1ull << bit)) {
                    int b = 0;
                    while (g.tb[b] != bit) ++b;
                    lay.code[n++] = (int16_t)b;
                } else if (!(used & (1ull << bit)) && ng < 6) {
                    lay.code[n++] = (int16_t)(16 + ng++);
                }
            }
            // fewer than 6 free bits (tiny states): the remaining group bits on top
            while (ng < 6) lay.code[n++] = (int16_t)(16 + ng++);
            lay.nloc = n;
        }
        if (g.k == 4)
            hipLaunchKernelGGL(k_dense_mfma<4>, dim3((unsigned)nblk), dim3(256), 0, s, d, a,
                               (const double*)scratch.ptr, tlo, thi, lay);
        else if (g.k == 5)
            hipLaunchKernelGGL(k_dense_mfma<5>, dim3((unsigned)nblk), dim3(256), 0, s, d, a,
                               (const double*)scratch.ptr, tlo, thi, lay);
        else {
            Mfma6Hi hi;
            for (int i = 0; i < 8; ++i) {
                int gc = 0, j = 0;
                for (int b = 9; b < lay.nloc; ++b) {
                    const int bit = ((512 * i) >> b) & 1, cd = lay.code[b];
                    if (cd >= 16) gc |= bit << (cd - 16);
                    else j |= bit << cd;
                }
                uint64_t off = group_base((uint64_t)gc, a) & ~a.cval;
                for (int b = 0; b < 6; ++b)
                    if ((j >> b) & 1) off |= 1ull << g.tb[b];
                hi.gc[i] = gc;
                hi.sl[i] = gc * 65 + j;
                hi.off[i] = off;
            }
            hipLaunchKernelGGL(k_dense_mfma6, dim3((unsigned)nblk), dim3(512), 0, s, d, a,
                               (const double*)scratch.ptr, tlo, thi6, lay, hi);
        }
    } else if (g.k == 3) {
        size_t D = (size_t)1 << g.k;
        if (int rc = ensure_ws(scratch, sizeof(double2) * D * D)) return rc;
        BSIM_HIP(hipMemcpyAsync(scratch.ptr, g.m.data(), sizeof(double2) * D * D, hipMemcpyHostToDevice, s));
        int tb[6] = {0};
        for (int b = 0; b < g.k; ++b) tb[b] = g.tb[b];
        int64_t gpb = 256 / (int64_t)D;
        int64_t nblk = grid_for((a.ngroups + gpb - 1) / gpb, 1, cap_blocks());
        size_t shmem = sizeof(double2) * (D * D + 256);  // 5 KiB
        hipLaunchKernelGGL(k_dense_lds, dim3((unsigned)nblk), dim3(256), shmem, s, d, a, (const double2*)scratch.ptr,
                           g.k, make_int4(tb[0], tb[1], tb[2], tb[3]), make_int2(tb[4], tb[5]));
    } else if (g.k <= 12) {
        const size_t D = (size_t)1 << g.k;
        if (int rc = ensure_ws(scratch, sizeof(double2) * D * D)) return rc;
        BSIM_HIP(hipMemcpyAsync(scratch.ptr, g.m.data(), sizeof(double2) * D * D, hipMemcpyHostToDevice, s));
        int tb[12] = {0};
        for (int b = 0; b < g.k; ++b) tb[b] = g.tb[b];
        const int4 tlo = make_int4(tb[0], tb[1], tb[2], tb[3]), tmid = make_int4(tb[4], tb[5], tb[6], tb[7]),
                   thi = make_int4(tb[8], tb[9], tb[10], tb[11]);
        const int64_t gb = D >= 2048 ? 1 : 2048 / (int64_t)D;
        const int64_t trips = (a.ngroups + gb - 1) / gb;
        const unsigned nblk = (unsigned)std::min<int64_t>(trips, cap_blocks());
        const size_t shmem = sizeof(double2) * (size_t)(gb * D);
        const double2* mg = (const double2*)scratch.ptr;
        switch (g.k) {
#define BIG_CASE(K)                                                                                            \
    case K:                                                                                                    \
        BSIM_HIP(hipFuncSetAttribute((const void*)k_dense_big<K>, hipFuncAttributeMaxDynamicSharedMemorySize, \
                                     (int)shmem));                                                             \
        hipLaunchKernelGGL(k_dense_big<K>, dim3(nblk), dim3(256), shmem, s, d, a, mg, tlo, tmid, thi);         \
        break;
            BIG_CASE(7)
            BIG_CASE(8)
            BIG_CASE(9)
            BIG_CASE(10)
            BIG_CASE(11)
            BIG_CASE(12)
#undef BIG_CASE
        }
    } else {
        return fail(BSIM_ENOTSUP, "dense gates act on at most 12 targets");
    }
    BSIM_HIP(hipGetLastError());
    return BSIM_OK;
}

int launch_double_excitation(double2* d, int vq, const int* e, double c, double sn, hipStream_t s) {
    int ord[4] = { e[0], e[1], e[2], e[3] };
    std::sort(ord, ord + 4);
    int64_t ng = (int64_t)1 << (vq - 4);
    int64_t grid = grid_for(ng, 256, cap_blocks());
    hipLaunchKernelGGL(k_double_exc, dim3((unsigned)grid), dim3(256), 0, s, d, ng,
                       make_int4(ord[0], ord[1], ord[2], ord[3]), make_int4(e[0], e[1], e[2], e[3]), c, sn);
    BSIM_HIP(hipGetLastError());
    return BSIM_OK;
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_set_max_blocks(int blocks, int* old) {
    if (blocks < 0) return fail(BSIM_EINVAL, "blocks must be >= 0");
    if (old) *old = g_max_blocks.load();
    g_max_blocks.store(blocks);
    return BSIM_OK;
}

int bsim_apply_1q(bsim_state* h, int q, const double* m, int conj) {
    if (int rc = check_state(h)) return rc;
    GateSpec g;
    if (int rc = make_spec(h->vq, 1, &q, 0, nullptr, nullptr, m, false, conj != 0, false, g)) return rc;
    h->alias_ready = false;
    return launch_gate(h->d, h->vq, g, h->scratch, h->stream);
}

int bsim_apply_2q(bsim_state* h, int t1, int t2, const double* m, int conj) {
    if (int rc = check_state(h)) return rc;
    int t[2] = { t1, t2 };
    GateSpec g;
    if (int rc = make_spec(h->vq, 2, t, 0, nullptr, nullptr, m, false, conj != 0, false, g)) return rc;
    h->alias_ready = false;
    return launch_gate(h->d, h->vq, g, h->scratch, h->stream);
}

int bsim_apply_controlled(bsim_state* h, int nc, const int* controls, const int* cvals, int nt, const int* targets,
                          const double* m, int conj) {
    if (int rc = check_state(h)) return rc;
    if (nt < 1 || nt > 2) return fail(BSIM_ENOTSUP, "controlled gates take 1 or 2 targets");
    if (nc < 1 || !controls) return fail(BSIM_EINVAL, "controlled gate without controls");
    GateSpec g;
    if (int rc = make_spec(h->vq, nt, targets, nc, controls, cvals, m, false, conj != 0, false, g)) return rc;
    h->alias_ready = false;
    return launch_gate(h->d, h->vq, g, h->scratch, h->stream);
}

int bsim_apply_kq(bsim_state* h, int k, const int* targets, const double* m, int conj) {
    if (int rc = check_state(h)) return rc;
    if (k < 1 || k > 12) return fail(BSIM_ENOTSUP, "k-qubit dense gates take 1..12 targets");
    GateSpec g;
    if (int rc = make_spec(h->vq, k, targets, 0, nullptr, nullptr, m, false, conj != 0, true, g)) return rc;
    h->alias_ready = false;
    if (int rc = launch_gate(h->d, h->vq, g, h->scratch, h->stream)) return rc;
    // the matrix was staged from g.m, which ends with this call
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_apply_diagonal(bsim_state* h, int k, const int* targets, const double* dvals, int conj) {
    if (int rc = check_state(h)) return rc;
    GateSpec g;
    if (int rc = make_spec(h->vq, k, targets, 0, nullptr, nullptr, dvals, true, conj != 0, false, g)) return rc;
    h->alias_ready = false;
    if (int rc = launch_gate(h->d, h->vq, g, h->scratch, h->stream)) return rc;
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_apply_double_excitation(bsim_state* h, const int* t, double angle) {
    if (int rc = check_state(h)) return rc;
    if (!t) return fail(BSIM_EINVAL, "targets is null");
    int e[4];
    uint64_t used = 0;
    for (int i = 0; i < 4; ++i) {
        if (t[i] < 0 || t[i] >= h->vq) return fail(BSIM_EINVAL, "target qubit out of range");
        e[i] = h->vq - 1 - t[i];
        if (used & (1ull << e[i])) return fail(BSIM_EINVAL, "repeated target qubit");
        used |= 1ull << e[i];
    }
    h->alias_ready = false;
    return launch_double_excitation(h->d, h->vq, e, cos(angle / 2.0), sin(angle / 2.0), h->stream);
}

}  // extern "C"
