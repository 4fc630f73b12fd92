// Density-matrix noise (src/noise_kernels.jl).
//
// Reference semantics (mode 0): apply_noise!(::Kraus, ...) visits the
// 2^(n-k) padded indices p and replaces the block rho[S(p), S(p)] -- the
// rows and columns that agree on every non-target bit -- by
// sum_m K_m B K_m^dagger (src/noise_kernels.jl:105-196).  One workgroup
// owns one block; the d x d block and each product stay in LDS (k <= 5) or,
// for k = 6..8, in a slice of device memory of the workgroup's own.
//
// Physical semantics (mode 1) and PhaseFlip are lowered by the planner to a
// superoperator / a 2-bit diagonal on the 4^n vector and run through the
// gate kernels (fused when planned).
#include <string.h>

#include <algorithm>
#include <vector>

#include "fused.h"

namespace bsim {

struct KrausArgs {
    int k;
    int et[8];   // endian target of channel qubit b, given order
    int od[8];   // sorted endian targets
};

template <bool GLOBAL>
__global__ __launch_bounds__(1024) void k_kraus_ref(double2* __restrict__ dm, int nq, KrausArgs a, int nK,
                                                    const double2* __restrict__ K, double2* __restrict__ work) {
    extern __shared__ double2 lds[];
    const int k = a.k;
    const int d = 1 << k, dd = d * d;
    double2* sh = GLOBAL ? work + (size_t)blockIdx.x * 3 * dd : lds;
    double2* B = sh;
    double2* Tm = sh + dd;
    double2* O = sh + 2 * dd;
    const int64_t N = (int64_t)1 << nq;
    const int64_t nblk = N >> k;
    for (int64_t pi = blockIdx.x; pi < nblk; pi += gridDim.x) {
        uint64_t p = (uint64_t)pi;
        for (int b = 0; b < k; ++b) p = pad_bit(p, a.od[b]);
        auto ix = [&](int j) -> uint64_t {
            uint64_t x = p;
            if (k == 1) return j ? (p ^ (1ull << a.et[0])) : p;
            if (k == 2) {  // (00, flip t2, flip t1, 11): noise_kernels.jl:145-155
                if (j & 1) x ^= 1ull << a.et[1];
                if (j & 2) x ^= 1ull << a.et[0];
                return x;
            }
            for (int b = 0; b < k; ++b)  // sorted endian targets (noise_kernels.jl:175-186)
                if ((j >> b) & 1) x ^= 1ull << a.od[b];
            return x;
        };
        for (int t = threadIdx.x; t < dd; t += blockDim.x) {
            B[t] = dm[ix(t % d) + ix(t / d) * N];
            O[t] = make_double2(0, 0);
        }
        __syncthreads();
        for (int m = 0; m < nK; ++m) {
            const double2* km = K + (size_t)m * dd;
            for (int t = threadIdx.x; t < dd; t += blockDim.x) {  // Tm = K_m B
                const int r = t % d, c = t / d;
                double2 acc = cmul(km[r], B[c * d]);
                for (int l = 1; l < d; ++l) acc = cfma(km[r + l * d], B[l + c * d], acc);
                Tm[t] = acc;
            }
            __syncthreads();
            for (int t = threadIdx.x; t < dd; t += blockDim.x) {  // O += Tm K_m^dagger
                const int r = t % d, c = t / d;
                double2 acc = cmul(Tm[r], cconj(km[c]));
                for (int l = 1; l < d; ++l) acc = cfma(Tm[r + l * d], cconj(km[c + l * d]), acc);
                O[t] = m == 0 ? acc : cadd(O[t], acc);
            }
            __syncthreads();
        }
        for (int t = threadIdx.x; t < dd; t += blockDim.x) dm[ix(t % d) + ix(t / d) * N] = O[t];
        __syncthreads();
    }
}

int launch_kraus_ref(bsim_state* h, double2* d, const EOp& e) {
    const int k = e.kq, nq = h->nq;
    if (k > 8) return fail(BSIM_ENOTSUP, "Kraus channels act on at most 8 qubits");
    const int D = 1 << k, dd = D * D;
    KrausArgs a;
    memset(&a, 0, sizeof(a));
    a.k = k;
    for (int b = 0; b < k; ++b) a.od[b] = a.et[b] = nq - 1 - e.ktargets[b];
    std::sort(a.od, a.od + k);
    size_t bytes = sizeof(double2) * e.kraus.size();
    if (int rc = ensure_ws(h->scratch, bytes)) return rc;
    BSIM_HIP(hipMemcpyAsync(h->scratch.ptr, e.kraus.data(), bytes, hipMemcpyHostToDevice, h->stream));
    int64_t nblk = ((int64_t)1 << nq) >> k;
    int block = dd < 64 ? 64 : (dd > 1024 ? 1024 : dd);
    if (k <= 5) {
        int64_t grid = nblk < 65536 ? nblk : 65536;
        hipLaunchKernelGGL(k_kraus_ref<false>, dim3((unsigned)grid), dim3(block), sizeof(double2) * 3 * dd, h->stream,
                           d, nq, a, e.nkraus, (const double2*)h->scratch.ptr, (double2*)nullptr);
    } else {
        // block, product and sum of each workgroup in device memory (3 d^2
        // amplitudes per workgroup: 12 MiB for 64 workgroups at k = 6)
        const int64_t grid = nblk < 64 ? nblk : 64;
        if (int rc = ensure_ws(h->alias_work, sizeof(double2) * 3 * (size_t)dd * (size_t)grid)) return rc;
        h->alias_ready = false;
        hipLaunchKernelGGL(k_kraus_ref<true>, dim3((unsigned)grid), dim3(block), 0, h->stream, d, nq, a, e.nkraus,
                           (const double2*)h->scratch.ptr, (double2*)h->alias_work.ptr);
    }
    BSIM_HIP(hipGetLastError());
    return BSIM_OK;
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_dm_kraus(bsim_state* h, int k, const int* targets, int n_kraus, const double* kraus, int mode) {
    if (int rc = check_state(h)) return rc;
    if (h->kind != BSIM_DM) return fail(BSIM_EINVAL, "Kraus channels need a density matrix");
    if (k < 1 || k > 8 || !targets || !kraus || n_kraus < 1) return fail(BSIM_EINVAL, "bad Kraus arguments");
    bsim_op op = {};
    op.type = mode ? BSIM_OP_KRAUS_PHYS : BSIM_OP_KRAUS_REF;
    op.n_targets = k;
    for (int b = 0; b < k; ++b) op.targets[b] = targets[b];
    op.n_kraus = n_kraus;
    op.matrix = kraus;
    std::vector<EOp> eops;
    if (int rc = lower_ops(h->nq, h->kind, &op, 1, eops, false)) return rc;
    h->alias_ready = false;
    for (const EOp& e : eops)
        if (int rc = launch_eop(h, h->d, e)) return rc;
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_dm_apply_superop(bsim_state* h, int k, const int* targets, const double* S) {
    if (int rc = check_state(h)) return rc;
    if (h->kind != BSIM_DM) return fail(BSIM_EINVAL, "superoperators need a density matrix");
    if (k < 1 || k > 5 || !targets || !S) return fail(BSIM_EINVAL, "bad superoperator arguments");
    bsim_op op = {};
    op.type = BSIM_OP_SUPEROP;
    op.n_targets = k;
    for (int b = 0; b < k; ++b) op.targets[b] = targets[b];
    op.matrix = S;
    std::vector<EOp> eops;
    if (int rc = lower_ops(h->nq, h->kind, &op, 1, eops, false)) return rc;
    h->alias_ready = false;
    for (const EOp& e : eops)
        if (int rc = launch_eop(h, h->d, e)) return rc;
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_dm_phaseflip(bsim_state* h, int q, double p) {
    if (int rc = check_state(h)) return rc;
    if (h->kind != BSIM_DM) return fail(BSIM_EINVAL, "PhaseFlip needs a density matrix");
    bsim_op op = {};
    op.type = BSIM_OP_PHASEFLIP;
    op.n_targets = 1;
    op.targets[0] = q;
    double pp[2] = { p, 0 };
    op.matrix = pp;
    std::vector<EOp> eops;
    if (int rc = lower_ops(h->nq, h->kind, &op, 1, eops, false)) return rc;
    h->alias_ready = false;
    for (const EOp& e : eops)
        if (int rc = launch_eop(h, h->d, e)) return rc;
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

}  // extern "C"
