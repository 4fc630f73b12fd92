// Density-matrix noise (src/noise_kernels.jl).
//
// Reference semantics (mode 0): apply_noise!(::Kraus, ...) visits the
// 2^(n-k) padded indices p and replaces the block rho[S(p), S(p)] -- the
// rows and columns that agree on every non-target bit -- by
// sum_m K_m B K_m^dagger (src/noise_kernels.jl:105-196).  One workgroup
// owns one block; the d x d block and each product stay in LDS.
//
// Physical semantics (mode 1) and PhaseFlip are lowered by the planner to a
// superoperator / a 2-bit diagonal on the 4^n vector and run through the
// gate kernels (fused when planned).
#include <algorithm>
#include <vector>

#include "fused.h"

namespace bsim {

__global__ void k_kraus_ref(double2* __restrict__ dm, int nq, int k, int4 e, int4 ord, int nK,
                            const double2* __restrict__ K) {
    extern __shared__ double2 sh[];
    const int d = 1 << k, dd = d * d;
    double2* B = sh;
    double2* Tm = sh + dd;
    double2* O = sh + 2 * dd;
    const int64_t N = (int64_t)1 << nq;
    const int64_t nblk = N >> k;
    const int et[4] = { e.x, e.y, e.z, e.w };
    const int od[4] = { ord.x, ord.y, ord.z, ord.w };
    const int t = threadIdx.x;
    const int r = t % d, c = t / d;
    for (int64_t pi = blockIdx.x; pi < nblk; pi += gridDim.x) {
        uint64_t p = (uint64_t)pi;
        for (int b = 0; b < k; ++b) p = pad_bit(p, od[b]);
        auto ix = [&](int j) -> uint64_t {
            if (k == 1) return j ? (p ^ (1ull << et[0])) : p;
            if (k == 2) {  // (00, flip t2, flip t1, 11): noise_kernels.jl:145-155
                uint64_t x = p;
                if (j & 1) x ^= 1ull << et[1];
                if (j & 2) x ^= 1ull << et[0];
                return x;
            }
            uint64_t x = p;  // sorted endian targets (noise_kernels.jl:175-186)
            for (int b = 0; b < k; ++b)
                if ((j >> b) & 1) x ^= 1ull << od[b];
            return x;
        };
        uint64_t row = 0, col = 0;
        if (t < dd) {
            row = ix(r);
            col = ix(c);
            B[t] = dm[row + col * N];
            O[t] = make_double2(0, 0);
        }
        __syncthreads();
        for (int m = 0; m < nK; ++m) {
            const double2* km = K + (size_t)m * dd;
            if (t < dd) {  // Tm = K_m B
                double2 acc = cmul(km[r], B[c * d]);
                for (int l = 1; l < d; ++l) acc = cfma(km[r + l * d], B[l + c * d], acc);
                Tm[t] = acc;
            }
            __syncthreads();
            if (t < dd) {  // O += Tm K_m^dagger
                double2 acc = cmul(Tm[r], cconj(km[c]));
                for (int l = 1; l < d; ++l) acc = cfma(Tm[r + l * d], cconj(km[c + l * d]), acc);
                O[t] = m == 0 ? acc : cadd(O[t], acc);
            }
            __syncthreads();
        }
        if (t < dd) dm[row + col * N] = O[t];
        __syncthreads();
    }
}

int launch_kraus_ref(bsim_state* h, double2* d, const EOp& e) {
    const int k = e.kq, nq = h->nq;
    const int D = 1 << k, dd = D * D;
    int et[4] = { 0, 0, 0, 0 }, od[4] = { 0, 0, 0, 0 };
    for (int b = 0; b < k; ++b) od[b] = et[b] = nq - 1 - e.ktargets[b];
    std::sort(od, od + k);
    size_t bytes = sizeof(double2) * e.kraus.size();
    if (int rc = ensure_ws(h->scratch, bytes)) return rc;
    BSIM_HIP(hipMemcpyAsync(h->scratch.ptr, e.kraus.data(), bytes, hipMemcpyHostToDevice, h->stream));
    int64_t nblk = ((int64_t)1 << nq) >> k;
    int block = dd < 64 ? 64 : dd;
    int64_t grid = nblk < 65536 ? nblk : 65536;
    hipLaunchKernelGGL(k_kraus_ref, dim3((unsigned)grid), dim3(block), sizeof(double2) * 3 * dd, h->stream, d, nq, k,
                       make_int4(et[0], et[1], et[2], et[3]), make_int4(od[0], od[1], od[2], od[3]), e.nkraus,
                       (const double2*)h->scratch.ptr);
    BSIM_HIP(hipGetLastError());
    return BSIM_OK;
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_dm_kraus(bsim_state* h, int k, const int* targets, int n_kraus, const double* kraus, int mode) {
    if (int rc = check_state(h)) return rc;
    if (h->kind != BSIM_DM) return fail(BSIM_EINVAL, "Kraus channels need a density matrix");
    if (k < 1 || k > 4 || !targets || !kraus || n_kraus < 1) return fail(BSIM_EINVAL, "bad Kraus arguments");
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
