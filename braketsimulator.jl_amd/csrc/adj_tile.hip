// Fused backward sweep of the state-vector adjoint gradient.
//
// The backward sweep visits the ops in reverse; for op k (matrix U, and for
// each angle p its derivative dU_p):
//     psi <- U^dagger psi;   g_p += 2 Re <lambda| dU_p |psi>;   lambda <- U^dagger lambda.
// With both vectors updated first the product is the same number,
// <lambda|dU|psi> = <U^dagger lambda| U^dagger dU |psi>, so each op runs on
// the two vectors in one dispatch and its payloads hold U^dagger dU.
// The reversed list of U^dagger ops is planned like a forward evolution
// (tile passes of 11 bits, sweeps of 3 register bits); a pass loads the tiles
// of both vectors into LDS, and every thread carries two 8-amplitude
// sub-cubes (psi and lambda) through the sweep's ops, computing the
// products of its sub-cube after each parametrized op.  Each
// wave's sum goes to partials[((slot * PM + p) * ntiles + tile) * W + wave]
// (PM: the largest parameter count of an op, W waves per workgroup), summed
// later in a fixed order.  Reordering ops on
// disjoint bits inside a pass does not change these terms: both vectors see
// the same unitaries.
#include <atomic>
#include <string.h>

#include <algorithm>
#include <functional>
#include <vector>

#include "adj_tile_body.h"
#include "fused.h"

namespace bsim {

__global__ __launch_bounds__(AT_TB) void k_adj_tile(double2* __restrict__ psi, double2* __restrict__ lam, int vq,
                                                    const Pass* __restrict__ pass, const Sweep* __restrict__ sweeps,
                                                    const TOp* __restrict__ tops, double* __restrict__ partials,
                                                    int pm, int64_t ntiles) {
    extern __shared__ double2 lds[];
    __shared__ uint64_t scat_lo[AT_TB];
    __shared__ uint64_t scat_hi[64];
    adj_tile_body(psi, lam, vq, pass, sweeps, tops, partials, pm, ntiles, blockIdx.x, lds, scat_lo, scat_hi);
}

// 11-bit passes: persistent workgroups (two per CU) taking tiles by ticket,
// the next tile prefetched during the last sweep.
__global__ __launch_bounds__(AT_TB, 2) void k_adj_tile_stream(double2* __restrict__ psi, double2* __restrict__ lam,
                                                              int vq, const Pass* __restrict__ pass,
                                                              const Sweep* __restrict__ sweeps,
                                                              const TOp* __restrict__ tops,
                                                              double* __restrict__ partials, int pm, int64_t ntiles,
                                                              uint64_t pass_tiles,
                                                              unsigned long long* __restrict__ ctr) {
    extern __shared__ double2 lds[];
    __shared__ uint64_t scat_lo[AT_TB];
    __shared__ uint64_t scat_hi[64];
    __shared__ uint64_t s_next;
    adj_stream_body(psi, lam, vq, pass, sweeps, tops, partials, pm, ntiles, pass_tiles, ctr, lds, scat_lo, scat_hi,
                    &s_next);
}

// out[slot * 3 + p] = sum over the n partial sums of row slot * pm + p, in a
// fixed order.
__global__ __launch_bounds__(256) void k_adj_tile_finish(const double* __restrict__ partials, int64_t n, int pm,
                                                         double* __restrict__ out) {
    __shared__ double sh[256];
    const double r = adj_row_sum(partials + (size_t)blockIdx.x * n, n, sh);
    if (threadIdx.x == 0) out[(blockIdx.x / pm) * 3 + blockIdx.x % pm] = r;
}

__global__ void k_adj_zero(double* __restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0;
}

int adjoint_back_plan(int vq, const std::vector<EOp>& eops, std::vector<EOp>& back, Plan& plan, int& nslots,
                      int& pm, const TileConfig* base) {
    const int L = (int)eops.size();
    nslots = 0;
    pm = 0;
    for (const EOp& e : eops) nslots += e.nparams > 0 ? 1 : 0;
    back.clear();
    back.reserve(eops.size());
    int slot = nslots;
    for (int k = L - 1; k >= 0; --k) {
        EOp b = eops[k];
        if (b.kind != E_DENSE && b.kind != E_DIAG) return fail(BSIM_EINVAL, "adjoint: state-vector ops only");
        if (b.g.diag) {
            for (auto& z : b.g.m) z.y = -z.y;
        } else {
            const int D = 1 << b.g.k;
            for (int r = 0; r < D; ++r)
                for (int c = 0; c < D; ++c) {
                    const double2 x = eops[k].g.m[c + (size_t)D * r];
                    b.g.m[r + (size_t)D * c] = make_double2(x.x, -x.y);
                }
        }
        if (b.nparams > 0) {
            b.aslot = --slot;
            pm = std::max(pm, b.nparams);
            // payload U^dagger dU: the products are taken after both vectors
            // are updated, <lam|dU|psi'> = <U^dagger lam| U^dagger dU |psi'>
            const int D = 1 << b.g.k;
            for (int p = 0; p < b.nparams; ++p) {
                std::vector<double2> t((size_t)D * D, make_double2(0.0, 0.0));
                for (int r = 0; r < D; ++r)
                    for (int c = 0; c < D; ++c) {
                        double2 acc = make_double2(0.0, 0.0);
                        for (int k2 = 0; k2 < D; ++k2) {
                            const double2 u = b.g.diag ? (k2 == r ? b.g.m[r] : make_double2(0.0, 0.0))
                                                       : b.g.m[r + (size_t)D * k2];
                            const double2 d = b.dm[p][k2 + (size_t)D * c];
                            acc.x += u.x * d.x - u.y * d.y;
                            acc.y += u.x * d.y + u.y * d.x;
                        }
                        t[r + (size_t)D * c] = acc;
                    }
                b.dm[p].swap(t);
            }
        }
        back.push_back(std::move(b));
    }
    TileConfig cfg = base ? *base : tile_config();
    cfg.tile_bits = 11;
    cfg.reg_bits = 3;
    cfg.merge_dense = false;  // single steps run by index in the reversed list
    return make_plan(vq, back, cfg, plan);
}

static std::atomic<bool> g_adj_shmem{ false };

static std::atomic<int> g_adj_slots{ 0 };

static int adjoint_tile_prepare() {
    if (g_adj_shmem) return BSIM_OK;
    const int shmem = (int)(2 * sizeof(double2) * ((1 << 11) + (1 << 7)));
    for (const void* k : { (const void*)k_adj_tile, (const void*)k_adj_tile_stream })
        BSIM_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, shmem));
    // workgroup slots of the streaming kernel: resident workgroups per CU
    // (two: 72 KiB of LDS each) times the CU count
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_adj_tile_stream, AT_TB, shmem) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    g_adj_slots = per_cu * device_cu_count();
    g_adj_shmem = true;
    return BSIM_OK;
}

size_t adjoint_tile_partial_bytes(const Plan& plan, int vq, int nslots, int pm) {
    int64_t nt = 1;
    for (const Pass& P : plan.passes) nt = std::max<int64_t>(nt, (int64_t)1 << (vq - P.T));
    return sizeof(double) * (size_t)nslots * (size_t)std::max(pm, 1) * (size_t)nt * AT_W;
}

int adjoint_tile_run(bsim_state* h, double2* psi, double2* lam, const Plan& plan, int nslots, int pm, double* out3,
                     const std::function<int(int)>& single) {
    if (int rc = adjoint_tile_prepare()) return rc;
    int64_t nt = 1;
    for (const Pass& P : plan.passes) {
        if (P.R != 3 || P.T > 11) return fail(BSIM_ENOTSUP, "adjoint tile pass shape");
        nt = std::max<int64_t>(nt, (int64_t)1 << (h->vq - P.T));
    }
    pm = std::max(pm, 1);
    const int64_t total = (int64_t)nslots * pm * nt * AT_W;
    if (int rc = ensure_ws(h->red, sizeof(double) * (size_t)std::max<int64_t>(total, 1))) return rc;
    double* partials = (double*)h->red.ptr;
    if (total > 0) {
        hipLaunchKernelGGL(k_adj_zero, dim3((unsigned)std::min<int64_t>((total + 255) / 256, 8192)), dim3(256), 0,
                           h->stream, partials, total);
        BSIM_HIP(hipGetLastError());
    }
    if (!plan.passes.empty()) {
        if (int rc = ensure_ws(h->ops, plan_table_bytes(plan))) return rc;
        if (int rc = upload_plan(plan, (char*)h->ops.ptr, h->stream)) return rc;
    }
    const char* tables = (const char*)h->ops.ptr;
    const size_t off_s = (sizeof(Pass) * plan.passes.size() + 255) & ~(size_t)255;
    const size_t off_o = off_s + ((sizeof(Sweep) * plan.sweeps.size() + 255) & ~(size_t)255);
    for (const Plan::Step& st : plan.steps) {
        if (st.tile) {
            const Pass& P = plan.passes[st.pass];
            const unsigned grid = (unsigned)(1ull << (h->vq - P.T));
            const size_t shmem = 2 * sizeof(double2) * (size_t)((1 << P.T) + ((1 << P.T) >> 4));
            if (P.T == 11 && h->vq >= 11) {
                const unsigned g = std::min<unsigned>(grid, (unsigned)g_adj_slots.load());
                hipLaunchKernelGGL(k_adj_tile_stream, dim3(g), dim3(AT_TB), shmem, h->stream, psi, lam, h->vq,
                                   (const Pass*)tables + st.pass, (const Sweep*)(tables + off_s),
                                   (const TOp*)(tables + off_o), partials, pm, nt, (uint64_t)grid, h->tile_ctr);
            } else {
                hipLaunchKernelGGL(k_adj_tile, dim3(grid), dim3(AT_TB), shmem, h->stream, psi, lam, h->vq,
                                   (const Pass*)tables + st.pass, (const Sweep*)(tables + off_s),
                                   (const TOp*)(tables + off_o), partials, pm, nt);
            }
            BSIM_HIP(hipGetLastError());
        } else {
            if (int rc = single(st.op)) return rc;
        }
    }
    if (total > 0) {
        hipLaunchKernelGGL(k_adj_tile_finish, dim3((unsigned)(nslots * pm)), dim3(256), 0, h->stream,
                           (const double*)partials, nt * AT_W, pm, out3);
        BSIM_HIP(hipGetLastError());
    }
    return BSIM_OK;
}

}  // namespace bsim
