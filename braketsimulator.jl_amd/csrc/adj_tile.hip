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
#include <array>
#include <cmath>
#include <complex>
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

namespace {

using cplx = std::complex<double>;
using M2 = std::array<cplx, 4>;  // 2 x 2, column-major: m[r + 2 c]

bool one_qubit(const EOp& e) {
    return (e.kind == E_DENSE || e.kind == E_DIAG) && e.g.k == 1 && e.g.cmask == 0;
}

M2 mat1(const EOp& e) {
    auto c = [](double2 x) { return cplx(x.x, x.y); };
    if (e.g.diag) return { c(e.g.m[0]), 0.0, 0.0, c(e.g.m[1]) };
    return { c(e.g.m[0]), c(e.g.m[1]), c(e.g.m[2]), c(e.g.m[3]) };
}

M2 mat1(const std::vector<double2>& d) {
    return { cplx(d[0].x, d[0].y), cplx(d[1].x, d[1].y), cplx(d[2].x, d[2].y), cplx(d[3].x, d[3].y) };
}

M2 mul(const M2& a, const M2& b) {
    M2 r;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) r[i + 2 * j] = a[i] * b[2 * j] + a[i + 2] * b[2 * j + 1];
    return r;
}

double maxabs(const M2& m) {
    double s = 0.0;
    for (const cplx& x : m) s = std::max({ s, std::fabs(x.real()), std::fabs(x.imag()) });
    return s;
}

// a m = m a to 1e-14 of m's size.  The payloads are formed in floating
// point from derivative matrices, so the exact commutation of a Pauli
// rotation's -i P / 2 with the rotation shows up as a residue of rounding.
bool commutes(const M2& a, const M2& m) {
    const M2 x = mul(a, m), y = mul(m, a);
    M2 d;
    for (int i = 0; i < 4; ++i) d[i] = x[i] - y[i];
    return maxabs(d) <= 1e-14 * maxabs(m);
}

}  // namespace

// Pauli form of a product payload (2 x 2, column-major): m = -i s P with
// P = X (1), Y (2) or Z (3) and s real; 0 when m has no such form.
// Entries below 1e-15 of the largest are rounding of an exact zero, and
// entries that agree to that much are taken as equal.
int pauli_form(const double2* m, double& s) {
    double sc = 0.0;
    for (int j = 0; j < 4; ++j) sc = std::max({ sc, std::fabs(m[j].x), std::fabs(m[j].y) });
    if (sc == 0.0) return 0;
    const double eps = 1e-15 * sc;
    auto z = [&](double x) { return std::fabs(x) <= eps; };
    auto eq = [&](double x, double y) { return std::fabs(x - y) <= eps; };
    const double2 m00 = m[0], m10 = m[1], m01 = m[2], m11 = m[3];
    const bool diag0 = z(m00.x) && z(m00.y) && z(m11.x) && z(m11.y);
    const bool off0 = z(m01.x) && z(m01.y) && z(m10.x) && z(m10.y);
    if (diag0 && z(m01.x) && z(m10.x) && eq(m01.y, m10.y)) {  // -i s X: m01 = m10 = -i s
        s = -m01.y;
        return 1;
    }
    if (diag0 && z(m01.y) && z(m10.y) && eq(m01.x, -m10.x)) {  // -i s Y = s [[0, -1], [1, 0]]
        s = m10.x;
        return 2;
    }
    if (off0 && z(m00.x) && z(m11.x) && eq(m00.y, -m11.y)) {  // -i s Z: m00 = -i s, m11 = i s
        s = -m00.y;
        return 3;
    }
    return 0;
}

// Runs of uncontrolled one-qubit ops on one qubit, each the next op on
// that qubit in the backward list (the ops between act on other qubits and
// commute with the run), become one dense op: the run's product W.  A
// product is taken after W when its payload commutes with the run's ops
// after its own, or before W when it commutes with the run's ops up to its
// own (its value is unchanged: <lam'|M|psi'> over the moved vectors equals
// the original when the unitaries between commute with M).  The kernels
// take the products of a merged op in two shapes: all after W (any form,
// at most 3), or one before W and at most one after, both of Pauli form
// (the coded path).  The RZ^dagger RY^dagger pairs of a hardware-efficient
// ansatz merge this way (the Z product before, the Y product after).
static void merge_one_qubit_runs(std::vector<EOp>& back, std::vector<int>& src) {
    struct Prod {
        M2 m;        // payload (after-form of its own op)
        M2 upto;     // the run's product through its own op
        int slot, prm;
        bool pre;
        std::vector<double2> dm;
    };
    const size_t n = back.size();
    std::vector<char> dead(n, 0);
    auto admissible = [](const std::vector<Prod>& ps) {
        int npre = 0, npost = 0, post_coded = 0;
        for (const Prod& p : ps) {
            double s;
            if (p.pre) ++npre;
            else {
                ++npost;
                post_coded += pauli_form(p.dm.data(), s) ? 1 : 0;
            }
        }
        if ((int)ps.size() > 3) return false;
        return npre == 0 || (npre == 1 && npost <= 1 && post_coded == npost);
    };
    for (size_t i = 0; i < n; ++i) {
        if (dead[i] || !one_qubit(back[i])) continue;
        const int q = back[i].g.tb[0];
        const uint64_t qb = 1ull << q;
        M2 W = mat1(back[i]);
        std::vector<Prod> prods;
        auto own = [&](const EOp& e, std::vector<Prod>& ps) {
            for (int p = 0; p < e.nparams; ++p)
                ps.push_back({ mat1(e.dm[p]), W, e.pslot[p] >= 0 ? e.pslot[p] : e.aslot,
                               e.pslot[p] >= 0 ? e.pprm[p] : p, e.ppre[p], e.dm[p] });
        };
        own(back[i], prods);
        int absorbed = 0;
        for (size_t k = i + 1; k < n; ++k) {
            if (dead[k] || !(back[k].touched() & qb)) continue;
            const EOp& x = back[k];
            if (!one_qubit(x) || x.g.tb[0] != q) break;
            const M2 X = mat1(x);
            std::vector<Prod> t = prods;
            bool ok = true;
            for (Prod& p : t) {
                if (p.pre || commutes(X, p.m)) continue;
                double s;
                if (commutes(p.upto, p.m) && pauli_form(p.dm.data(), s)) p.pre = true;
                else ok = false;
            }
            if (!ok) break;
            const M2 W2 = mul(X, W);
            const M2 W0 = W;
            W = W2;
            own(x, t);
            if (!admissible(t)) {
                W = W0;
                break;
            }
            prods.swap(t);
            dead[k] = 1;
            ++absorbed;
        }
        if (!absorbed) continue;
        EOp m;
        m.kind = E_DENSE;
        m.g.k = 1;
        m.g.tb[0] = q;
        m.g.diag = false;
        m.g.m.resize(4);
        for (int j = 0; j < 4; ++j) m.g.m[j] = make_double2(W[j].real(), W[j].imag());
        m.nparams = (int)prods.size();
        for (int p = 0; p < m.nparams; ++p) {
            m.dm[p] = prods[p].dm;
            m.pslot[p] = prods[p].slot;
            m.pprm[p] = prods[p].prm;
            m.ppre[p] = prods[p].pre;
        }
        m.aslot = m.nparams ? prods[0].slot : -1;
        back[i] = std::move(m);
        src[i] = -1;
    }
    size_t w = 0;
    for (size_t i = 0; i < n; ++i)
        if (!dead[i]) {
            if (w != i) {
                back[w] = std::move(back[i]);
                src[w] = src[i];
            }
            ++w;
        }
    back.resize(w);
    src.resize(w);
}

// Pauli-coded products (the TOp comment gives the layout): every
// uncontrolled dense 1-register op whose products are one before its
// update and at most one after, or one after, all of Pauli form.
static int code_products(Plan& plan, int pm) {
    for (size_t o = 0; o < plan.tops.size(); ++o) {
        TOp& op = plan.tops[o];
        if (op.kind != 0 || op.cmask != 0) continue;
        op.cval = 0;
        const int np = op.nparam;
        if (np < 1) continue;
        int npre = 0, npost = 0;
        bool coded = np <= 2;
        uint32_t w[2] = { 0, 0 };
        double sc[2] = { 0.0, 0.0 };
        for (int p = 0; p < np && coded; ++p) {
            const TOp& d = plan.tops[o + 1 + p];
            double s = 0.0;
            const int f = pauli_form(d.m, s);
            const int side = d.pre ? 1 : 0;
            (d.pre ? npre : npost) += 1;
            const int64_t row = (int64_t)d.a * pm + d.b;
            coded = f != 0 && row < (1ll << 30) && w[side] == 0;
            w[side] = (uint32_t)(row << 2) | (uint32_t)f;
            sc[side] = s;
        }
        if (!coded) {
            if (npre) return fail(BSIM_EINVAL, "adjoint plan: a product before its op must be of Pauli form");
            continue;
        }
        op.cval = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
        op.m[4] = make_double2(sc[1], sc[0]);
    }
    return BSIM_OK;
}

int adjoint_back_plan(int vq, const std::vector<EOp>& eops, std::vector<EOp>& back, Plan& plan, int& nslots,
                      int& pm, const TileConfig* base, std::vector<int>* src) {
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
    std::vector<int> idx(back.size());
    for (int i = 0; i < (int)back.size(); ++i) idx[i] = L - 1 - i;
    merge_one_qubit_runs(back, idx);
    if (src) src->swap(idx);
    TileConfig cfg = base ? *base : tile_config();
    cfg.tile_bits = 11;
    cfg.reg_bits = 3;
    cfg.merge_dense = false;  // single steps run by index in `back`
    if (int rc = make_plan(vq, back, cfg, plan)) return rc;
    return code_products(plan, pm);
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
