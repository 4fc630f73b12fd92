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

#include "fused.h"
#include "tile_kernel.h"

namespace bsim {

constexpr int AT_TB = 256;   // threads per workgroup (2^(11 - 3) sub-cubes)
constexpr int AT_BITS = 8;   // log2(AT_TB)
constexpr int AT_W = AT_TB / 64;

// DPP move of a double (two 32-bit moves): lanes whose source lies outside
// their row, or whose row is masked off, receive 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROW_MASK, 0xF, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xF, true);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the wavefront, complete in lane 63: an inclusive scan inside each
// row of 16 lanes (row_shr 1, 2, 4, 8), then the row totals carried by
// row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2, 3).  VALU moves only,
// no LDS round trips.
__device__ __forceinline__ double wave_sum63(double v) {
    v += dpp_d<0x111, 0xF>(v);
    v += dpp_d<0x112, 0xF>(v);
    v += dpp_d<0x114, 0xF>(v);
    v += dpp_d<0x118, 0xF>(v);
    v += dpp_d<0x142, 0xA>(v);
    v += dpp_d<0x143, 0xC>(v);
    return v;
}

// Re sum_j conj(l_j) (dU psi)_j over the sub-cube for a dense 1-bit dU on
// register A, restricted to the registers where the op's controls hold.
template <int A>
__device__ __forceinline__ double dot_dense1(const SubCube<3>& sp, const SubCube<3>& sl, const TOp* __restrict__ op,
                                             const TOp* __restrict__ d) {
    const c2 m00 = to_c2(d->m[0]), m10 = to_c2(d->m[1]), m01 = to_c2(d->m[2]), m11 = to_c2(d->m[3]);
    const bool tok = op->cmask == 0 || (sp.gbase & op->tmask) == op->tval;
    const uint32_t jm = op->cmask ? op->jmask : 0xFFu;
    double acc = 0.0;
    sfor<0, 8>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & (1 << A)) == 0) {
            constexpr int k = j | (1 << A);
            const c2 a = getv<j>(sp.v), b = getv<k>(sp.v);
            const c2 x = cfma(m01, b, cmul(m00, a)), y = cfma(m11, b, cmul(m10, a));
            const c2 la = getv<j>(sl.v), lb = getv<k>(sl.v);
            const double t = la.x * x.x + la.y * x.y + lb.x * y.x + lb.y * y.y;
            acc += (tok && ((jm >> j) & 1u)) ? t : 0.0;
        }
    });
    return acc;
}

template <int A, int B>
__device__ __forceinline__ double dot_dense2(const SubCube<3>& sp, const SubCube<3>& sl, const TOp* __restrict__ op,
                                             const TOp* __restrict__ d) {
    const bool tok = op->cmask == 0 || (sp.gbase & op->tmask) == op->tval;
    const uint32_t jm = op->cmask ? op->jmask : 0xFFu;
    double acc = 0.0;
    sfor<0, 8>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & ((1 << A) | (1 << B))) == 0) {
            constexpr int i1 = j | (1 << A), i2 = j | (1 << B), i3 = j | (1 << A) | (1 << B);
            const c2 a[4] = { getv<j>(sp.v), getv<i1>(sp.v), getv<i2>(sp.v), getv<i3>(sp.v) };
            const c2 l[4] = { getv<j>(sl.v), getv<i1>(sl.v), getv<i2>(sl.v), getv<i3>(sl.v) };
            double t = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                c2 mu = cmul(to_c2(d->m[r]), a[0]);
                mu = cfma(to_c2(d->m[r + 4]), a[1], mu);
                mu = cfma(to_c2(d->m[r + 8]), a[2], mu);
                mu = cfma(to_c2(d->m[r + 12]), a[3], mu);
                t += l[r].x * mu.x + l[r].y * mu.y;
            }
            acc += (tok && ((jm >> j) & 1u)) ? t : 0.0;
        }
    });
    return acc;
}

__device__ __forceinline__ double dot_diag(const SubCube<3>& sp, const SubCube<3>& sl, const TOp* __restrict__ op,
                                           const TOp* __restrict__ d) {
    const int nd = op->nd;
    int base_idx = 0, r0 = 0, r1 = 0, r2 = 0;
    for (int i = 0; i < nd; ++i) {
        const uint64_t bit = 1ull << op->dbit[i];
        base_idx |= (sp.gbase & bit) ? (1 << i) : 0;
        r0 |= (sp.g0 == bit) ? (1 << i) : 0;
        r1 |= (sp.g1 == bit) ? (1 << i) : 0;
        r2 |= (sp.g2 == bit) ? (1 << i) : 0;
    }
    const uint64_t cm = op->cmask, cv = op->cval;
    const uint64_t tb = sp.gbase & cm;
    double acc = 0.0;
    sfor<0, 8>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const int idx = base_idx | ((j & 1) ? r0 : 0) | ((j & 2) ? r1 : 0) | ((j & 4) ? r2 : 0);
        const c2 mu = cmul(to_c2(d->m[idx]), getv<j>(sp.v));
        const c2 l = getv<j>(sl.v);
        const uint64_t uj = (((j & 1) ? sp.g0 : 0) | ((j & 2) ? sp.g1 : 0) | ((j & 4) ? sp.g2 : 0)) & cm;
        acc += (cm == 0 || (tb | uj) == cv) ? (l.x * mu.x + l.y * mu.y) : 0.0;
    });
    return acc;
}

__device__ __forceinline__ double dot_rdiag(const SubCube<3>& sp, const SubCube<3>& sl, const TOp* __restrict__ op,
                                            const TOp* __restrict__ d) {
    const bool tok = op->cmask == 0 || (sp.gbase & op->tmask) == op->tval;
    const uint32_t jm = op->cmask ? op->jmask : 0xFFu;
    double acc = 0.0;
    sfor<0, 8>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const c2 mu = cmul(to_c2(d->m[j]), getv<j>(sp.v));
        const c2 l = getv<j>(sl.v);
        acc += (tok && ((jm >> j) & 1u)) ? (l.x * mu.x + l.y * mu.y) : 0.0;
    });
    return acc;
}

__device__ __forceinline__ double adj_dot(const SubCube<3>& sp, const SubCube<3>& sl, const TOp* __restrict__ op,
                                          const TOp* __restrict__ d) {
    switch (op->kind * 16 + op->a * 4 + op->b) {
        case 0: return dot_dense1<0>(sp, sl, op, d);
        case 4: return dot_dense1<1>(sp, sl, op, d);
        case 8: return dot_dense1<2>(sp, sl, op, d);
        case 16 + 1: return dot_dense2<0, 1>(sp, sl, op, d);
        case 16 + 2: return dot_dense2<0, 2>(sp, sl, op, d);
        case 16 + 6: return dot_dense2<1, 2>(sp, sl, op, d);
        case 96: return dot_rdiag(sp, sl, op, d);
        default: return dot_diag(sp, sl, op, d);
    }
}

// Both sub-cubes through one op, one dispatch: the op's index tests (sign
// parities, diagonal indices) are computed once for the two.
__device__ __forceinline__ void sign_both(SubCube<3>& sp, SubCube<3>& sl, const TOp* __restrict__ op) {
    const uint64_t a = sp.gbase;
    uint32_t m = op->jmask;
    uint32_t par = 0;
    for (int k = 0; k < op->npairs; ++k) par += __popcll(a & (a >> op->pd[k]) & op->pm[k]);
    auto flip = [&](uint32_t p, uint32_t pattern) { m = xor_and(m, (uint32_t)__builtin_amdgcn_sbfe(p, 0, 1), pattern); };
    flip(par, 0xFFu);
    flip(__popcll(a & op->sgm[0]), reg_pattern<3>(0));
    flip(__popcll(a & op->sgm[1]), reg_pattern<3>(1));
    flip(__popcll(a & op->sgm[2]), reg_pattern<3>(2));
    auto neg = [&](SubCube<3>& s) {
        sfor<0, 8>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const uint32_t t = m << (31 - j);
            c2 v = getv<j>(s.v);
            const uint64_t bx = __builtin_bit_cast(uint64_t, v.x), by = __builtin_bit_cast(uint64_t, v.y);
            const uint64_t hx = xor_and((uint32_t)(bx >> 32), t, 0x80000000u);
            const uint64_t hy = xor_and((uint32_t)(by >> 32), t, 0x80000000u);
            v.x = __builtin_bit_cast(double, (bx & 0xFFFFFFFFull) | (hx << 32));
            v.y = __builtin_bit_cast(double, (by & 0xFFFFFFFFull) | (hy << 32));
            setv<j>(s.v, v);
        });
    };
    neg(sp);
    neg(sl);
}

__device__ __forceinline__ void diag_both(SubCube<3>& sp, SubCube<3>& sl, const TOp* __restrict__ op) {
    const int nd = op->nd;
    int base_idx = 0, r0 = 0, r1 = 0, r2 = 0;
    for (int i = 0; i < nd; ++i) {
        const uint64_t bit = 1ull << op->dbit[i];
        base_idx |= (sp.gbase & bit) ? (1 << i) : 0;
        r0 |= (sp.g0 == bit) ? (1 << i) : 0;
        r1 |= (sp.g1 == bit) ? (1 << i) : 0;
        r2 |= (sp.g2 == bit) ? (1 << i) : 0;
    }
    const uint64_t cm = op->cmask, cv = op->cval;
    const uint64_t tb = sp.gbase & cm;
    sfor<0, 8>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const int idx = base_idx | ((j & 1) ? r0 : 0) | ((j & 2) ? r1 : 0) | ((j & 4) ? r2 : 0);
        const c2 f = to_c2(op->m[idx]);
        const uint64_t uj = (((j & 1) ? sp.g0 : 0) | ((j & 2) ? sp.g1 : 0) | ((j & 4) ? sp.g2 : 0)) & cm;
        const bool on = cm == 0 || (tb | uj) == cv;
        setv<j>(sp.v, csel(on, cmul(f, getv<j>(sp.v)), getv<j>(sp.v)));
        setv<j>(sl.v, csel(on, cmul(f, getv<j>(sl.v)), getv<j>(sl.v)));
    });
}

__device__ __forceinline__ void run_op_both(SubCube<3>& sp, SubCube<3>& sl, const TOp* __restrict__ op) {
    const OpHead hd = load_head(op);
    switch (hd.kind * 16 + hd.a * 4 + hd.b) {
        case 48: sign_both(sp, sl, op); break;
        case 0: sw_dense1<3, 0>(sp, op, hd); sw_dense1<3, 0>(sl, op, hd); break;
        case 4: sw_dense1<3, 1>(sp, op, hd); sw_dense1<3, 1>(sl, op, hd); break;
        case 8: sw_dense1<3, 2>(sp, op, hd); sw_dense1<3, 2>(sl, op, hd); break;
        case 16 + 1: sw_dense2<3, 0, 1>(sp, op, hd); sw_dense2<3, 0, 1>(sl, op, hd); break;
        case 16 + 2: sw_dense2<3, 0, 2>(sp, op, hd); sw_dense2<3, 0, 2>(sl, op, hd); break;
        case 16 + 6: sw_dense2<3, 1, 2>(sp, op, hd); sw_dense2<3, 1, 2>(sl, op, hd); break;
        default:
            if (hd.kind == 6) {
                sw_rdiag<3>(sp, op, hd);
                sw_rdiag<3>(sl, op, hd);
            } else if (hd.kind == 7) {
                switch (hd.a) {
                    case 0: sw_x<3, 0>(sp, op, hd); sw_x<3, 0>(sl, op, hd); break;
                    case 1: sw_x<3, 1>(sp, op, hd); sw_x<3, 1>(sl, op, hd); break;
                    default: sw_x<3, 2>(sp, op, hd); sw_x<3, 2>(sl, op, hd); break;
                }
            } else {
                diag_both(sp, sl, op);
            }
            break;
    }
}

// One tile of one backward pass (T <= 11 bits, 3 register bits per sweep).
__global__ __launch_bounds__(AT_TB) void k_adj_tile(double2* __restrict__ psi, double2* __restrict__ lam, int vq,
                                                    const Pass* __restrict__ pass, const Sweep* __restrict__ sweeps,
                                                    const TOp* __restrict__ tops, double* __restrict__ partials,
                                                    int pm, int64_t ntiles) {
    extern __shared__ double2 lds[];
    __shared__ uint64_t scat_lo[AT_TB];
    __shared__ uint64_t scat_hi[64];
    const int T = pass->T;
    const int nloc = 1 << T;
    const int stride_t = nloc + (nloc >> 4);
    double2* tp = lds;             // psi tile
    double2* tl = lds + stride_t;  // lambda tile
    const uint64_t lmask = pass->lmask;
    const uint64_t tile_index = blockIdx.x;
    {
        const int t = threadIdx.x;
        uint64_t s = 0;
        for (int b = 0; b < AT_BITS && b < T; ++b)
            if ((t >> b) & 1) s |= 1ull << pass->lbits[b];
        scat_lo[t] = s;
        if (t < 64) {
            uint64_t h = 0;
            for (int b = AT_BITS; b < T; ++b)
                if ((t >> (b - AT_BITS)) & 1) h |= 1ull << pass->lbits[b];
            scat_hi[t] = h;
        }
    }
    uint64_t w = tile_index, base = 0;
    for (int bit = 0, k = 0; bit < vq; ++bit)
        if (!((lmask >> bit) & 1)) {
            base |= ((w >> k) & 1) << bit;
            ++k;
        }
    __syncthreads();
    const uint64_t s_t = base | scat_lo[threadIdx.x];
    const int iters = nloc >> AT_BITS;
    if (iters == 0) {
        if ((int)threadIdx.x < nloc) {
            tp[phys(threadIdx.x)] = psi[s_t];
            tl[phys(threadIdx.x)] = lam[s_t];
        }
    }
    for (int i = 0; i < iters; ++i) {
        const uint64_t g = s_t | scat_hi[i];
        const int l = phys((int)threadIdx.x + (i << AT_BITS));
        tp[l] = psi[g];
        tl[l] = lam[g];
    }
    __syncthreads();
    const int nsub = nloc >> 3;
    const int c = threadIdx.x;
    const bool valid = c < nsub;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int sidx = 0; sidx < pass->nsweeps; ++sidx) {
        const Sweep* sw = sweeps + pass->sweep_begin + sidx;
        const int sb0 = sw->sb[0], sb1 = sw->sb[1], sb2 = sw->sb[2];
        const int ob = sw->op_begin, oe = sw->op_end;
        auto gl = [&](int l) { return scat_lo[(1 << l) & (AT_TB - 1)] | scat_hi[(1 << l) >> AT_BITS]; };
        const uint64_t g0 = gl(sb0), g1 = gl(sb1), g2 = gl(sb2);
        const int o0 = 1 << sb0, o1 = 1 << sb1, o2 = 1 << sb2;
        auto ins = [](int x, int b) { const int lo = x & ((1 << b) - 1); return ((x - lo) << 1) | lo; };
        const int lb = valid ? ins(ins(ins(c, sb0), sb1), sb2) : 0;
        const int pl = phys(lb);
        SubCube<3> sp, sl;
        sp.gbase = sl.gbase = base | scat_lo[lb & (AT_TB - 1)] | scat_hi[lb >> AT_BITS];
        sp.g0 = sl.g0 = g0;
        sp.g1 = sl.g1 = g1;
        sp.g2 = sl.g2 = g2;
        sp.g3 = sl.g3 = 0;
        sfor<0, 8>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const int o = ((j & 1) ? o0 : 0) + ((j & 2) ? o1 : 0) + ((j & 4) ? o2 : 0);
            setv<j>(sp.v, valid ? to_c2(tp[pl + phys(o)]) : c2{ 0.0, 0.0 });
            setv<j>(sl.v, valid ? to_c2(tl[pl + phys(o)]) : c2{ 0.0, 0.0 });
        });
        for (int o = ob; o < oe;) {
            const TOp* op = tops + o;
            const int np = op->nparam;
            run_op_both(sp, sl, op);
            for (int p = 0; p < np; ++p) {
                const TOp* d = op + 1 + p;
                const double v = wave_sum63(adj_dot(sp, sl, op, d));
                if (lane == 63)
                    partials[((size_t)(d->a * pm + d->b) * ntiles + tile_index) * AT_W + wave] = v;
            }
            o += 1 + np;
        }
        if (valid) {
            sfor<0, 8>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const int o = ((j & 1) ? o0 : 0) + ((j & 2) ? o1 : 0) + ((j & 4) ? o2 : 0);
                tp[pl + phys(o)] = to_d2(getv<j>(sp.v));
                tl[pl + phys(o)] = to_d2(getv<j>(sl.v));
            });
        }
        __syncthreads();
    }
    if (iters == 0) {
        if ((int)threadIdx.x < nloc) {
            psi[s_t] = tp[phys(threadIdx.x)];
            lam[s_t] = tl[phys(threadIdx.x)];
        }
    }
    for (int i = 0; i < iters; ++i) {
        const uint64_t g = s_t | scat_hi[i];
        const int l = phys((int)threadIdx.x + (i << AT_BITS));
        psi[g] = tp[l];
        lam[g] = tl[l];
    }
}

// out[slot * 3 + p] = sum over the n partial sums of row slot * pm + p, in a
// fixed order.
__global__ __launch_bounds__(256) void k_adj_tile_finish(const double* __restrict__ partials, int64_t n, int pm,
                                                         double* __restrict__ out) {
    __shared__ double sh[256];
    const double* p = partials + (size_t)blockIdx.x * n;
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[(blockIdx.x / pm) * 3 + blockIdx.x % pm] = sh[0];
}

__global__ void k_adj_zero(double* __restrict__ p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0;
}

int adjoint_back_plan(int vq, const std::vector<EOp>& eops, std::vector<EOp>& back, Plan& plan, int& nslots,
                      int& pm) {
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
    TileConfig cfg = tile_config();
    cfg.tile_bits = 11;
    cfg.reg_bits = 3;
    return make_plan(vq, back, cfg, plan);
}

static std::atomic<bool> g_adj_shmem{ false };

static int adjoint_tile_prepare() {
    if (g_adj_shmem) return BSIM_OK;
    BSIM_HIP(hipFuncSetAttribute((const void*)k_adj_tile, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)(2 * sizeof(double2) * ((1 << 11) + (1 << 7)))));
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
            hipLaunchKernelGGL(k_adj_tile, dim3(grid), dim3(AT_TB), shmem, h->stream, psi, lam, h->vq,
                               (const Pass*)tables + st.pass, (const Sweep*)(tables + off_s),
                               (const TOp*)(tables + off_o), partials, pm, nt);
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
