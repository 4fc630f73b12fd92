// Device side of the fused state-vector adjoint backward sweep: the tile
// body shared by k_adj_tile (one launch per backward pass) and the batch
// kernel (one ticket per tile).  See adj_tile.hip for the method.
#pragma once

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

// dot_dense1 for an uncontrolled op whose payload is real (nd = 1).
template <int A>
__device__ __forceinline__ double dot_dense1_real(const SubCube<3>& sp, const SubCube<3>& sl,
                                                  const TOp* __restrict__ d) {
    const double m00 = d->m[0].x, m10 = d->m[1].x, m01 = d->m[2].x, m11 = d->m[3].x;
    double acc = 0.0;
    sfor<0, 8>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & (1 << A)) == 0) {
            constexpr int k = j | (1 << A);
            const c2 a = getv<j>(sp.v), b = getv<k>(sp.v);
            const c2 la = getv<j>(sl.v), lb = getv<k>(sl.v);
            const double xr = m00 * a.x + m01 * b.x, xi = m00 * a.y + m01 * b.y;
            const double yr = m10 * a.x + m11 * b.x, yi = m10 * a.y + m11 * b.y;
            acc += la.x * xr + la.y * xi + lb.x * yr + lb.y * yi;
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

// dot_rdiag for an uncontrolled register diagonal whose payload is
// imaginary (nd = 2): Re conj(l) (i mu v) = mu (l.y v.x - l.x v.y).
__device__ __forceinline__ double dot_rdiag_imag(const SubCube<3>& sp, const SubCube<3>& sl,
                                                 const TOp* __restrict__ d) {
    double acc = 0.0;
    sfor<0, 8>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const c2 v = getv<j>(sp.v), l = getv<j>(sl.v);
        acc += d->m[j].y * (l.y * v.x - l.x * v.y);
    });
    return acc;
}

// Im <l| P |v> over the sub-cube for a Pauli P on register A (P = 1 X,
// 2 Y, 3 Z): with Im(x* y) = x.x y.y - x.y y.x and Re(x* y) = x.x y.x + x.y y.y
// over each pair (a, b) of v and (la, lb) of l,
//   X: Im(la* b) + Im(lb* a),  Y: Re(lb* a) - Re(la* b),  Z: Im(la* a) - Im(lb* b).
// Re <l| -i s P |v> = s Im <l| P |v>: the product of a Pauli-coded payload.
template <int A, int P>
__device__ __forceinline__ double pauli_im(const SubCube<3>& sp, const SubCube<3>& sl) {
    double acc = 0.0;
    sfor<0, 8>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & (1 << A)) == 0) {
            constexpr int k = j | (1 << A);
            const c2 a = getv<j>(sp.v), b = getv<k>(sp.v);
            const c2 la = getv<j>(sl.v), lb = getv<k>(sl.v);
            if constexpr (P == 1) acc += (la.x * b.y - la.y * b.x) + (lb.x * a.y - lb.y * a.x);
            else if constexpr (P == 2) acc += (lb.x * a.x + lb.y * a.y) - (la.x * b.x + la.y * b.y);
            else acc += (la.x * a.y - la.y * a.x) - (lb.x * b.y - lb.y * b.x);
        }
    });
    return acc;
}

template <int A>
__device__ __forceinline__ double pauli_im_a(const SubCube<3>& sp, const SubCube<3>& sl, uint32_t p) {
    switch (p) {
        case 1: return pauli_im<A, 1>(sp, sl);
        case 2: return pauli_im<A, 2>(sp, sl);
        default: return pauli_im<A, 3>(sp, sl);
    }
}

__device__ __forceinline__ double pauli_im_any(const SubCube<3>& sp, const SubCube<3>& sl, int a, uint32_t p) {
    switch (a) {
        case 0: return pauli_im_a<0>(sp, sl, p);
        case 1: return pauli_im_a<1>(sp, sl, p);
        default: return pauli_im_a<2>(sp, sl, p);
    }
}

__device__ __forceinline__ double adj_dot(const SubCube<3>& sp, const SubCube<3>& sl, const TOp* __restrict__ op,
                                          const TOp* __restrict__ d) {
    switch (op->kind * 16 + op->a * 4 + op->b) {
        case 0: return d->nd == 1 ? dot_dense1_real<0>(sp, sl, d) : dot_dense1<0>(sp, sl, op, d);
        case 4: return d->nd == 1 ? dot_dense1_real<1>(sp, sl, d) : dot_dense1<1>(sp, sl, op, d);
        case 8: return d->nd == 1 ? dot_dense1_real<2>(sp, sl, d) : dot_dense1<2>(sp, sl, op, d);
        case 16 + 1: return dot_dense2<0, 1>(sp, sl, op, d);
        case 16 + 2: return dot_dense2<0, 2>(sp, sl, op, d);
        case 16 + 6: return dot_dense2<1, 2>(sp, sl, op, d);
        case 96: return d->nd == 2 ? dot_rdiag_imag(sp, sl, d) : dot_rdiag(sp, sl, op, d);
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
        case 0:
            if (hd.nd == 1) {
                sw_dense1_real<3, 0>(sp, hd);
                sw_dense1_real<3, 0>(sl, hd);
            } else {
                sw_dense1<3, 0>(sp, op, hd);
                sw_dense1<3, 0>(sl, op, hd);
            }
            break;
        case 4:
            if (hd.nd == 1) {
                sw_dense1_real<3, 1>(sp, hd);
                sw_dense1_real<3, 1>(sl, hd);
            } else {
                sw_dense1<3, 1>(sp, op, hd);
                sw_dense1<3, 1>(sl, op, hd);
            }
            break;
        case 8:
            if (hd.nd == 1) {
                sw_dense1_real<3, 2>(sp, hd);
                sw_dense1_real<3, 2>(sl, hd);
            } else {
                sw_dense1<3, 2>(sp, op, hd);
                sw_dense1<3, 2>(sl, op, hd);
            }
            break;
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

// One sweep of a backward pass over the tile in LDS (tp: psi, tl: lambda,
// 2^T amplitudes each, T <= 11): every thread carries one sub-cube of each
// through the sweep's ops and writes its wave's product sums to `partials`
// under `tile_index`.  Ends with a barrier.
__device__ __forceinline__ void adj_sweep(const Pass* __restrict__ pass, const Sweep* __restrict__ sweeps,
                                          const TOp* __restrict__ tops, int sidx, uint64_t base, int nloc,
                                          double2* __restrict__ tp, double2* __restrict__ tl,
                                          const uint64_t* __restrict__ scat_lo, const uint64_t* __restrict__ scat_hi,
                                          double* __restrict__ partials, int pm, int64_t ntiles,
                                          uint64_t tile_index) {
    const int nsub = nloc >> 3;
    const int c = threadIdx.x;
    const bool valid = c < nsub;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
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
    auto put = [&](uint32_t row, double v) {
        if (lane == 63) partials[((size_t)row * ntiles + tile_index) * AT_W + wave] = v;
    };
    for (int o = ob; o < oe;) {
        const TOp* op = tops + o;
        const int np = op->nparam;
        const uint64_t pc = op->kind == 0 && op->cmask == 0 ? op->cval : 0;
        if (pc) {
            // Pauli-coded products, read from the op's head (the payload
            // TOps that follow are skipped): one before the update, one after
            const int a = op->a;
            const uint32_t wpre = (uint32_t)(pc >> 32), wpost = (uint32_t)pc;
            const double2 sc = op->m[4];
            if (wpre) put(wpre >> 2, sc.x * wave_sum63(pauli_im_any(sp, sl, a, wpre & 3u)));
            run_op_both(sp, sl, op);
            if (wpost) put(wpost >> 2, sc.y * wave_sum63(pauli_im_any(sp, sl, a, wpost & 3u)));
        } else {
            run_op_both(sp, sl, op);
            for (int p = 0; p < np; ++p) {
                const TOp* d = op + 1 + p;
                put((uint32_t)(d->a * pm + d->b), wave_sum63(adj_dot(sp, sl, op, d)));
            }
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

// The scatter tables of a pass: scat_lo[t] the vector bits of local bits
// 0..AT_BITS-1 of t, scat_hi[h] those of local bits AT_BITS.. of h.
__device__ __forceinline__ void adj_scatter(const Pass* __restrict__ pass, int T, uint64_t* __restrict__ scat_lo,
                                            uint64_t* __restrict__ scat_hi) {
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

// Global base of tile w: w deposited into the non-tile bits.
__device__ __forceinline__ uint64_t adj_tile_base(uint64_t lmask, int vq, uint64_t w) {
    uint64_t base = 0;
    for (int bit = 0, k = 0; bit < vq; ++bit)
        if (!((lmask >> bit) & 1)) {
            base |= ((w >> k) & 1) << bit;
            ++k;
        }
    return base;
}

// One tile of one backward pass (T <= 11 bits, 3 register bits per sweep),
// AT_TB threads.  LDS: `lds` holds the two tiles (2 (2^T + 2^T/16)
// amplitudes), scat_lo[AT_TB] and scat_hi[64] the scatter tables.
__device__ __forceinline__ void adj_tile_body(double2* __restrict__ psi, double2* __restrict__ lam, int vq,
                                              const Pass* __restrict__ pass, const Sweep* __restrict__ sweeps,
                                              const TOp* __restrict__ tops, double* __restrict__ partials, int pm,
                                              int64_t ntiles, uint64_t tile_index, double2* __restrict__ lds,
                                              uint64_t* __restrict__ scat_lo, uint64_t* __restrict__ scat_hi) {
    const int T = pass->T;
    const int nloc = 1 << T;
    const int stride_t = nloc + (nloc >> 4);
    double2* tp = lds;             // psi tile
    double2* tl = lds + stride_t;  // lambda tile
    adj_scatter(pass, T, scat_lo, scat_hi);
    const uint64_t base = adj_tile_base(pass->lmask, vq, tile_index);
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
    for (int sidx = 0; sidx < pass->nsweeps; ++sidx)
        adj_sweep(pass, sweeps, tops, sidx, base, nloc, tp, tl, scat_lo, scat_hi, partials, pm, ntiles, tile_index);
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

// The streaming backward pass (T = 11): a persistent workgroup per slot
// takes tiles by ticket -- its first tile is its index, then each ticket v
// = atomicAdd(ctr, 1) names tile gridDim.x + v, and the workgroup drawing
// the pass's last ticket (v = ntiles - 1) resets the counter -- and loads
// the next tile's psi and lambda rows into registers while the current
// tile runs its last sweep.  The same sweeps, products and partial slots
// as adj_tile_body, tile by tile.  LDS as adj_tile_body plus s_next.
__device__ __forceinline__ void adj_stream_body(double2* __restrict__ psi, double2* __restrict__ lam, int vq,
                                                const Pass* __restrict__ pass, const Sweep* __restrict__ sweeps,
                                                const TOp* __restrict__ tops, double* __restrict__ partials, int pm,
                                                int64_t ntiles, uint64_t pass_tiles,
                                                unsigned long long* __restrict__ ctr, double2* __restrict__ lds,
                                                uint64_t* __restrict__ scat_lo, uint64_t* __restrict__ scat_hi,
                                                uint64_t* s_next) {
    constexpr int T = 11, NR = 1 << (T - AT_BITS);  // 8 rows of AT_TB amplitudes
    constexpr int nloc = 1 << T, stride_t = nloc + (nloc >> 4);
    double2* tp = lds;
    double2* tl = lds + stride_t;
    adj_scatter(pass, T, scat_lo, scat_hi);
    const uint64_t lmask = pass->lmask;
    __syncthreads();
    const int nsw = pass->nsweeps;
    dvec<2 * NR> pp, pq;  // the next tile's psi and lambda rows (row i = local tid + AT_TB i)
    auto load = [&](uint64_t b) {
        const uint64_t s_t = b | scat_lo[threadIdx.x];
        sfor<0, NR>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const uint64_t g = s_t | scat_hi[i];
            setv<i>(pp, to_c2(psi[g]));
            setv<i>(pq, to_c2(lam[g]));
        });
    };
    auto to_lds = [&]() {
        sfor<0, NR>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int l = phys((int)threadIdx.x + (i << AT_BITS));
            tp[l] = to_d2(getv<i>(pp));
            tl[l] = to_d2(getv<i>(pq));
        });
        __syncthreads();
    };
    uint64_t t = blockIdx.x;  // < pass_tiles (grid <= pass_tiles)
    uint64_t base = adj_tile_base(lmask, vq, t);
    load(base);
    to_lds();
    for (;;) {
        if (threadIdx.x == 0) {
            const unsigned long long v = atomicAdd(ctr, 1ull);
            if (v == pass_tiles - 1) atomicExch(ctr, 0ull);  // the pass's last ticket
            *s_next = gridDim.x + v;
        }
        __syncthreads();
        const uint64_t nt = *s_next;  // rewritten only after the store barrier below
        const uint64_t nbase = nt < pass_tiles ? adj_tile_base(lmask, vq, nt) : base;
        for (int sidx = 0; sidx + 1 < nsw; ++sidx)
            adj_sweep(pass, sweeps, tops, sidx, base, nloc, tp, tl, scat_lo, scat_hi, partials, pm, ntiles, t);
        load(nbase);  // in flight during the last sweep and the stores
        if (nsw > 0)
            adj_sweep(pass, sweeps, tops, nsw - 1, base, nloc, tp, tl, scat_lo, scat_hi, partials, pm, ntiles, t);
        const uint64_t s_t = base | scat_lo[threadIdx.x];
        sfor<0, NR>([&](auto I) {
            constexpr int i = decltype(I)::value;
            const int l = phys((int)threadIdx.x + (i << AT_BITS));
            const uint64_t g = s_t | scat_hi[i];
            psi[g] = tp[l];
            lam[g] = tl[l];
        });
        __syncthreads();  // the next tile rewrites LDS
        if (nt >= pass_tiles) break;
        t = nt;
        base = nbase;
        to_lds();
    }
}

// Fixed-order sum of one row of n partial sums by a 256-thread workgroup:
// thread t sums entries t, t + 256, ..., then a tree in `sh` (256 doubles).
// The result is valid in thread 0.
__device__ __forceinline__ double adj_row_sum(const double* __restrict__ p, int64_t n, double* __restrict__ sh) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) s += p[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

}  // namespace bsim
