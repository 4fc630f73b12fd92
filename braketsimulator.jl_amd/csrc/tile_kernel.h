// The fused tile kernel (one HBM pass of a planned instruction block).
//
// One workgroup, one tile of 2^T amplitudes in LDS (padded by one element
// per 16 so that strided sub-cube reads spread over the banks).  A pass
// works with R register bits per sweep (R = 3 or 4, the planner's choice):
// the workgroup has TB(R) = 2^(12 - R) threads, so a 12-bit tile gives each
// thread one sub-cube of 2^R amplitudes per sweep (R = 3: 512 threads with
// 8 amplitudes, four waves per SIMD; R = 4: 256 threads with 16, two).
//   1. load: local index j = tid + TB i; the global index is the block's
//      base plus the scatter of j over the tile's vector bits (two small LDS
//      tables: the low log2(TB) local bits and the rest).  Local bits
//      0..low_bits-1 are vector bits 0..low_bits-1, so each wavefront load
//      covers contiguous runs of 2^low_bits amplitudes.
//   2. sweeps: a sweep names R local bits; each thread gathers its sub-cube
//      into registers, applies the sweep's gates there (dense 1- and 2-qubit
//      on the register bits, diagonal/controls on any vector bit through the
//      global index) and writes it back.
//   3. store: the load pattern in reverse.
// Every register index is a compile-time constant (sfor below), so the
// sub-cube never leaves the VGPR file.
#pragma once

#include <type_traits>

#include "fused.h"

namespace bsim {

template <int R>
constexpr int tile_threads() { return 4096 >> R; }
template <int R>
constexpr int tile_thread_bits() { return 12 - R; }

template <int I, int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        sfor<I + 1, N>(f);
    }
}

__device__ __forceinline__ int phys(int j) { return j + (j >> 4); }

struct c2 {
    double x, y;
};
// The sub-cube is one vector value of 2^(R+1) doubles, so it stays in VGPRs:
// every element access below uses a compile-time index.
template <int N>
using dvec = double __attribute__((ext_vector_type(N)));
template <int J, int N>
__device__ __forceinline__ c2 getv(const dvec<N>& v) {
    return { v[2 * J], v[2 * J + 1] };
}
template <int J, int N>
__device__ __forceinline__ void setv(dvec<N>& v, c2 a) {
    v[2 * J] = a.x;
    v[2 * J + 1] = a.y;
}
__device__ __forceinline__ c2 to_c2(double2 a) { return { a.x, a.y }; }
__device__ __forceinline__ double2 to_d2(c2 a) { return make_double2(a.x, a.y); }
__device__ __forceinline__ c2 cmul(c2 a, c2 b) { return { a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x }; }
__device__ __forceinline__ c2 cfma(c2 a, c2 b, c2 c) {
    return { fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)) };
}
// Component-wise select: `on ? a : b` on the struct itself becomes a select
// between two stack addresses, i.e. scratch memory.
__device__ __forceinline__ c2 csel(bool on, c2 a, c2 b) { return { on ? a.x : b.x, on ? a.y : b.y }; }

template <int R>
struct SubCube {
    static constexpr int NR = 1 << R;
    dvec<2 * NR> v;
    uint64_t gbase;            // global index of register 0
    uint64_t g0, g1, g2, g3;   // global bit of register bits 0..R-1 (g3 = 0 when R = 3)
};

// registers of a sub-cube whose index has bit i set (the sign op's patterns)
template <int R>
constexpr uint32_t reg_pattern(int i) {
    uint32_t p = 0;
    for (int j = 0; j < (1 << R); ++j)
        if ((j >> i) & 1) p |= 1u << j;
    return p;
}

// The fields every dense op needs, read with the op's kind before the
// dispatch so that a 1-bit dense op costs one scalar-load round trip.
struct OpHead {
    int32_t kind, a, b, nd;
    uint64_t cmask;
    c2 m0, m1, m2, m3;
};
__device__ __forceinline__ OpHead load_head(const TOp* __restrict__ op) {
    OpHead h;
    h.kind = op->kind;
    h.a = op->a;
    h.b = op->b;
    h.nd = op->nd;
    h.cmask = op->cmask;
    h.m0 = to_c2(op->m[0]);
    h.m1 = to_c2(op->m[1]);
    h.m2 = to_c2(op->m[2]);
    h.m3 = to_c2(op->m[3]);
    return h;
}

// One 2x2 complex matrix on the pair (a, b) in place: x = m00 a + m01 b,
// y = m10 a + m11 b with P = m01 b and Q = m10 a formed first, so that each
// output's last instruction is the last read of the input it replaces.
__device__ __forceinline__ void pair_inplace(double& ar, double& ai, double& br, double& bi, c2 m00, c2 m10, c2 m01,
                                             c2 m11) {
    double pr, pi, qr, qi, t, t2;
    asm("v_mul_f64 %[pr], %[br], %[m01r]\n\t"
        "v_mul_f64 %[qr], %[ar], %[m10r]\n\t"
        "v_mul_f64 %[pi], %[bi], %[m01r]\n\t"
        "v_mul_f64 %[qi], %[ai], %[m10r]\n\t"
        "v_fma_f64 %[pr], -%[bi], %[m01i], %[pr]\n\t"
        "v_fma_f64 %[qr], -%[ai], %[m10i], %[qr]\n\t"
        "v_fma_f64 %[pi], %[br], %[m01i], %[pi]\n\t"
        "v_fma_f64 %[qi], %[ar], %[m10i], %[qi]\n\t"
        "v_fma_f64 %[t], -%[ai], %[m00i], %[pr]\n\t"
        "v_fma_f64 %[t2], -%[bi], %[m11i], %[qr]\n\t"
        "v_fma_f64 %[ai], %[ai], %[m00r], %[pi]\n\t"
        "v_fma_f64 %[bi], %[bi], %[m11r], %[qi]\n\t"
        "v_fma_f64 %[ai], %[ar], %[m00i], %[ai]\n\t"
        "v_fma_f64 %[bi], %[br], %[m11i], %[bi]\n\t"
        "v_fma_f64 %[ar], %[ar], %[m00r], %[t]\n\t"
        "v_fma_f64 %[br], %[br], %[m11r], %[t2]"
        : [ar] "+v"(ar), [ai] "+v"(ai), [br] "+v"(br), [bi] "+v"(bi), [pr] "=&v"(pr), [pi] "=&v"(pi),
          [qr] "=&v"(qr), [qi] "=&v"(qi), [t] "=&v"(t), [t2] "=&v"(t2)
        : [m00r] "s"(m00.x), [m00i] "s"(m00.y), [m10r] "s"(m10.x), [m10i] "s"(m10.y), [m01r] "s"(m01.x),
          [m01i] "s"(m01.y), [m11r] "s"(m11.x), [m11i] "s"(m11.y));
}

// The same with a real matrix (m00, m10, m01, m11 real): 8 instructions.
__device__ __forceinline__ void pair_inplace_real(double& ar, double& ai, double& br, double& bi, double m00,
                                                  double m10, double m01, double m11) {
    double pr, pi, qr, qi;
    asm("v_mul_f64 %[pr], %[br], %[m01]\n\t"
        "v_mul_f64 %[pi], %[bi], %[m01]\n\t"
        "v_mul_f64 %[qr], %[ar], %[m10]\n\t"
        "v_mul_f64 %[qi], %[ai], %[m10]\n\t"
        "v_fma_f64 %[ar], %[ar], %[m00], %[pr]\n\t"
        "v_fma_f64 %[ai], %[ai], %[m00], %[pi]\n\t"
        "v_fma_f64 %[br], %[br], %[m11], %[qr]\n\t"
        "v_fma_f64 %[bi], %[bi], %[m11], %[qi]"
        : [ar] "+v"(ar), [ai] "+v"(ai), [br] "+v"(br), [bi] "+v"(bi), [pr] "=&v"(pr), [pi] "=&v"(pi),
          [qr] "=&v"(qr), [qi] "=&v"(qi)
        : [m00] "s"(m00), [m10] "s"(m10), [m01] "s"(m01), [m11] "s"(m11));
}

// An uncontrolled dense 1-register op with a real matrix (the planner marks
// it with nd = 1).
template <int R, int A>
__device__ __forceinline__ void sw_dense1_real(SubCube<R>& s, const OpHead& hd) {
    const double m00 = hd.m0.x, m10 = hd.m1.x, m01 = hd.m2.x, m11 = hd.m3.x;
    sfor<0, (1 << R)>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & (1 << A)) == 0) {
            constexpr int k = j | (1 << A);
            double ar = s.v[2 * j], ai = s.v[2 * j + 1], br = s.v[2 * k], bi = s.v[2 * k + 1];
            pair_inplace_real(ar, ai, br, bi, m00, m10, m01, m11);
            s.v[2 * j] = ar;
            s.v[2 * j + 1] = ai;
            s.v[2 * k] = br;
            s.v[2 * k + 1] = bi;
        }
    });
}

// Dense ops come in two instantiations: without controls (the common case)
// and with the split control test; the choice is made once per op.
template <int R, int A, bool CTL>
__device__ __forceinline__ void sw_dense1_t(SubCube<R>& s, const TOp* __restrict__ op, const OpHead& hd) {
    const c2 m00 = hd.m0, m10 = hd.m1, m01 = hd.m2, m11 = hd.m3;
    bool tok = true;
    uint32_t jm = 0xFFFFu;
    if constexpr (CTL) {
        tok = (s.gbase & op->tmask) == op->tval;  // this thread's part of the control test
        jm = op->jmask;                           // the registers' part
    }
    if constexpr (!CTL) {
        sfor<0, (1 << R)>([&](auto J) {
            constexpr int j = decltype(J)::value;
            if constexpr ((j & (1 << A)) == 0) {
                constexpr int k = j | (1 << A);
                double ar = s.v[2 * j], ai = s.v[2 * j + 1], br = s.v[2 * k], bi = s.v[2 * k + 1];
                pair_inplace(ar, ai, br, bi, m00, m10, m01, m11);
                s.v[2 * j] = ar;
                s.v[2 * j + 1] = ai;
                s.v[2 * k] = br;
                s.v[2 * k + 1] = bi;
            }
        });
        return;
    }
    sfor<0, (1 << R)>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & (1 << A)) == 0) {
            constexpr int k = j | (1 << A);
            const c2 a = getv<j>(s.v), b = getv<k>(s.v);
            c2 x = cfma(m01, b, cmul(m00, a));
            c2 y = cfma(m11, b, cmul(m10, a));
            if constexpr (CTL) {
                const bool on = tok && ((jm >> j) & 1u);
                x = csel(on, x, a);
                y = csel(on, y, b);
            }
            setv<j>(s.v, x);
            setv<k>(s.v, y);
        }
    });
}

template <int R, int A>
__device__ __forceinline__ void sw_dense1(SubCube<R>& s, const TOp* __restrict__ op, const OpHead& hd) {
    if (hd.cmask) sw_dense1_t<R, A, true>(s, op, hd);
    else sw_dense1_t<R, A, false>(s, op, hd);
}

// acc = sum over c != i of M[i][c] a_c (12 operands at most: asm operand
// lists stay short); M column-major m[r + 4c].
__device__ __forceinline__ void offdiag_row(double& tr, double& ti, double a0r, double a0i, double a1r, double a1i,
                                            double a2r, double a2i, c2 m0, c2 m1, c2 m2) {
    asm("v_mul_f64 %[tr], %[a0r], %[m0r]\n\t"
        "v_mul_f64 %[ti], %[a0i], %[m0r]\n\t"
        "v_fma_f64 %[tr], -%[a0i], %[m0i], %[tr]\n\t"
        "v_fma_f64 %[ti], %[a0r], %[m0i], %[ti]\n\t"
        "v_fma_f64 %[tr], %[a1r], %[m1r], %[tr]\n\t"
        "v_fma_f64 %[ti], %[a1i], %[m1r], %[ti]\n\t"
        "v_fma_f64 %[tr], -%[a1i], %[m1i], %[tr]\n\t"
        "v_fma_f64 %[ti], %[a1r], %[m1i], %[ti]\n\t"
        "v_fma_f64 %[tr], %[a2r], %[m2r], %[tr]\n\t"
        "v_fma_f64 %[ti], %[a2i], %[m2r], %[ti]\n\t"
        "v_fma_f64 %[tr], -%[a2i], %[m2i], %[tr]\n\t"
        "v_fma_f64 %[ti], %[a2r], %[m2i], %[ti]"
        : [tr] "=&v"(tr), [ti] "=&v"(ti)
        : [a0r] "v"(a0r), [a0i] "v"(a0i), [a1r] "v"(a1r), [a1i] "v"(a1i), [a2r] "v"(a2r), [a2i] "v"(a2i),
          [m0r] "s"(m0.x), [m0i] "s"(m0.y), [m1r] "s"(m1.x), [m1i] "s"(m1.y), [m2r] "s"(m2.x), [m2i] "s"(m2.y));
}

// a <- m a + t in place (one scratch register pair).
__device__ __forceinline__ void diag_inplace(double& ar, double& ai, double tr, double ti, c2 m) {
    double u;
    asm("v_fma_f64 %[u], -%[ai], %[mi], %[tr]\n\t"
        "v_fma_f64 %[ai], %[ai], %[mr], %[ti]\n\t"
        "v_fma_f64 %[ai], %[ar], %[mi], %[ai]\n\t"
        "v_fma_f64 %[ar], %[ar], %[mr], %[u]"
        : [ar] "+v"(ar), [ai] "+v"(ai), [u] "=&v"(u)
        : [tr] "v"(tr), [ti] "v"(ti), [mr] "s"(m.x), [mi] "s"(m.y));
}

template <int R, int A, int B, bool CTL>
__device__ __forceinline__ void sw_dense2_t(SubCube<R>& s, const TOp* __restrict__ op, const OpHead& hd) {
    if constexpr (!CTL) {
        // off-diagonal parts of every output first (all inputs intact), then
        // each output in place of its input: no copies of the group
        const c2 M[16] = { hd.m0, hd.m1, hd.m2, hd.m3, to_c2(op->m[4]), to_c2(op->m[5]), to_c2(op->m[6]),
                           to_c2(op->m[7]), to_c2(op->m[8]), to_c2(op->m[9]), to_c2(op->m[10]), to_c2(op->m[11]),
                           to_c2(op->m[12]), to_c2(op->m[13]), to_c2(op->m[14]), to_c2(op->m[15]) };
        sfor<0, (1 << R)>([&](auto J) {
            constexpr int j = decltype(J)::value;
            if constexpr ((j & ((1 << A) | (1 << B))) == 0) {
                constexpr int ix[4] = { j, j | (1 << A), j | (1 << B), j | (1 << A) | (1 << B) };
                double ar[4], ai[4], tr[4], ti[4];
                for (int q = 0; q < 4; ++q) {
                    ar[q] = s.v[2 * ix[q]];
                    ai[q] = s.v[2 * ix[q] + 1];
                }
                offdiag_row(tr[0], ti[0], ar[1], ai[1], ar[2], ai[2], ar[3], ai[3], M[4], M[8], M[12]);
                offdiag_row(tr[1], ti[1], ar[0], ai[0], ar[2], ai[2], ar[3], ai[3], M[1], M[9], M[13]);
                offdiag_row(tr[2], ti[2], ar[0], ai[0], ar[1], ai[1], ar[3], ai[3], M[2], M[6], M[14]);
                offdiag_row(tr[3], ti[3], ar[0], ai[0], ar[1], ai[1], ar[2], ai[2], M[3], M[7], M[11]);
                diag_inplace(ar[0], ai[0], tr[0], ti[0], M[0]);
                diag_inplace(ar[1], ai[1], tr[1], ti[1], M[5]);
                diag_inplace(ar[2], ai[2], tr[2], ti[2], M[10]);
                diag_inplace(ar[3], ai[3], tr[3], ti[3], M[15]);
                for (int q = 0; q < 4; ++q) {
                    s.v[2 * ix[q]] = ar[q];
                    s.v[2 * ix[q] + 1] = ai[q];
                }
            }
        });
        return;
    }
    bool tok = true;
    uint32_t jm = 0xFFFFu;
    if constexpr (CTL) {
        tok = (s.gbase & op->tmask) == op->tval;
        jm = op->jmask;
    }
    sfor<0, (1 << R)>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & ((1 << A) | (1 << B))) == 0) {
            constexpr int i1 = j | (1 << A), i2 = j | (1 << B), i3 = j | (1 << A) | (1 << B);
            const c2 a0 = getv<j>(s.v), a1 = getv<i1>(s.v), a2 = getv<i2>(s.v), a3 = getv<i3>(s.v);
            const c2 col0[4] = { hd.m0, hd.m1, hd.m2, hd.m3 };
            auto row = [&](int r) {
                c2 acc = cmul(col0[r], a0);
                acc = cfma(to_c2(op->m[r + 4]), a1, acc);
                acc = cfma(to_c2(op->m[r + 8]), a2, acc);
                return cfma(to_c2(op->m[r + 12]), a3, acc);
            };
            c2 r0 = row(0), r1 = row(1), r2 = row(2), r3 = row(3);
            if constexpr (CTL) {
                const bool on = tok && ((jm >> j) & 1u);
                r0 = csel(on, r0, a0);
                r1 = csel(on, r1, a1);
                r2 = csel(on, r2, a2);
                r3 = csel(on, r3, a3);
            }
            setv<j>(s.v, r0);
            setv<i1>(s.v, r1);
            setv<i2>(s.v, r2);
            setv<i3>(s.v, r3);
        }
    });
}

template <int R, int A, int B>
__device__ __forceinline__ void sw_dense2(SubCube<R>& s, const TOp* __restrict__ op, const OpHead& hd) {
    if (hd.cmask) sw_dense2_t<R, A, B, true>(s, op, hd);
    else sw_dense2_t<R, A, B, false>(s, op, hd);
}

// Diagonal index of register j = (index bits read from the thread's base)
// | (index bits carried by the register bits of j).  The second part is the
// same for every thread, so it is built from uniform values.
template <int R>
__device__ __forceinline__ void sw_diag(SubCube<R>& s, const TOp* __restrict__ op) {
    const int nd = op->nd;
    int base_idx = 0, r0 = 0, r1 = 0, r2 = 0, r3 = 0;
    for (int i = 0; i < nd; ++i) {
        const uint64_t bit = 1ull << op->dbit[i];
        base_idx |= (s.gbase & bit) ? (1 << i) : 0;
        r0 |= (s.g0 == bit) ? (1 << i) : 0;
        r1 |= (s.g1 == bit) ? (1 << i) : 0;
        r2 |= (s.g2 == bit) ? (1 << i) : 0;
        r3 |= (s.g3 == bit) ? (1 << i) : 0;
    }
    const uint64_t cm = op->cmask, cv = op->cval;
    const uint64_t tb = s.gbase & cm;
    sfor<0, (1 << R)>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const int idx = base_idx | ((j & 1) ? r0 : 0) | ((j & 2) ? r1 : 0) | ((j & 4) ? r2 : 0) | ((j & 8) ? r3 : 0);
        const c2 f = to_c2(op->m[idx]);
        const c2 x = cmul(f, getv<j>(s.v));
        const uint64_t uj = (((j & 1) ? s.g0 : 0) | ((j & 2) ? s.g1 : 0) | ((j & 4) ? s.g2 : 0) | ((j & 8) ? s.g3 : 0)) & cm;
        setv<j>(s.v, csel(cm == 0 || (tb | uj) == cv, x, getv<j>(s.v)));
    });
}

// X on register A (kind 7, 3-register sweeps): the register pairs swap
// where the controls hold.
template <int R, int A, bool CTL>
__device__ __forceinline__ void sw_x_t(SubCube<R>& s, const TOp* __restrict__ op) {
    bool tok = true;
    uint32_t jm = 0xFFFFu;
    if constexpr (CTL) {
        tok = (s.gbase & op->tmask) == op->tval;
        jm = op->jmask;
    }
    sfor<0, (1 << R)>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & (1 << A)) == 0) {
            constexpr int k = j | (1 << A);
            const c2 a = getv<j>(s.v), b = getv<k>(s.v);
            if constexpr (CTL) {
                const bool on = tok && ((jm >> j) & 1u);
                setv<j>(s.v, csel(on, b, a));
                setv<k>(s.v, csel(on, a, b));
            } else {
                setv<j>(s.v, b);
                setv<k>(s.v, a);
            }
        }
    });
}

template <int R, int A>
__device__ __forceinline__ void sw_x(SubCube<R>& s, const TOp* __restrict__ op, const OpHead& hd) {
    if (hd.cmask) sw_x_t<R, A, true>(s, op);
    else sw_x_t<R, A, false>(s, op);
}

// Diagonal on register bits only (kind 6): register j is scaled by m[j], the
// same entries for every thread (scalar loads).
template <int R>
__device__ __forceinline__ void sw_rdiag(SubCube<R>& s, const TOp* __restrict__ op, const OpHead& hd) {
    bool tok = true;
    uint32_t jm = 0xFFFFu;
    if (hd.cmask) {
        tok = (s.gbase & op->tmask) == op->tval;
        jm = op->jmask;
    }
    sfor<0, (1 << R)>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const c2 a = getv<j>(s.v);
        const c2 x = cmul(to_c2(op->m[j]), a);
        setv<j>(s.v, csel(tok && ((jm >> j) & 1u), x, a));
    });
}

__device__ __forceinline__ uint32_t xor_and(uint32_t a, uint32_t b, uint32_t c) {
    // a ^ (b & c) in one v_bitop3 (truth table over a = 0xF0, b = 0xCC, c = 0xAA)
    return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)(0xF0 ^ (0xCC & 0xAA)));
}

// +-1 diagonal (CZ layers): negate where
// P(i) = sum_k parity(i & (i >> pd[k]) & pm[k]) is 1.  P is a quadratic form
// over GF(2): for i = gbase ^ r_j,
// P(i) = P(gbase) ^ P(r_j) ^ sum over the register bits of j of
// parity(gbase & sgm[bit]); the planner stores P(r_j) for every register j
// in jmask.
template <int R>
__device__ __forceinline__ void sw_sign(SubCube<R>& s, const OpHead& hd) {
    // the parameters come with the op head (the planner's copies): pair
    // distances in nd, jmask and the pair count in cmask, pair masks and
    // register-bit masks in m0..m3; unused pair masks are zero
    uint32_t m = (uint32_t)hd.cmask;
    const int np = (int)(hd.cmask >> 32);
    const int pd = hd.nd;
    const uint64_t pm0 = __builtin_bit_cast(uint64_t, hd.m0.x), pm1 = __builtin_bit_cast(uint64_t, hd.m0.y);
    const uint64_t pm2 = __builtin_bit_cast(uint64_t, hd.m1.x), pm3 = __builtin_bit_cast(uint64_t, hd.m1.y);
    const uint64_t sg0 = __builtin_bit_cast(uint64_t, hd.m2.x), sg1 = __builtin_bit_cast(uint64_t, hd.m2.y);
    const uint64_t sg2 = __builtin_bit_cast(uint64_t, hd.m3.x), sg3 = __builtin_bit_cast(uint64_t, hd.m3.y);
    const uint64_t a = s.gbase;
    uint32_t par = __popcll(a & (a >> (pd & 0xFF)) & pm0);
    if (np > 1) {
        par += __popcll(a & (a >> ((pd >> 8) & 0xFF)) & pm1);
        if (np > 2) {
            par += __popcll(a & (a >> ((pd >> 16) & 0xFF)) & pm2);
            if (np > 3) par += __popcll(a & (a >> ((pd >> 24) & 0xFF)) & pm3);
        }
    }
    // all-ones where the parity bit is set (sign-extended bit 0), masked into m
    auto flip = [&](uint32_t p, uint32_t pattern) { m = xor_and(m, (uint32_t)__builtin_amdgcn_sbfe(p, 0, 1), pattern); };
    flip(par, (1u << (1 << R)) - 1);
    flip(__popcll(a & sg0), reg_pattern<R>(0));
    flip(__popcll(a & sg1), reg_pattern<R>(1));
    flip(__popcll(a & sg2), reg_pattern<R>(2));
    if constexpr (R == 4) flip(__popcll(a & sg3), reg_pattern<R>(3));
    sfor<0, (1 << R)>([&](auto J) {
        constexpr int j = decltype(J)::value;
        const uint32_t t = m << (31 - j);  // bit j of m at the sign position (garbage below, masked off)
        c2 v = getv<j>(s.v);
        const uint64_t bx = __builtin_bit_cast(uint64_t, v.x), by = __builtin_bit_cast(uint64_t, v.y);
        const uint64_t hx = xor_and((uint32_t)(bx >> 32), t, 0x80000000u);
        const uint64_t hy = xor_and((uint32_t)(by >> 32), t, 0x80000000u);
        v.x = __builtin_bit_cast(double, (bx & 0xFFFFFFFFull) | (hx << 32));
        v.y = __builtin_bit_cast(double, (by & 0xFFFFFFFFull) | (hy << 32));
        setv<j>(s.v, v);
    });
}

// Reference-semantics Kraus block update as a predicated superoperator
// (kind 4, four register bits only).  Q(gbase) is this thread's part of the
// block test; a register group is updated when Q(gbase) equals the uniform
// part of its register bits.  On two registers the 4x4 runs like a dense
// 2-bit op with a select per group.
__device__ __forceinline__ uint64_t kraus_q(const SubCube<4>& s, const TOp* __restrict__ op) {
    const uint64_t g = s.gbase;
    return (g ^ (g >> op->pd[0])) & op->pm[0];
}

template <int A, int B>
__device__ __forceinline__ void sw_kraus2(SubCube<4>& s, const TOp* __restrict__ op, const OpHead& hd) {
    const uint64_t q = kraus_q(s, op);
    const uint64_t s0 = op->sgm[0], s1 = op->sgm[1], s2 = op->sgm[2], s3 = op->sgm[3];
    // Reference-semantics blocks hold the registers whose row and column
    // agree on every other qubit, a small fraction: a wave none of whose
    // groups is in a block leaves the sub-cubes as they are (the selects
    // below would keep every value).
    bool any = false;
    sfor<0, 16>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & ((1 << A) | (1 << B))) == 0) {
            const uint64_t uj = ((j & 1) ? s0 : 0) ^ ((j & 2) ? s1 : 0) ^ ((j & 4) ? s2 : 0) ^ ((j & 8) ? s3 : 0);
            any = any || q == uj;
        }
    });
    if (!__any(any)) return;
    const c2 col0[4] = { hd.m0, hd.m1, hd.m2, hd.m3 };
    sfor<0, 16>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & ((1 << A) | (1 << B))) == 0) {
            constexpr int i1 = j | (1 << A), i2 = j | (1 << B), i3 = j | (1 << A) | (1 << B);
            const uint64_t uj = ((j & 1) ? s0 : 0) ^ ((j & 2) ? s1 : 0) ^ ((j & 4) ? s2 : 0) ^ ((j & 8) ? s3 : 0);
            const bool on = q == uj;
            const c2 a0 = getv<j>(s.v), a1 = getv<i1>(s.v), a2 = getv<i2>(s.v), a3 = getv<i3>(s.v);
            auto row = [&](int r) {
                c2 acc = cmul(col0[r], a0);
                acc = cfma(to_c2(op->m[r + 4]), a1, acc);
                acc = cfma(to_c2(op->m[r + 8]), a2, acc);
                return cfma(to_c2(op->m[r + 12]), a3, acc);
            };
            const c2 r0 = row(0), r1 = row(1), r2 = row(2), r3 = row(3);
            setv<j>(s.v, csel(on, r0, a0));
            setv<i1>(s.v, csel(on, r1, a1));
            setv<i2>(s.v, csel(on, r2, a2));
            setv<i3>(s.v, csel(on, r3, a3));
        }
    });
}

// On all four registers (a 2-qubit channel: both qubits' row and column
// bits) the block is the whole sub-cube, updated only where Q(gbase) == 0 --
// a few threads of a few tiles.  Those threads put the sub-cube back into
// their own LDS slots and accumulate one output at a time from there, so
// the 16x16 product needs no extra registers.
__device__ __forceinline__ void sw_kraus4(SubCube<4>& s, const TOp* __restrict__ op, double2* __restrict__ tile,
                                          int pl, int o0, int o1, int o2, int o3) {
    if (kraus_q(s, op) != 0) return;
    auto slot = [&](int j) {
        const int o = ((j & 1) ? o0 : 0) + ((j & 2) ? o1 : 0) + ((j & 4) ? o2 : 0) + ((j & 8) ? o3 : 0);
        return pl + phys(o);
    };
    sfor<0, 16>([&](auto J) {
        constexpr int j = decltype(J)::value;
        tile[slot(j)] = to_d2(getv<j>(s.v));
    });
    sfor<0, 16>([&](auto J) {
        constexpr int r = decltype(J)::value;
        c2 acc = { 0.0, 0.0 };
        for (int c = 0; c < 16; ++c) acc = cfma(to_c2(op[c].m[r]), to_c2(tile[slot(c)]), acc);
        setv<r>(s.v, acc);
    });
}

// A 16x16 superoperator of orbit form (kind 4, b = 1 + pairing): entries
// only where the row and column registers differ by an element of
// X = {0, P1, P2, P1 ^ P2}, so each orbit {o ^ X[r]} is updated by its own
// 4x4 block (op + k for the k-th orbit), in registers.
template <int P1, int P2>
__device__ __forceinline__ void sw_orbit4(SubCube<4>& s, const TOp* __restrict__ op) {
    constexpr int A1 = __builtin_ctz(P1), A2 = __builtin_ctz(P2);
    constexpr int REST = 0xF & ~((1 << A1) | (1 << A2));
    constexpr int B1 = __builtin_ctz(REST), B2 = __builtin_ctz(REST & (REST - 1));
    sfor<0, 16>([&](auto J) {
        constexpr int o = decltype(J)::value;
        if constexpr (((o >> A1) & 1) == 0 && ((o >> A2) & 1) == 0) {
            constexpr int k = ((o >> B1) & 1) | (((o >> B2) & 1) << 1);
            constexpr int i0 = o, i1 = o ^ P1, i2 = o ^ P2, i3 = o ^ P1 ^ P2;
            const TOp* __restrict__ blk = op + k;
            const c2 x0 = getv<i0>(s.v), x1 = getv<i1>(s.v), x2 = getv<i2>(s.v), x3 = getv<i3>(s.v);
            auto row = [&](int r) {
                c2 acc = cmul(to_c2(blk->m[r]), x0);
                acc = cfma(to_c2(blk->m[r + 4]), x1, acc);
                acc = cfma(to_c2(blk->m[r + 8]), x2, acc);
                return cfma(to_c2(blk->m[r + 12]), x3, acc);
            };
            const c2 y0 = row(0), y1 = row(1), y2 = row(2), y3 = row(3);
            setv<i0>(s.v, y0);
            setv<i1>(s.v, y1);
            setv<i2>(s.v, y2);
            setv<i3>(s.v, y3);
        }
    });
}

__device__ __forceinline__ void sw_kraus16(SubCube<4>& s, const TOp* __restrict__ op, const OpHead& hd,
                                           double2* __restrict__ tile, int pl, int o0, int o1, int o2, int o3) {
    switch (hd.b) {
        case 0: sw_kraus4(s, op, tile, pl, o0, o1, o2, o3); break;
        case 1: sw_orbit4<0x3, 0xC>(s, op); break;
        case 2: sw_orbit4<0x5, 0xA>(s, op); break;
        default: sw_orbit4<0x9, 0x6>(s, op); break;
    }
}

// Apply one tile op; returns the number of TOp entries it occupies.  KRAUS:
// the plan may hold reference Kraus ops (density-matrix plans only, R = 4;
// other instantiations carry no code for them).
template <int R, bool KRAUS>
__device__ __forceinline__ int run_op(SubCube<R>& s, const TOp* __restrict__ op, double2* __restrict__ tile, int pl,
                                      int o0, int o1, int o2, int o3) {
    const OpHead hd = load_head(op);
    if constexpr (KRAUS && R == 4) {
        if (hd.kind == 4) {
            if (hd.nd) {
                sw_kraus16(s, op, hd, tile, pl, o0, o1, o2, o3);
                return 1 + hd.nd;
            }
            switch (hd.a * 4 + hd.b) {
                case 1: sw_kraus2<0, 1>(s, op, hd); break;
                case 2: sw_kraus2<0, 2>(s, op, hd); break;
                case 3: sw_kraus2<0, 3>(s, op, hd); break;
                case 6: sw_kraus2<1, 2>(s, op, hd); break;
                case 7: sw_kraus2<1, 3>(s, op, hd); break;
                default: sw_kraus2<2, 3>(s, op, hd); break;
            }
            return 1;
        }
    }
    if constexpr (R == 4) {
        switch (hd.kind * 16 + hd.a * 4 + hd.b) {
            case 48: sw_sign<R>(s, hd); break;
            case 0: sw_dense1<R, 0>(s, op, hd); break;
            case 4: sw_dense1<R, 1>(s, op, hd); break;
            case 8: sw_dense1<R, 2>(s, op, hd); break;
            case 12: sw_dense1<R, 3>(s, op, hd); break;
            case 16 + 1: sw_dense2<R, 0, 1>(s, op, hd); break;
            case 16 + 2: sw_dense2<R, 0, 2>(s, op, hd); break;
            case 16 + 3: sw_dense2<R, 0, 3>(s, op, hd); break;
            case 16 + 6: sw_dense2<R, 1, 2>(s, op, hd); break;
            case 16 + 7: sw_dense2<R, 1, 3>(s, op, hd); break;
            case 16 + 11: sw_dense2<R, 2, 3>(s, op, hd); break;
            default:  // kinds 2 and 6 (kept out of the jump table)
                if (hd.kind == 6) sw_rdiag<R>(s, op, hd);
                else sw_diag<R>(s, op);
                break;
        }
    } else {
        switch (hd.kind * 16 + hd.a * 4 + hd.b) {
            case 48: sw_sign<R>(s, hd); break;
            case 0: sw_dense1<R, 0>(s, op, hd); break;
            case 4: sw_dense1<R, 1>(s, op, hd); break;
            case 8: sw_dense1<R, 2>(s, op, hd); break;
            case 16 + 1: sw_dense2<R, 0, 1>(s, op, hd); break;
            case 16 + 2: sw_dense2<R, 0, 2>(s, op, hd); break;
            case 16 + 6: sw_dense2<R, 1, 2>(s, op, hd); break;
            default:  // kinds 2, 6 and 7 (kept out of the jump table)
                if (hd.kind == 6) {
                    sw_rdiag<R>(s, op, hd);
                } else if (hd.kind == 7) {
                    switch (hd.a) {
                        case 0: sw_x<R, 0>(s, op, hd); break;
                        case 1: sw_x<R, 1>(s, op, hd); break;
                        default: sw_x<R, 2>(s, op, hd); break;
                    }
                } else {
                    sw_diag<R>(s, op);
                }
                break;
        }
    }
    return 1;
}

// One tile of one pass.  `init`: the state is |0...0> and is not read (the
// tile is built in LDS).  LDS: tile (2^T + 2^T/16 amplitudes), scat_lo[TB],
// scat_hi[64].
template <int R, bool KRAUS>
__device__ __forceinline__ void tile_body(double2* __restrict__ sv, int vq, const Pass* __restrict__ pass,
                                          const Sweep* __restrict__ sweeps, const TOp* __restrict__ tops,
                                          uint64_t tile_index, bool init, double2* __restrict__ tile,
                                          uint64_t* __restrict__ scat_lo, uint64_t* __restrict__ scat_hi) {
    constexpr int TB = tile_threads<R>(), TBB = tile_thread_bits<R>(), NR = 1 << R;
    const int T = pass->T;
    const int nloc = 1 << T;
    const uint64_t lmask = pass->lmask;
    {
        const int t = threadIdx.x;
        uint64_t s = 0;
        for (int b = 0; b < TBB && b < T; ++b)
            if ((t >> b) & 1) s |= 1ull << pass->lbits[b];
        scat_lo[t] = s;
        if (t < 64) {
            uint64_t h = 0;
            for (int b = TBB; b < T; ++b)
                if ((t >> (b - TBB)) & 1) h |= 1ull << pass->lbits[b];
            scat_hi[t] = h;
        }
    }
    // global base: deposit the block index into the non-tile bits
    uint64_t w = tile_index, base = 0;
    for (int bit = 0, k = 0; bit < vq; ++bit)
        if (!((lmask >> bit) & 1)) {
            base |= ((w >> k) & 1) << bit;
            ++k;
        }
    __syncthreads();
    const uint64_t s_t = base | scat_lo[threadIdx.x & (TB - 1)];
    const int iters = nloc >> TBB;  // rows of TB amplitudes (T >= TBB)
    if (init) {
        for (int j = threadIdx.x; j < nloc; j += TB)
            tile[phys(j)] = make_double2((base == 0 && j == 0) ? 1.0 : 0.0, 0.0);
    } else if (iters == 0) {
        if ((int)threadIdx.x < nloc) tile[phys(threadIdx.x)] = sv[s_t];
    }
    // issue up to NR loads per thread before the first LDS write
    for (int i0 = 0; i0 < iters && !init; i0 += NR) {
        dvec<2 * NR> buf;
        sfor<0, NR>([&](auto J) {
            constexpr int jj = decltype(J)::value;
            if (i0 + jj < iters) setv<jj>(buf, to_c2(sv[s_t | scat_hi[i0 + jj]]));
        });
        sfor<0, NR>([&](auto J) {
            constexpr int jj = decltype(J)::value;
            if (i0 + jj < iters) tile[phys((int)threadIdx.x + ((i0 + jj) << TBB))] = to_d2(getv<jj>(buf));
        });
    }
    __syncthreads();
    const int nsub = nloc >> R;
    for (int sidx = 0; sidx < pass->nsweeps; ++sidx) {
        const Sweep* sw = sweeps + pass->sweep_begin + sidx;
        const int sb0 = sw->sb[0], sb1 = sw->sb[1], sb2 = sw->sb[2], sb3 = R == 4 ? sw->sb[3] : 0;
        const int ob = sw->op_begin, oe = sw->op_end;
        // global bits of the register bits (uniform)
        auto gl = [&](int l) { return scat_lo[(1 << l) & (TB - 1)] | scat_hi[(1 << l) >> TBB]; };
        const uint64_t g0 = gl(sb0), g1 = gl(sb1), g2 = gl(sb2), g3 = R == 4 ? gl(sb3) : 0;
        // local offsets of the registers: the register bits are disjoint from the
        // sub-cube's base, so phys(lb | o) = phys(lb) + phys(o)
        const int o0 = 1 << sb0, o1 = 1 << sb1, o2 = 1 << sb2, o3 = R == 4 ? 1 << sb3 : 0;
        for (int c = threadIdx.x; c < nsub; c += TB) {
            // base of sub-cube c: zero bits inserted at the (ascending) register bits
            auto ins = [](int x, int b) { const int lo = x & ((1 << b) - 1); return ((x - lo) << 1) | lo; };
            int lb = ins(ins(ins(c, sb0), sb1), sb2);
            if constexpr (R == 4) lb = ins(lb, sb3);
            const int pl = phys(lb);
            SubCube<R> s;
            s.gbase = base | scat_lo[lb & (TB - 1)] | scat_hi[lb >> TBB];
            s.g0 = g0;
            s.g1 = g1;
            s.g2 = g2;
            s.g3 = g3;
            sfor<0, NR>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const int o = ((j & 1) ? o0 : 0) + ((j & 2) ? o1 : 0) + ((j & 4) ? o2 : 0) + ((j & 8) ? o3 : 0);
                setv<j>(s.v, to_c2(tile[pl + phys(o)]));
            });
            for (int o = ob; o < oe;) o += run_op<R, KRAUS>(s, tops + o, tile, pl, o0, o1, o2, o3);
            sfor<0, NR>([&](auto J) {
                constexpr int j = decltype(J)::value;
                const int o = ((j & 1) ? o0 : 0) + ((j & 2) ? o1 : 0) + ((j & 4) ? o2 : 0) + ((j & 8) ? o3 : 0);
                tile[pl + phys(o)] = to_d2(getv<j>(s.v));
            });
        }
        __syncthreads();
    }
    if (iters == 0) {
        if ((int)threadIdx.x < nloc) sv[s_t] = tile[phys(threadIdx.x)];
    }
    // read up to NR rows from LDS, then stream them out (non-temporal: the
    // pass never re-reads what it wrote)
    for (int i0 = 0; i0 < iters; i0 += NR) {
        dvec<2 * NR> buf;
        sfor<0, NR>([&](auto J) {
            constexpr int jj = decltype(J)::value;
            if (i0 + jj < iters) setv<jj>(buf, to_c2(tile[phys((int)threadIdx.x + ((i0 + jj) << TBB))]));
        });
        sfor<0, NR>([&](auto J) {
            constexpr int jj = decltype(J)::value;
            if (i0 + jj < iters) {
                double* p = (double*)(sv + (s_t | scat_hi[i0 + jj]));
                const c2 a = getv<jj>(buf);
                __builtin_nontemporal_store(a.x, p);
                __builtin_nontemporal_store(a.y, p + 1);
            }
        });
    }
}

// ---------------------------------------------------------------------------
// The streaming pass (T = 12, R = 4, no reference Kraus ops): a persistent
// workgroup per slot (two per CU) takes tiles from a ticket counter and
// loads the next tile's amplitudes into registers while the current tile
// runs its last sweep, so one tile's HBM latency hides behind the previous
// tile's arithmetic.  Tiles are handed out in ticket order: a slot that
// runs behind (HBM channel contention, a busier XCD) takes fewer tiles
// instead of holding up the end of the pass.
//
// Tickets: the first tile of workgroup b is tile b; after that each tile
// costs one ticket v = atomicAdd(ctr, 1) and names tile gridDim.x + v.  A
// pass of n tiles draws exactly n tickets (one per tile run, the last of
// each workgroup naming a tile >= n, which ends it), so the workgroup that
// draws v = n - 1 holds the last ticket and resets the counter to zero for
// the next pass on the stream.  No host-side clearing is needed, which
// keeps the pass valid inside a captured graph and across its replays.
//
// LEAN: the pass holds only uncontrolled dense 1-/2-register ops and sign
// ops (the brick-wall shape): only those op paths are compiled in, which
// frees registers and instruction cache for the rest.  LEAN + KRAUS: the
// same op set plus reference Kraus ops (density-matrix plans).
// ---------------------------------------------------------------------------

// Register chains (kinds 8..11, planned for the lean passes): the thread's
// 2x2 accumulator A (column-major), multiplied with uniform matrices and
// per-thread Z factors t = parity(gbase & mask), then applied once.
struct Acc {
    double a00r, a00i, a10r, a10i, a01r, a01i, a11r, a11i;
};

// parity(g & m) in bit 31, the sign bit of a double's high word
__device__ __forceinline__ uint32_t par31(uint64_t g, uint64_t m) { return (uint32_t)__popcll(g & m) << 31; }

__device__ __forceinline__ double negif(double x, uint32_t s31) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    return __builtin_bit_cast(double, b ^ ((uint64_t)s31 << 32));
}

// u a + v b (complex, uniform u and v)
__device__ __forceinline__ void cmadd2(double& re, double& im, c2 u, double ar, double ai, c2 v, double br,
                                       double bi) {
    re = fma(v.x, br, fma(-v.y, bi, fma(u.x, ar, -u.y * ai)));
    im = fma(v.x, bi, fma(v.y, br, fma(u.x, ai, u.y * ar)));
}

// The 2x2 per-thread matrix on register pair (a, b) in place, as
// pair_inplace with the matrix in VGPRs.
__device__ __forceinline__ void pair_inplace_v(double& ar, double& ai, double& br, double& bi, const Acc& A) {
    double pr, pi, qr, qi, t, t2;
    asm("v_mul_f64 %[pr], %[br], %[m01r]\n\t"
        "v_mul_f64 %[qr], %[ar], %[m10r]\n\t"
        "v_mul_f64 %[pi], %[bi], %[m01r]\n\t"
        "v_mul_f64 %[qi], %[ai], %[m10r]\n\t"
        "v_fma_f64 %[pr], -%[bi], %[m01i], %[pr]\n\t"
        "v_fma_f64 %[qr], -%[ai], %[m10i], %[qr]\n\t"
        "v_fma_f64 %[pi], %[br], %[m01i], %[pi]\n\t"
        "v_fma_f64 %[qi], %[ar], %[m10i], %[qi]\n\t"
        "v_fma_f64 %[t], -%[ai], %[m00i], %[pr]\n\t"
        "v_fma_f64 %[t2], -%[bi], %[m11i], %[qr]\n\t"
        "v_fma_f64 %[ai], %[ai], %[m00r], %[pi]\n\t"
        "v_fma_f64 %[bi], %[bi], %[m11r], %[qi]\n\t"
        "v_fma_f64 %[ai], %[ar], %[m00i], %[ai]\n\t"
        "v_fma_f64 %[bi], %[br], %[m11i], %[bi]\n\t"
        "v_fma_f64 %[ar], %[ar], %[m00r], %[t]\n\t"
        "v_fma_f64 %[br], %[br], %[m11r], %[t2]"
        : [ar] "+v"(ar), [ai] "+v"(ai), [br] "+v"(br), [bi] "+v"(bi), [pr] "=&v"(pr), [pi] "=&v"(pi),
          [qr] "=&v"(qr), [qi] "=&v"(qi), [t] "=&v"(t), [t2] "=&v"(t2)
        : [m00r] "v"(A.a00r), [m00i] "v"(A.a00i), [m10r] "v"(A.a10r), [m10i] "v"(A.a10i), [m01r] "v"(A.a01r),
          [m01i] "v"(A.a01i), [m11r] "v"(A.a11r), [m11i] "v"(A.a11i));
}

template <int R, int AB>
__device__ __forceinline__ void chain_apply(SubCube<R>& s, const Acc& A) {
    sfor<0, (1 << R)>([&](auto J) {
        constexpr int j = decltype(J)::value;
        if constexpr ((j & (1 << AB)) == 0) {
            constexpr int k = j | (1 << AB);
            double ar = s.v[2 * j], ai = s.v[2 * j + 1], br = s.v[2 * k], bi = s.v[2 * k + 1];
            pair_inplace_v(ar, ai, br, bi, A);
            s.v[2 * j] = ar;
            s.v[2 * j + 1] = ai;
            s.v[2 * k] = br;
            s.v[2 * k + 1] = bi;
        }
    });
}

// A = U Z^t A (one further matrix of a chain)
__device__ __forceinline__ void chain_mul(Acc& A, uint32_t z, c2 u0, c2 u1, c2 u2, c2 u3) {
    const double b10r = negif(A.a10r, z), b10i = negif(A.a10i, z), b11r = negif(A.a11r, z), b11i = negif(A.a11i, z);
    double n00r, n00i, n10r, n10i, n01r, n01i, n11r, n11i;
    cmadd2(n00r, n00i, u0, A.a00r, A.a00i, u2, b10r, b10i);
    cmadd2(n10r, n10i, u1, A.a00r, A.a00i, u3, b10r, b10i);
    cmadd2(n01r, n01i, u0, A.a01r, A.a01i, u2, b11r, b11i);
    cmadd2(n11r, n11i, u1, A.a01r, A.a01i, u3, b11r, b11i);
    A = Acc{ n00r, n00i, n10r, n10i, n01r, n01i, n11r, n11i };
}

__device__ __forceinline__ void run_chain(SubCube<4>& s, Acc& A, const TOp* __restrict__ op, const OpHead& hd) {
    if (hd.kind == 12) {  // a whole chain
        const int n = hd.nd, np = op->npairs;
        const uint64_t a = s.gbase;
        uint32_t z = par31(a, hd.cmask);
        if (np) {  // the sweep's per-thread sign rides on this chain
            uint32_t par = __popcll(a & (a >> op->pd[0]) & op->pm[0]);
            if (np > 1) {
                par += __popcll(a & (a >> op->pd[1]) & op->pm[1]);
                if (np > 2) {
                    par += __popcll(a & (a >> op->pd[2]) & op->pm[2]);
                    if (np > 3) par += __popcll(a & (a >> op->pd[3]) & op->pm[3]);
                }
            }
            const uint32_t p = par << 31;
            A.a00r = negif(hd.m0.x, p);
            A.a00i = negif(hd.m0.y, p);
            A.a10r = negif(hd.m1.x, p);
            A.a10i = negif(hd.m1.y, p);
            z ^= p;
        } else {
            A.a00r = hd.m0.x;
            A.a00i = hd.m0.y;
            A.a10r = hd.m1.x;
            A.a10i = hd.m1.y;
        }
        A.a01r = negif(hd.m2.x, z);
        A.a01i = negif(hd.m2.y, z);
        A.a11r = negif(hd.m3.x, z);
        A.a11i = negif(hd.m3.y, z);
        if (n > 1) {
            chain_mul(A, par31(a, op->sgm[0]), to_c2(op->m[4]), to_c2(op->m[5]), to_c2(op->m[6]), to_c2(op->m[7]));
            if (n > 2) {
                chain_mul(A, par31(a, op->sgm[1]), to_c2(op->m[8]), to_c2(op->m[9]), to_c2(op->m[10]),
                          to_c2(op->m[11]));
                if (n > 3)
                    chain_mul(A, par31(a, op->sgm[2]), to_c2(op->m[12]), to_c2(op->m[13]), to_c2(op->m[14]),
                              to_c2(op->m[15]));
            }
        }
        const uint64_t ze = op->sgm[3];
        if (ze) {
            const uint32_t zz = par31(a, ze);
            A.a10r = negif(A.a10r, zz);
            A.a10i = negif(A.a10i, zz);
            A.a11r = negif(A.a11r, zz);
            A.a11i = negif(A.a11i, zz);
        }
    } else if (hd.kind == 8) {  // A = U Z^t
        const uint32_t z = par31(s.gbase, hd.cmask);
        A.a00r = hd.m0.x;
        A.a00i = hd.m0.y;
        A.a10r = hd.m1.x;
        A.a10i = hd.m1.y;
        A.a01r = negif(hd.m2.x, z);
        A.a01i = negif(hd.m2.y, z);
        A.a11r = negif(hd.m3.x, z);
        A.a11i = negif(hd.m3.y, z);
        return;
    } else if (hd.kind == 9) {  // A = U Z^t A
        chain_mul(A, par31(s.gbase, hd.cmask), hd.m0, hd.m1, hd.m2, hd.m3);
        return;
    } else if (hd.kind == 10) {  // A = Z^t A, applied to register a
        if (hd.cmask) {
            const uint32_t z = par31(s.gbase, hd.cmask);
            A.a10r = negif(A.a10r, z);
            A.a10i = negif(A.a10i, z);
            A.a11r = negif(A.a11r, z);
            A.a11i = negif(A.a11i, z);
        }
    } else {  // 11: A = (-1)^P A
        const int np = (int)(hd.cmask >> 32);
        const int pd = hd.nd;
        const uint64_t a = s.gbase;
        uint32_t par = __popcll(a & (a >> (pd & 0xFF)) & __builtin_bit_cast(uint64_t, hd.m0.x));
        if (np > 1) {
            par += __popcll(a & (a >> ((pd >> 8) & 0xFF)) & __builtin_bit_cast(uint64_t, hd.m0.y));
            if (np > 2) {
                par += __popcll(a & (a >> ((pd >> 16) & 0xFF)) & __builtin_bit_cast(uint64_t, hd.m1.x));
                if (np > 3) par += __popcll(a & (a >> ((pd >> 24) & 0xFF)) & __builtin_bit_cast(uint64_t, hd.m1.y));
            }
        }
        const uint32_t z = par << 31;
        A.a00r = negif(A.a00r, z);
        A.a00i = negif(A.a00i, z);
        A.a10r = negif(A.a10r, z);
        A.a10i = negif(A.a10i, z);
        A.a01r = negif(A.a01r, z);
        A.a01i = negif(A.a01i, z);
        A.a11r = negif(A.a11r, z);
        A.a11i = negif(A.a11i, z);
        return;
    }
    // kinds 10 and 12: apply A (the switch stays in this function: moved
    // into a helper, the sub-cube went to scratch memory)
    switch (hd.a) {
        case 0: chain_apply<4, 0>(s, A); break;
        case 1: chain_apply<4, 1>(s, A); break;
        case 2: chain_apply<4, 2>(s, A); break;
        default: chain_apply<4, 3>(s, A); break;
    }
}

__device__ __forceinline__ int run_op_lean(SubCube<4>& s, Acc& A, const TOp* __restrict__ op) {
    const OpHead hd = load_head(op);
    // one key for both tests, so that every head field is read in the op's
    // first scalar round trip (testing the kind alone let the compiler load
    // b in a second one)
    const int key = hd.kind * 16 + hd.a * 4 + hd.b;
    if (key >= 8 * 16) {  // register chains: kind >= 8 (a < 4, b = 0)
        run_chain(s, A, op, hd);
        return 1;
    }
    switch (key) {
        case 48: sw_sign<4>(s, hd); break;
        case 0: sw_dense1_t<4, 0, false>(s, op, hd); break;
        case 4: sw_dense1_t<4, 1, false>(s, op, hd); break;
        case 8: sw_dense1_t<4, 2, false>(s, op, hd); break;
        case 12: sw_dense1_t<4, 3, false>(s, op, hd); break;
        case 16 + 1: sw_dense2_t<4, 0, 1, false>(s, op, hd); break;
        case 16 + 2: sw_dense2_t<4, 0, 2, false>(s, op, hd); break;
        case 16 + 3: sw_dense2_t<4, 0, 3, false>(s, op, hd); break;
        case 16 + 6: sw_dense2_t<4, 1, 2, false>(s, op, hd); break;
        case 16 + 7: sw_dense2_t<4, 1, 3, false>(s, op, hd); break;
        default: sw_dense2_t<4, 2, 3, false>(s, op, hd); break;
    }
    return 1;
}

// The lean op set plus the reference Kraus ops of density-matrix plans.
__device__ __forceinline__ int run_op_lean_kraus(SubCube<4>& s, Acc& A, const TOp* __restrict__ op,
                                                 double2* __restrict__ tile, int pl, int o0, int o1, int o2,
                                                 int o3) {
    const OpHead hd = load_head(op);
    const int key = hd.kind * 16 + hd.a * 4 + hd.b;  // as in run_op_lean
    if (key >= 8 * 16) {
        run_chain(s, A, op, hd);
        return 1;
    }
    if (hd.kind == 4) {
        if (hd.nd) {
            sw_kraus16(s, op, hd, tile, pl, o0, o1, o2, o3);
            return 1 + hd.nd;
        }
        switch (hd.a * 4 + hd.b) {
            case 1: sw_kraus2<0, 1>(s, op, hd); break;
            case 2: sw_kraus2<0, 2>(s, op, hd); break;
            case 3: sw_kraus2<0, 3>(s, op, hd); break;
            case 6: sw_kraus2<1, 2>(s, op, hd); break;
            case 7: sw_kraus2<1, 3>(s, op, hd); break;
            default: sw_kraus2<2, 3>(s, op, hd); break;
        }
        return 1;
    }
    switch (key) {
        case 48: sw_sign<4>(s, hd); break;
        case 0: sw_dense1_t<4, 0, false>(s, op, hd); break;
        case 4: sw_dense1_t<4, 1, false>(s, op, hd); break;
        case 8: sw_dense1_t<4, 2, false>(s, op, hd); break;
        case 12: sw_dense1_t<4, 3, false>(s, op, hd); break;
        case 16 + 1: sw_dense2_t<4, 0, 1, false>(s, op, hd); break;
        case 16 + 2: sw_dense2_t<4, 0, 2, false>(s, op, hd); break;
        case 16 + 3: sw_dense2_t<4, 0, 3, false>(s, op, hd); break;
        case 16 + 6: sw_dense2_t<4, 1, 2, false>(s, op, hd); break;
        case 16 + 7: sw_dense2_t<4, 1, 3, false>(s, op, hd); break;
        default: sw_dense2_t<4, 2, 3, false>(s, op, hd); break;
    }
    return 1;
}

// Workgroup barrier over LDS traffic only: this wave's LDS reads and writes
// (and scalar loads) have completed when it passes; its global loads and
// stores may still be in flight.  __syncthreads() also waits for those
// (vmcnt(0)), which would make every sweep wait for the next tile's
// prefetch.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// LDS: tile (2^12 + 2^8 amplitudes), scat_lo[256], scat_hi[16], s_next[2].
// The launch has min(ntiles, slots) workgroups of 256 threads.
template <bool LEAN, bool KRAUS>
__device__ __forceinline__ void stream_body(double2* __restrict__ sv, int vq, const Pass* __restrict__ pass,
                                            const Sweep* __restrict__ sweeps, const TOp* __restrict__ tops,
                                            uint64_t ntiles, unsigned long long* __restrict__ ctr,
                                            double2* __restrict__ tile, uint64_t* __restrict__ scat_lo,
                                            uint64_t* __restrict__ scat_hi, uint64_t* s_next) {
    constexpr int R = 4, T = 12, TB = tile_threads<R>(), TBB = tile_thread_bits<R>(), NR = 1 << R;
    static_assert(T - TBB == R, "one row of the tile per register");
    const uint64_t lmask = pass->lmask;
    {
        const int t = threadIdx.x;
        uint64_t sc = 0;
        for (int b = 0; b < TBB; ++b)
            if ((t >> b) & 1) sc |= 1ull << pass->lbits[b];
        scat_lo[t] = sc;
        if (t < NR) {
            uint64_t h = 0;
            for (int b = TBB; b < T; ++b)
                if ((t >> (b - TBB)) & 1) h |= 1ull << pass->lbits[b];
            scat_hi[t] = h;
        }
    }
    lds_barrier();
    // global base of tile w: w deposited into the non-tile bits
    auto tile_base = [&](uint64_t w) {
        uint64_t b = 0;
        for (int bit = 0, k = 0; bit < vq; ++bit)
            if (!((lmask >> bit) & 1)) {
                b |= ((w >> k) & 1) << bit;
                ++k;
            }
        return b;
    };
    const int nsw = pass->nsweeps;
    Acc acc{};  // register chains of the lean op sets
    dvec<2 * NR> pre;  // this thread's amplitudes of the next tile (row jj = local j = tid + TB jj)
    auto load = [&](uint64_t b) {
        const uint64_t s_t = b | scat_lo[threadIdx.x];
        sfor<0, NR>([&](auto J) {
            constexpr int jj = decltype(J)::value;
            setv<jj>(pre, to_c2(sv[s_t | scat_hi[jj]]));
        });
    };
    auto to_lds = [&]() {
        sfor<0, NR>([&](auto J) {
            constexpr int jj = decltype(J)::value;
            tile[phys((int)threadIdx.x + (jj << TBB))] = to_d2(getv<jj>(pre));
        });
        lds_barrier();
    };
    // one sweep: a 12-bit tile holds exactly one sub-cube per thread
    auto sweep = [&](int sidx, uint64_t base) {
        const Sweep* sw = sweeps + pass->sweep_begin + sidx;
        const int sb0 = sw->sb[0], sb1 = sw->sb[1], sb2 = sw->sb[2], sb3 = sw->sb[3];
        const int ob = sw->op_begin, oe = sw->op_end;
        auto gl = [&](int l) { return scat_lo[(1 << l) & (TB - 1)] | scat_hi[(1 << l) >> TBB]; };
        const int o0 = 1 << sb0, o1 = 1 << sb1, o2 = 1 << sb2, o3 = 1 << sb3;
        auto ins = [](int x, int b) { const int lo = x & ((1 << b) - 1); return ((x - lo) << 1) | lo; };
        const int lb = ins(ins(ins(ins((int)threadIdx.x, sb0), sb1), sb2), sb3);
        const int pl = phys(lb);
        SubCube<R> sc;
        sc.gbase = base | scat_lo[lb & (TB - 1)] | scat_hi[lb >> TBB];
        sc.g0 = gl(sb0);
        sc.g1 = gl(sb1);
        sc.g2 = gl(sb2);
        sc.g3 = gl(sb3);
        sfor<0, NR>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const int o = ((j & 1) ? o0 : 0) + ((j & 2) ? o1 : 0) + ((j & 4) ? o2 : 0) + ((j & 8) ? o3 : 0);
            setv<j>(sc.v, to_c2(tile[pl + phys(o)]));
        });
        if constexpr (LEAN && KRAUS) {
            for (int o = ob; o < oe;) o += run_op_lean_kraus(sc, acc, tops + o, tile, pl, o0, o1, o2, o3);
        } else if constexpr (LEAN) {
            for (int o = ob; o < oe;) o += run_op_lean(sc, acc, tops + o);
        } else {
            for (int o = ob; o < oe;) o += run_op<R, false>(sc, tops + o, tile, pl, o0, o1, o2, o3);
        }
        sfor<0, NR>([&](auto J) {
            constexpr int j = decltype(J)::value;
            const int o = ((j & 1) ? o0 : 0) + ((j & 2) ? o1 : 0) + ((j & 4) ? o2 : 0) + ((j & 8) ? o3 : 0);
            tile[pl + phys(o)] = to_d2(getv<j>(sc.v));
        });
        lds_barrier();
    };
    // Tiles: the workgroup's first tile is blockIdx.x; each ticket v names
    // tile gridDim.x + v.  The ticket of the tile after next is drawn once
    // the current tile's sweeps are done and before its stores, so the
    // atomic's wait (an in-order vmcnt) finds the memory pipe nearly idle.
    uint64_t base = tile_base(blockIdx.x);  // blockIdx.x < ntiles (grid <= ntiles)
    auto ticket = [&](int slot) {
        if (threadIdx.x == 0) {
            const unsigned long long v = atomicAdd(ctr, 1ull);
            if (v == ntiles - 1) atomicExch(ctr, 0ull);  // the pass's last ticket
            s_next[slot] = gridDim.x + v;
        }
    };
    load(base);
    ticket(0);
    to_lds();  // its barrier publishes s_next[0]
    uint64_t next = s_next[0];
    for (int it = 1;; ++it) {
        // after the last tile, the prefetch re-reads this tile's amplitudes
        // (in bounds, discarded), so every path issues the same memory
        // operations in the same order
        const uint64_t nbase = next < ntiles ? tile_base(next) : base;
        load(nbase);  // in flight during every sweep (their barriers wait for LDS traffic only)
        for (int sidx = 0; sidx < nsw; ++sidx) sweep(sidx, base);
        // a workgroup stops drawing after its first ticket past the end, so
        // a pass draws exactly ntiles tickets; slots alternate, so a write
        // never meets a read of the previous value that has not passed the
        // store barrier
        if (next < ntiles) ticket(it & 1);
        const uint64_t s_t = base | scat_lo[threadIdx.x];
        sfor<0, NR>([&](auto J) {
            constexpr int jj = decltype(J)::value;
            const c2 a = to_c2(tile[phys((int)threadIdx.x + (jj << TBB))]);
            double* p = (double*)(sv + (s_t | scat_hi[jj]));
            __builtin_nontemporal_store(a.x, p);
            __builtin_nontemporal_store(a.y, p + 1);
        });
        lds_barrier();  // the next tile rewrites LDS; s_next[it & 1] visible
        if (next >= ntiles) break;
        base = nbase;
        next = s_next[it & 1];
        to_lds();
    }
}

}  // namespace bsim
