#include <string.h>
// Test helper (CPU only): plans an instruction list with libbsim's own host
// planner (lower_ops + make_plan) and interprets the plan -- every pass,
// sweep and tile op with its documented meaning on global indices -- on a
// full state vector, without tiles.  Comparing the result with the oracle
// checks the planner's lowering, ordering, fusion and index-test tables on
// the CPU.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>
#include "fused.h"
using namespace bsim;
typedef std::complex<double> cd;
// Apply tile op o of plan p (R register bits vb[0..R), their union G) to the
// full vector v.
static void apply_top(const Plan& p, int o, int R, const int* vb, uint64_t G, uint64_t N, cd* v) {
    const TOp& op = p.tops[o];
    auto M = [&](int j) { return cd(op.m[j].x, op.m[j].y); };
    auto pred = [&](uint64_t i) { return ((i ^ (i >> op.pd[0])) & op.pm[0]) == 0; };
    if (op.kind == 5) return;  // matrix columns of the preceding kind-4 op
    if (op.kind == 4 && op.nd == 0) {
        uint64_t ba = 1ull << vb[op.a], bb = 1ull << vb[op.b];
        for (uint64_t i = 0; i < N; ++i) if (!(i & ba) && !(i & bb) && pred(i)) {
            uint64_t id[4] = {i, i | ba, i | bb, i | ba | bb};
            cd x[4], y[4];
            for (int k = 0; k < 4; ++k) x[k] = v[id[k]];
            for (int r = 0; r < 4; ++r) { y[r] = 0; for (int c = 0; c < 4; ++c) y[r] += M(r + 4 * c) * x[c]; }
            for (int k = 0; k < 4; ++k) v[id[k]] = y[k];
        }
    } else if (op.kind == 4) {
        // the 16x16 in register order: columns in the 16 TOps, or (b = 1 +
        // pairing) the orbit blocks: entry (o ^ X[r], o ^ X[c]) of orbit k
        // in TOp k, m[r + 4 c], every other entry zero
        std::vector<cd> F(256, cd(0.0, 0.0));
        if (op.b == 0) {
            for (int c = 0; c < 16; ++c)
                for (int r = 0; r < 16; ++r) F[r + 16 * c] = cd(p.tops[o + c].m[r].x, p.tops[o + c].m[r].y);
        } else {
            const int P1[3] = { 0x3, 0x5, 0x9 }, P2[3] = { 0xC, 0xA, 0x6 };
            const int q1 = P1[op.b - 1], q2 = P2[op.b - 1], X[4] = { 0, q1, q2, q1 ^ q2 };
            const int a1 = __builtin_ctz(q1), a2 = __builtin_ctz(q2);
            int k = 0;
            for (int ob = 0; ob < 16; ++ob) {
                if (((ob >> a1) & 1) || ((ob >> a2) & 1)) continue;
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c)
                        F[(ob ^ X[r]) + 16 * (ob ^ X[c])] =
                            cd(p.tops[o + k].m[r + 4 * c].x, p.tops[o + k].m[r + 4 * c].y);
                ++k;
            }
        }
        for (uint64_t i = 0; i < N; ++i) {
            if (i & G) continue;
            if (!pred(i)) continue;
            uint64_t id[16];
            for (int j = 0; j < 16; ++j) {
                id[j] = i;
                for (int r = 0; r < 4; ++r) if ((j >> r) & 1) id[j] |= 1ull << vb[r];
            }
            cd x[16], y[16];
            for (int k = 0; k < 16; ++k) x[k] = v[id[k]];
            for (int r = 0; r < 16; ++r) {
                y[r] = 0;
                for (int c = 0; c < 16; ++c) y[r] += F[r + 16 * c] * x[c];
            }
            for (int k = 0; k < 16; ++k) v[id[k]] = y[k];
        }
    } else if (op.kind == 0 || op.kind == 7) {  // 7: m holds X
        uint64_t bit = 1ull << vb[op.a];
        for (uint64_t i = 0; i < N; ++i) if (!(i & bit)) {
            if (op.cmask && ((i & op.cmask) != op.cval)) continue;
            cd a = v[i], b = v[i | bit];
            v[i] = M(0) * a + M(2) * b; v[i | bit] = M(1) * a + M(3) * b;
        }
    } else if (op.kind == 1) {
        uint64_t ba = 1ull << vb[op.a], bb = 1ull << vb[op.b];
        for (uint64_t i = 0; i < N; ++i) if (!(i & ba) && !(i & bb)) {
            if (op.cmask && ((i & op.cmask) != op.cval)) continue;
            uint64_t id[4] = {i, i | ba, i | bb, i | ba | bb};
            cd x[4], y[4];
            for (int k = 0; k < 4; ++k) x[k] = v[id[k]];
            for (int r = 0; r < 4; ++r) { y[r] = 0; for (int c = 0; c < 4; ++c) y[r] += M(r + 4 * c) * x[c]; }
            for (int k = 0; k < 4; ++k) v[id[k]] = y[k];
        }
    } else if (op.kind == 6) {
        for (uint64_t i = 0; i < N; ++i) {
            if (op.cmask && ((i & op.cmask) != op.cval)) continue;
            int j = 0;
            for (int r = 0; r < R; ++r) j |= (int)((i >> vb[r]) & 1) << r;
            v[i] *= M(j);
        }
    } else if (op.kind == 3 && op.split_only) {
        // defined by the split form: jmask, per-thread pairs, register terms
        for (uint64_t i = 0; i < N; ++i) {
            const uint64_t g = i & ~G;
            int j = 0;
            for (int r = 0; r < R; ++r) j |= (int)((i >> vb[r]) & 1) << r;
            int par = (int)((op.jmask >> j) & 1);
            for (int k = 0; k < op.npairs; ++k) par ^= __builtin_popcountll(g & (g >> op.pd[k]) & op.pm[k]) & 1;
            for (int r = 0; r < R; ++r)
                if ((j >> r) & 1) par ^= __builtin_popcountll(g & op.sgm[r]) & 1;
            if (par) v[i] = -v[i];
        }
    } else if (op.kind == 3) {
        for (uint64_t i = 0; i < N; ++i) {
            int par = 0;
            for (int k = 0; k < op.npairs; ++k) par ^= __builtin_popcountll(i & (i >> op.pd[k]) & op.pm[k]) & 1;
            if (par) v[i] = -v[i];
        }
    } else {
        for (uint64_t i = 0; i < N; ++i) {
            if (op.cmask && ((i & op.cmask) != op.cval)) continue;
            int idx = 0;
            for (int b = 0; b < op.nd; ++b) idx |= (int)((i >> op.dbit[b]) & 1) << b;
            v[i] *= M(idx);
        }
    }
    }

// The per-thread / per-register split of every index test of sweep S must
// agree with the direct test on every index.
static int check_sweep(const Plan& p, const Sweep& S, int R, const int* vb, uint64_t G, uint64_t N) {
    for (int o = S.op_begin; o < S.op_end; ++o) {
        const TOp& op = p.tops[o];
        if ((op.kind == 0 || op.kind == 1 || op.kind == 7 || (op.kind == 4 && op.nd == 0)) && (op.a >= R || op.b >= R))
            return -15;  // a register bit the pass's sweeps do not have
        for (uint64_t i = 0; i < N; ++i) {
            const uint64_t gbase = i & ~G;
            int j = 0;
            for (int r = 0; r < R; ++r) j |= (int)((i >> vb[r]) & 1) << r;
            if (op.kind == 3) {
                // the head copies the kernel reads (nd, cmask, m[0..3])
                uint64_t w[8];
                memcpy(w, op.m, sizeof(w));
                for (int k = 0; k < 4; ++k)
                    if (((op.nd >> (8 * k)) & 0xFF) != (op.pd[k] & 0xFF) || w[k] != op.pm[k] || w[4 + k] != op.sgm[k])
                        return -16;
                if ((uint32_t)op.cmask != op.jmask || (int)(op.cmask >> 32) != op.npairs) return -16;
                if (op.split_only) continue;  // no pair form to compare with
                int direct = 0, split = (int)((op.jmask >> j) & 1);
                for (int k = 0; k < op.npairs; ++k) {
                    direct ^= __builtin_popcountll(i & (i >> op.pd[k]) & op.pm[k]) & 1;
                    split ^= __builtin_popcountll(gbase & (gbase >> op.pd[k]) & op.pm[k]) & 1;
                }
                for (int r = 0; r < R; ++r)
                    if ((j >> r) & 1) split ^= __builtin_popcountll(gbase & op.sgm[r]) & 1;
                if (direct != split) return -10;
            } else if ((op.kind <= 1 || op.kind == 6 || op.kind == 7) && op.cmask) {
                const bool direct = (i & op.cmask) == op.cval;
                const bool split = ((gbase & op.tmask) == op.tval) && ((op.jmask >> j) & 1);
                if (direct != split) return -11;
            } else if (op.kind == 4) {
                const bool direct = ((i ^ (i >> op.pd[0])) & op.pm[0]) == 0;
                if (R != 4) return -14;  // kind 4 runs in 4-register passes only
                uint64_t u = 0;
                for (int r = 0; r < 4; ++r)
                    if ((j >> r) & 1) u ^= op.sgm[r];
                const bool split = ((gbase ^ (gbase >> op.pd[0])) & op.pm[0]) == u;
                if (direct != split) return -12;
            }
        }
    }
    return 0;
}

static void apply_eop_cpu(const EOp& e, uint64_t N, cd* v);

// A register chain (kinds 8, 9 and 11 up to its kind 10 at index `end`) with
// its documented per-sub-cube meaning: A = U Z^t ..., t = parity(gbase &
// mask), applied to register a of every sub-cube.  Returns the index after
// the chain, or -1 when it is malformed.
// A whole chain in one op (kind 12).
static int apply_chain12(const Plan& p, int o, int R, const int* vb, uint64_t G, uint64_t N, cd* v) {
    const TOp& op = p.tops[o];
    if (op.a < 0 || op.a >= R || op.nd < 1 || op.nd > 4 || op.npairs < 0 || op.npairs > 4) return -17;
    auto M = [&](int j) { return cd(op.m[j].x, op.m[j].y); };
    for (uint64_t g = 0; g < N; ++g) {
        if (g & G) continue;
        auto t = [&](uint64_t mask) { return __builtin_popcountll(g & mask) & 1; };
        const int t0 = t(op.cmask);
        cd A[4] = {M(0), M(1), t0 ? -M(2) : M(2), t0 ? -M(3) : M(3)};
        for (int k = 1; k < op.nd; ++k) {
            const int tk = t(op.sgm[k - 1]);
            const cd B[4] = {A[0], tk ? -A[1] : A[1], A[2], tk ? -A[3] : A[3]};
            A[0] = M(4 * k) * B[0] + M(4 * k + 2) * B[1];
            A[1] = M(4 * k + 1) * B[0] + M(4 * k + 3) * B[1];
            A[2] = M(4 * k) * B[2] + M(4 * k + 2) * B[3];
            A[3] = M(4 * k + 1) * B[2] + M(4 * k + 3) * B[3];
        }
        if (t(op.sgm[3])) { A[1] = -A[1]; A[3] = -A[3]; }
        int par = 0;
        for (int k = 0; k < op.npairs; ++k) par ^= __builtin_popcountll(g & (g >> op.pd[k]) & op.pm[k]) & 1;
        if (par)
            for (int k = 0; k < 4; ++k) A[k] = -A[k];
        for (int j = 0; j < (1 << R); ++j) {
            if ((j >> op.a) & 1) continue;
            uint64_t x = g;
            for (int r = 0; r < R; ++r)
                if ((j >> r) & 1) x |= 1ull << vb[r];
            const uint64_t y = x | (1ull << vb[op.a]);
            const cd u = v[x], w = v[y];
            v[x] = A[0] * u + A[2] * w;
            v[y] = A[1] * u + A[3] * w;
        }
    }
    return 0;
}

static int apply_chain(const Plan& p, int o, int oe, int R, const int* vb, uint64_t G, uint64_t N, cd* v) {
    int end = o;
    while (end < oe && p.tops[end].kind != 10) ++end;
    if (end == oe) return -1;
    const int a = p.tops[end].a;
    if (a < 0 || a >= R) return -1;
    for (uint64_t g = 0; g < N; ++g) {
        if (g & G) continue;
        cd A[4] = {1.0, 0.0, 0.0, 1.0};
        for (int q = o; q <= end; ++q) {
            const TOp& op = p.tops[q];
            const int t = __builtin_popcountll(g & op.cmask) & 1;
            auto M = [&](int j) { return cd(op.m[j].x, op.m[j].y); };
            if (op.kind == 8) {
                if (q != o) return -1;
                A[0] = M(0); A[1] = M(1); A[2] = t ? -M(2) : M(2); A[3] = t ? -M(3) : M(3);
            } else if (op.kind == 9 || op.kind == 10) {
                if (op.kind == 9 && !op.cmask) return -1;
                cd B[4] = {A[0], t ? -A[1] : A[1], A[2], t ? -A[3] : A[3]};
                if (op.kind == 10) {
                    for (int k = 0; k < 4; ++k) A[k] = B[k];
                } else {
                    A[0] = M(0) * B[0] + M(2) * B[1];
                    A[1] = M(1) * B[0] + M(3) * B[1];
                    A[2] = M(0) * B[2] + M(2) * B[3];
                    A[3] = M(1) * B[2] + M(3) * B[3];
                }
            } else if (op.kind == 11) {
                const int np = (int)(op.cmask >> 32);
                uint64_t pm[4];
                memcpy(pm, op.m, sizeof(pm));
                int par = 0;
                for (int k = 0; k < np; ++k) par ^= __builtin_popcountll(g & (g >> ((op.nd >> (8 * k)) & 0xFF)) & pm[k]) & 1;
                if (par)
                    for (int k = 0; k < 4; ++k) A[k] = -A[k];
            } else {
                return -1;
            }
        }
        for (int j = 0; j < (1 << R); ++j) {
            if ((j >> a) & 1) continue;
            uint64_t x = g;
            for (int r = 0; r < R; ++r)
                if ((j >> r) & 1) x |= 1ull << vb[r];
            const uint64_t y = x | (1ull << vb[a]);
            const cd u = v[x], w = v[y];
            v[x] = A[0] * u + A[2] * w;
            v[y] = A[1] * u + A[3] * w;
        }
    }
    return end + 1;
}

static int run(int nq, int kind, const bsim_op* ops, int n, double* state) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, kind, ops, n, e, false)) return rc;
    const int vq = kind == BSIM_SV ? nq : 2 * nq;
    Plan p;
    if (int rc = make_plan(vq, e, tile_config(), p)) return rc;
    cd* v = (cd*)state;
    const uint64_t N = 1ull << vq;
    for (auto& st : p.steps) {
        if (!st.tile) {
            // a single launch: the op itself, or the product of a merged run
            const EOp& se = st.merged >= 0 ? p.merged[st.merged] : e[st.op];
            if (se.kind == E_KRAUS_REF) { printf("Kraus step\n"); return -99; }
            apply_eop_cpu(se, N, v);
            continue;
        }
        const Pass& P = p.passes[st.pass];
        for (int s = 0; s < P.nsweeps; ++s) {
            const Sweep& S = p.sweeps[P.sweep_begin + s];
            const int R = P.R;  // register bits of the pass's sweeps
            if (R != 3 && R != 4) return -13;
            int vb[4] = {0, 0, 0, 0};
            for (int r = 0; r < R; ++r) vb[r] = P.lbits[S.sb[r]];
            uint64_t G = 0;
            for (int r = 0; r < R; ++r) G |= 1ull << vb[r];
            if (int rc = check_sweep(p, S, R, vb, G, N)) return rc;
            for (int o = S.op_begin; o < S.op_end;) {
                const int k = p.tops[o].kind;
                if (k == 12) {
                    if (int rc = apply_chain12(p, o, R, vb, G, N, v)) return rc;
                    ++o;
                } else if (k >= 8 && k <= 11) {
                    if (k != 8) return -17;  // a chain starts with kind 8
                    const int nx = apply_chain(p, o, S.op_end, R, vb, G, N, v);
                    if (nx < 0) return -17;
                    o = nx;
                } else {
                    apply_top(p, o, R, vb, G, N, v);
                    ++o;
                }
            }
        }
    }
    return 0;
}

extern "C" int planexec(int nq, const bsim_op* ops, int n, double* state) { return run(nq, BSIM_SV, ops, n, state); }
// density matrix: state is the stored 4^n vector (column-major)
extern "C" int planexec_dm(int nq, const bsim_op* ops, int n, double* state) { return run(nq, BSIM_DM, ops, n, state); }
// number of plan steps that are not tile passes
extern "C" int plan_singles(int nq, int kind, const bsim_op* ops, int n) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, kind, ops, n, e, false)) return rc;
    Plan p;
    if (int rc = make_plan(kind == BSIM_SV ? nq : 2 * nq, e, tile_config(), p)) return rc;
    int k = 0;
    for (auto& st : p.steps) k += st.tile ? 0 : 1;
    return k;
}
// number of tile passes
extern "C" int plan_passes(int nq, int kind, const bsim_op* ops, int n) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, kind, ops, n, e, false)) return rc;
    Plan p;
    if (int rc = make_plan(kind == BSIM_SV ? nq : 2 * nq, e, tile_config(), p)) return rc;
    return (int)p.passes.size();
}
// number of single steps that launch a merged run of ops
extern "C" int plan_merged(int nq, int kind, const bsim_op* ops, int n) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, kind, ops, n, e, false)) return rc;
    Plan p;
    if (int rc = make_plan(kind == BSIM_SV ? nq : 2 * nq, e, tile_config(), p)) return rc;
    int k = 0;
    for (auto& st : p.steps) k += (!st.tile && st.merged >= 0) ? 1 : 0;
    return k;
}

// ---- fused adjoint backward plan (state vectors) ----------------------------

// Re sum_i conj(lam_i) (D psi)_i for tile op `op` with product payload d
// (D = d.m in the op's own layout), over the indices where the op's
// controls hold.
static double dot_top(const TOp& op, const TOp& d, const int* vb, uint64_t N, const cd* psi, const cd* lam) {
    auto D = [&](int j) { return cd(d.m[j].x, d.m[j].y); };
    auto ctl = [&](uint64_t i) { return op.cmask == 0 || (i & op.cmask) == op.cval; };
    double acc = 0.0;
    if (op.kind == 0) {
        const uint64_t bit = 1ull << vb[op.a];
        for (uint64_t i = 0; i < N; ++i)
            if (!(i & bit) && ctl(i)) {
                const cd a = psi[i], b = psi[i | bit];
                acc += (std::conj(lam[i]) * (D(0) * a + D(2) * b) + std::conj(lam[i | bit]) * (D(1) * a + D(3) * b)).real();
            }
    } else if (op.kind == 1) {
        const uint64_t ba = 1ull << vb[op.a], bb = 1ull << vb[op.b];
        for (uint64_t i = 0; i < N; ++i)
            if (!(i & ba) && !(i & bb) && ctl(i)) {
                const uint64_t id[4] = { i, i | ba, i | bb, i | ba | bb };
                for (int r = 0; r < 4; ++r) {
                    cd y = 0;
                    for (int c = 0; c < 4; ++c) y += D(r + 4 * c) * psi[id[c]];
                    acc += (std::conj(lam[id[r]]) * y).real();
                }
            }
    } else if (op.kind == 6) {
        for (uint64_t i = 0; i < N; ++i)
            if (ctl(i)) {
                int j = 0;
                for (int r = 0; r < 3; ++r) j |= (int)((i >> vb[r]) & 1) << r;
                acc += (std::conj(lam[i]) * D(j) * psi[i]).real();
            }
    } else if (op.kind == 2) {
        for (uint64_t i = 0; i < N; ++i)
            if (ctl(i)) {
                int idx = 0;
                for (int b = 0; b < op.nd; ++b) idx |= (int)((i >> op.dbit[b]) & 1) << b;
                acc += (std::conj(lam[i]) * D(idx) * psi[i]).real();
            }
    }
    return acc;
}

// Im <lam| P |psi> for a Pauli P (1 X, 2 Y, 3 Z) on vector bit `bit`, from
// its definition: (X v)_i = v_(i^b), (Y v)_i = -i v_(i^b) if bit b of i is 0
// else i v_(i^b), (Z v)_i = -v_i if bit b of i is set.
static double pauli_im(int bit, int P, uint64_t N, const cd* psi, const cd* lam) {
    const uint64_t b = 1ull << bit;
    const cd I(0.0, 1.0);
    cd acc = 0.0;
    for (uint64_t i = 0; i < N; ++i) {
        const bool one = (i & b) != 0;
        cd y;
        if (P == 1) y = psi[i ^ b];
        else if (P == 2) y = (one ? I : -I) * psi[i ^ b];
        else y = one ? -psi[i] : psi[i];
        acc += std::conj(lam[i]) * y;
    }
    return acc.imag();
}

// A non-tile step: the elementary op itself (dense or diagonal, any width).
static void apply_eop_cpu(const EOp& e, uint64_t N, cd* v) {
    const GateSpec& g = e.g;
    const int D = 1 << g.k;
    uint64_t tm = 0;
    for (int b = 0; b < g.k; ++b) tm |= 1ull << g.tb[b];
    std::vector<uint64_t> id(D);
    std::vector<cd> x(D);
    for (uint64_t i = 0; i < N; ++i) {
        if ((i & tm) || (g.cmask && (i & g.cmask) != g.cval)) continue;
        for (int j = 0; j < D; ++j) {
            id[j] = i;
            for (int b = 0; b < g.k; ++b)
                if ((j >> b) & 1) id[j] |= 1ull << g.tb[b];
            x[j] = v[id[j]];
        }
        for (int r = 0; r < D; ++r) {
            cd y = 0;
            if (g.diag) y = cd(g.m[r].x, g.m[r].y) * x[r];
            else
                for (int c = 0; c < D; ++c) y += cd(g.m[r + (size_t)D * c].x, g.m[r + (size_t)D * c].y) * x[c];
            v[id[r]] = y;
        }
    }
}

// The backward sweep of the fused adjoint, interpreted on full vectors:
// psi = the final state, lam = H psi on entry; grad[param] as
// bsim_adjoint_gradient returns it.  Negative returns name a plan defect.
extern "C" int adjexec(int nq, const bsim_op* ops, int n, int n_params, double* psi_, double* lam_, double* grad) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, BSIM_SV, ops, n, e, true)) return rc;
    std::vector<EOp> back;
    Plan p;
    int ns = 0, pm = 0;
    if (int rc = adjoint_back_plan(nq, e, back, p, ns, pm)) return rc;
    cd* psi = (cd*)psi_;
    cd* lam = (cd*)lam_;
    const uint64_t N = 1ull << nq;
    std::vector<double> out(3 * (size_t)ns, 0.0);
    std::vector<int> seen(3 * (size_t)ns, 0);
    for (auto& st : p.steps) {
        if (!st.tile) {
            if (back[st.op].nparams) return -20;  // parametrized ops belong in tile passes
            apply_eop_cpu(back[st.op], N, psi);
            apply_eop_cpu(back[st.op], N, lam);
            continue;
        }
        const Pass& P = p.passes[st.pass];
        if (P.R != 3 || P.T > 11) return -21;  // the shape the backward tile kernel takes
        for (int s = 0; s < P.nsweeps; ++s) {
            const Sweep& S = p.sweeps[P.sweep_begin + s];
            int vb[4] = { 0, 0, 0, 0 };
            uint64_t G = 0;
            for (int r = 0; r < 3; ++r) {
                vb[r] = P.lbits[S.sb[r]];
                G |= 1ull << vb[r];
            }
            if (int rc = check_sweep(p, S, 3, vb, G, N)) return rc;
            for (int o = S.op_begin; o < S.op_end;) {
                const TOp& op = p.tops[o];
                if (op.kind == 5 || op.kind == 4) return -22;
                if (o + 1 + op.nparam > S.op_end) return -23;
                // Pauli-coded products: the head's words must name the
                // payloads that follow (one before the update, at most one
                // after), and the coded form must give the payload's value
                const bool coded = op.kind == 0 && op.cmask == 0 && op.cval != 0;
                const uint32_t words[2] = { (uint32_t)op.cval, (uint32_t)(op.cval >> 32) };  // after, before
                int named[2] = { 0, 0 };
                int rc = 0;
                auto product = [&](const TOp& d) {
                    double v = dot_top(op, d, vb, N, psi, lam);
                    if (coded) {
                        const int side = d.pre ? 1 : 0;
                        const uint32_t w = words[side];
                        ++named[side];
                        if (w == 0 || (w >> 2) != (uint32_t)(d.a * pm + d.b)) rc = -26;
                        const double sc = d.pre ? op.m[4].x : op.m[4].y;
                        const double c = sc * pauli_im(vb[op.a], (int)(w & 3u), N, psi, lam);
                        if (std::fabs(c - v) > 1e-12 * std::max(1.0, std::fabs(v))) rc = -27;
                        v = c;  // the kernel's number
                    } else if (d.pre) {
                        rc = -28;  // products before their op are coded only
                    }
                    out[3 * (size_t)d.a + d.b] += v;
                    ++seen[3 * (size_t)d.a + d.b];
                };
                for (int q = 0; q < op.nparam; ++q) {
                    const TOp& d = p.tops[o + 1 + q];
                    if (d.kind != 5 || d.a < 0 || d.a >= ns || d.b < 0 || d.b >= pm) return -24;
                }
                for (int q = 0; q < op.nparam; ++q)
                    if (p.tops[o + 1 + q].pre) product(p.tops[o + 1 + q]);
                apply_top(p, o, 3, vb, G, N, psi);
                apply_top(p, o, 3, vb, G, N, lam);
                for (int q = 0; q < op.nparam; ++q)
                    if (!p.tops[o + 1 + q].pre) product(p.tops[o + 1 + q]);
                if (rc) return rc;
                if (coded && ((words[0] != 0) != (named[0] == 1) || (words[1] != 0) != (named[1] == 1) ||
                              named[0] > 1 || named[1] > 1))
                    return -26;
                o += 1 + op.nparam;
            }
        }
    }
    for (int q = 0; q < n_params; ++q) grad[q] = 0.0;
    int slot = 0;
    for (const EOp& x : e) {
        if (x.nparams == 0) continue;
        for (int q = 0; q < x.nparams; ++q) {
            if (seen[3 * (size_t)slot + q] != 1) return -25;  // every product exactly once
            grad[x.pidx[q]] += 2.0 * out[3 * (size_t)slot + q];
        }
        ++slot;
    }
    return 0;
}

// number of tile ops of kind `tkind` in the plan (with controls only, if `ctl`)
extern "C" int plan_kind_count(int nq, int kind, const bsim_op* ops, int n, int tkind, int ctl) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, kind, ops, n, e, false)) return rc;
    Plan p;
    if (int rc = make_plan(kind == BSIM_SV ? nq : 2 * nq, e, tile_config(), p)) return rc;
    int k = 0;
    for (auto& t : p.tops) k += (t.kind == tkind && (!ctl || t.cmask)) ? 1 : 0;
    return k;
}

// register-chain statistics of the plan: out[0..4] = tile ops of kinds
// 8..12, out[5] = split-only sign ops, out[6] = the longest kind-12 chain
extern "C" int plan_chain_stats(int nq, int kind, const bsim_op* ops, int n, int* out) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, kind, ops, n, e, false)) return rc;
    Plan p;
    if (int rc = make_plan(kind == BSIM_SV ? nq : 2 * nq, e, tile_config(), p)) return rc;
    for (int i = 0; i < 7; ++i) out[i] = 0;
    for (auto& t : p.tops) {
        if (t.kind >= 8 && t.kind <= 12) ++out[t.kind - 8];
        if (t.kind == 3 && t.split_only) ++out[5];
        if (t.kind == 12 && t.nd > out[6]) out[6] = t.nd;
    }
    return 0;
}

// backward-plan statistics of the fused adjoint: out[0] = ops of the
// backward list, out[1] = merged one-qubit runs among them, out[2] = tile
// ops with Pauli-coded products, out[3] = coded products taken before
// their op, out[4] = coded products taken after
extern "C" int plan_adj_stats(int nq, const bsim_op* ops, int n, int* out) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, BSIM_SV, ops, n, e, true)) return rc;
    std::vector<EOp> back;
    std::vector<int> src;
    Plan p;
    int ns = 0, pm = 0;
    if (int rc = adjoint_back_plan(nq, e, back, p, ns, pm, nullptr, &src)) return rc;
    for (int k = 0; k < 5; ++k) out[k] = 0;
    out[0] = (int)back.size();
    for (int s : src) out[1] += s < 0 ? 1 : 0;
    for (size_t o = 0; o < p.tops.size(); ++o) {
        const TOp& t = p.tops[o];
        if (t.kind != 0 || t.cmask != 0 || t.cval == 0) continue;
        ++out[2];
        out[3] += (t.cval >> 32) ? 1 : 0;
        out[4] += (uint32_t)t.cval ? 1 : 0;
    }
    return 0;
}

// 16x16 superoperator tile ops of the plan: out[0] in orbit form (kind 4,
// b > 0), out[1] as 16 columns (kind 4, b = 0, nd = 15)
extern "C" int plan_superop_forms(int nq, int kind, const bsim_op* ops, int n, int* out) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, kind, ops, n, e, false)) return rc;
    Plan p;
    if (int rc = make_plan(kind == BSIM_SV ? nq : 2 * nq, e, tile_config(), p)) return rc;
    out[0] = out[1] = 0;
    for (const TOp& t : p.tops)
        if (t.kind == 4 && t.nd) ++out[t.b ? 0 : 1];
    return 0;
}
