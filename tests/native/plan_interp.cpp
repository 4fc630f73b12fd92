// Test helper (CPU only): plans an instruction list with libbsim's own host
// planner (lower_ops + make_plan) and interprets the plan -- every pass,
// sweep and tile op with its documented meaning on global indices -- on a
// full state vector, without tiles.  Comparing the result with the oracle
// checks the planner's lowering, ordering, fusion and index-test tables on
// the CPU.
#include <complex>
#include <cstdio>
#include <vector>
#include "fused.h"
using namespace bsim;
typedef std::complex<double> cd;
extern "C" int planexec(int nq, const bsim_op* ops, int n, double* state) {
    std::vector<EOp> e;
    if (int rc = lower_ops(nq, BSIM_SV, ops, n, e, false)) return rc;
    Plan p;
    if (int rc = make_plan(nq, e, tile_config(), p)) return rc;
    cd* v = (cd*)state;
    const uint64_t N = 1ull << nq;
    for (auto& st : p.steps) {
        if (!st.tile) { printf("non-tile step\n"); return -99; }
        const Pass& P = p.passes[st.pass];
        for (int s = 0; s < P.nsweeps; ++s) {
            const Sweep& S = p.sweeps[P.sweep_begin + s];
            int vb[4];
            for (int r = 0; r < 4; ++r) vb[r] = P.lbits[S.sb[r]];
            uint64_t G = 0;
            for (int r = 0; r < 4; ++r) G |= 1ull << vb[r];
            // the per-thread / per-register split of every index test must
            // agree with the direct test on every index
            for (int o = S.op_begin; o < S.op_end; ++o) {
                const TOp& op = p.tops[o];
                for (uint64_t i = 0; i < N; ++i) {
                    const uint64_t gbase = i & ~G;
                    int j = 0;
                    for (int r = 0; r < 4; ++r) j |= (int)((i >> vb[r]) & 1) << r;
                    if (op.kind == 3) {
                        int direct = 0, split = (int)((op.jmask >> j) & 1);
                        for (int k = 0; k < op.npairs; ++k) {
                            direct ^= __builtin_popcountll(i & (i >> op.pd[k]) & op.pm[k]) & 1;
                            split ^= __builtin_popcountll(gbase & (gbase >> op.pd[k]) & op.pm[k]) & 1;
                        }
                        for (int r = 0; r < 4; ++r)
                            if ((j >> r) & 1) split ^= __builtin_popcountll(gbase & op.sgm[r]) & 1;
                        if (direct != split) return -10;
                    } else if (op.kind <= 1 && op.cmask) {
                        const bool direct = (i & op.cmask) == op.cval;
                        const bool split = ((gbase & op.tmask) == op.tval) && ((op.jmask >> j) & 1);
                        if (direct != split) return -11;
                    }
                }
            }
            for (int o = S.op_begin; o < S.op_end; ++o) {
                const TOp& op = p.tops[o];
                auto M = [&](int j) { return cd(op.m[j].x, op.m[j].y); };
                if (op.kind == 0) {
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
        }
    }
    return 0;
}
