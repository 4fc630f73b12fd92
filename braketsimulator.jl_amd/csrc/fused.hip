// Fused evolution (evolve!, src/sv_simulator.jl:140-148 and
// src/dm_simulator.jl:170-179): the instruction stream is planned on the
// host into HBM passes.  A pass loads a tile of 2^T amplitudes -- the low
// `low_bits` vector bits (contiguous HBM runs) plus up to T-low_bits other
// bits chosen by the planner -- into LDS, applies every planned gate whose
// target bits lie in the tile, and writes the tile back: one HBM read and
// one write for many gates.  Inside a pass gates are grouped in sweeps of 4
// register bits: each thread pulls a 16-amplitude sub-cube from LDS into
// registers, applies all gates of the sweep there, and pushes it back.
// Diagonal gates and controls read the global index and need no tile bit.
#include <atomic>
#include <string.h>

#include <algorithm>
#include <complex>
#include <mutex>
#include <vector>

#include "fused.h"

namespace bsim {

// The library-wide tile configuration.  Entry points take one snapshot per
// call (tile_config()), so a call never plans with a half-updated value and
// setters on one thread do not race with planning on another.
static std::mutex g_cfg_mu;
static TileConfig g_cfg;

TileConfig tile_config() {
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    return g_cfg;
}

void set_tile_config(const TileConfig& c) {
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    g_cfg = c;
}

uint64_t EOp::touched() const {
    if (kind == E_KRAUS_REF && !kpres) return ~0ull;
    uint64_t t = g.cmask;
    for (int b = 0; b < g.k; ++b) t |= 1ull << g.tb[b];
    return t;
}

bool EOp::pair_diag() const { return kind == E_DIAG || (kind == E_KRAUS_REF && kpres); }

bool EOp::fusable() const {
    if (kind == E_DENSE) return g.k <= 2 || (g.k == 4 && tile4);
    if (kind == E_DIAG) return g.k <= 4;
    if (kind == E_KRAUS_REF) return kq <= 2 && kshift > 0;
    return false;
}

static bool is_diagonal(const double* m, int k) {
    int D = 1 << k;
    for (int c = 0; c < D; ++c)
        for (int r = 0; r < D; ++r)
            if (r != c && (m[2 * (r + D * c)] != 0.0 || m[2 * (r + D * c) + 1] != 0.0)) return false;
    return true;
}

// ---------------------------------------------------------------------------
// lowering
// ---------------------------------------------------------------------------

int lower_ops(int nq, int kind, const bsim_op* ops, int n_ops, std::vector<EOp>& out, bool keep_params) {
    const int vq = kind == BSIM_SV ? nq : 2 * nq;
    out.clear();
    out.reserve((size_t)n_ops * (kind == BSIM_SV ? 1 : 2));
    std::vector<double> tmp;
    for (int i = 0; i < n_ops; ++i) {
        const bsim_op& op = ops[i];
        int k = op.n_targets, nc = op.n_controls;
        if (k < 1 || k > BSIM_MAX_TARGETS) return fail(BSIM_EINVAL, "op: n_targets must be 1..8");
        if (nc < 0 || nc > BSIM_MAX_CONTROLS) return fail(BSIM_EINVAL, "op: n_controls must be 0..8");
        if (!op.matrix) return fail(BSIM_EINVAL, "op: matrix is null");
        for (int b = 0; b < k; ++b)
            if (op.targets[b] < 0 || op.targets[b] >= nq) return fail(BSIM_EINVAL, "op: target out of range");
        for (int c = 0; c < nc; ++c)
            if (op.controls[c] < 0 || op.controls[c] >= nq) return fail(BSIM_EINVAL, "op: control out of range");
        if (op.type == BSIM_OP_GATE || op.type == BSIM_OP_DIAG) {
            bool diag = op.type == BSIM_OP_DIAG;
            const double* m = op.matrix;
            // a gate that is diagonal at this angle runs as a diagonal op only
            // when its derivatives are diagonal too (Rx(0) = I, dRx/dtheta is not)
            bool d_diag = true;
            if (keep_params)
                for (int j = 0; j < op.n_params && j < 3; ++j)
                    if (op.dmatrix[j] && !is_diagonal(op.dmatrix[j], k)) d_diag = false;
            if (!diag && k <= 4 && d_diag && is_diagonal(m, k)) {
                int D = 1 << k;
                tmp.resize(2 * (size_t)D);
                for (int j = 0; j < D; ++j) {
                    tmp[2 * j] = m[2 * (j + D * j)];
                    tmp[2 * j + 1] = m[2 * (j + D * j) + 1];
                }
                m = tmp.data();
                diag = true;
            }
            int passes = kind == BSIM_SV ? 1 : 2;
            for (int p = 0; p < passes; ++p) {
                // DM: pass 0 = U on the column bits (vector qubit q), pass 1 =
                // conj(U) on the row bits (vector qubit nq + q), dm_simulator.jl:136-138
                int t[8], c[8];
                for (int b = 0; b < k; ++b) t[b] = op.targets[b] + (p ? nq : 0);
                for (int b = 0; b < nc; ++b) c[b] = op.controls[b] + (p ? nq : 0);
                EOp e;
                e.kind = diag ? E_DIAG : E_DENSE;
                if (int rc = make_spec(vq, k, t, nc, c, op.control_values, m, diag, p == 1, false, e.g)) return rc;
                if (keep_params && p == 0) {
                    e.nparams = op.n_params;
                    if (e.nparams < 0 || e.nparams > 3) return fail(BSIM_EINVAL, "op: n_params must be 0..3");
                    int D = 1 << k;
                    for (int j = 0; j < e.nparams; ++j) {
                        e.pidx[j] = op.param_index[j];
                        if (!op.dmatrix[j]) return fail(BSIM_EINVAL, "op: dmatrix is null");
                        e.dm[j].resize((size_t)D * D);
                        for (int x = 0; x < D * D; ++x)
                            e.dm[j][x] = make_double2(op.dmatrix[j][2 * x], op.dmatrix[j][2 * x + 1]);
                    }
                }
                out.push_back(std::move(e));
            }
        } else if (op.type == BSIM_OP_KRAUS_REF || op.type == BSIM_OP_KRAUS_PHYS || op.type == BSIM_OP_PHASEFLIP) {
            if (kind != BSIM_DM) return fail(BSIM_EINVAL, "noise ops need a density matrix");
            if (nc) return fail(BSIM_EINVAL, "noise ops take no controls");
            uint64_t used = 0;
            for (int b = 0; b < k; ++b) {
                if (used & (1ull << op.targets[b])) return fail(BSIM_EINVAL, "op: repeated target");
                used |= 1ull << op.targets[b];
            }
            if (op.type == BSIM_OP_PHASEFLIP) {
                if (k != 1) return fail(BSIM_EINVAL, "PhaseFlip acts on one qubit");
                double p = op.matrix[0];
                double same = (1.0 - p) + p, diff = (1.0 - p) - p;  // noise_kernels.jl:18-20
                double d[8] = { same, 0, diff, 0, diff, 0, same, 0 };
                int t[2] = { op.targets[0] + nq, op.targets[0] };  // (row bit, column bit)
                EOp e;
                e.kind = E_DIAG;
                if (int rc = make_spec(vq, 2, t, 0, nullptr, nullptr, d, true, false, false, e.g)) return rc;
                out.push_back(std::move(e));
            } else if (op.type == BSIM_OP_KRAUS_REF) {
                if (k > 8) return fail(BSIM_ENOTSUP, "Kraus channels act on at most 8 qubits");
                if (op.n_kraus < 1) return fail(BSIM_EINVAL, "Kraus channel without matrices");
                EOp e;
                e.kind = E_KRAUS_REF;
                e.kq = k;
                for (int b = 0; b < k; ++b) e.ktargets[b] = op.targets[b];
                e.nkraus = op.n_kraus;
                size_t D = (size_t)1 << k, n = (size_t)op.n_kraus * D * D;
                e.kraus.resize(n);
                for (size_t x = 0; x < n; ++x) e.kraus[x] = make_double2(op.matrix[2 * x], op.matrix[2 * x + 1]);
                if (k <= 2) {
                    // Tile form.  The reference updates block B[r, c] = stored[ix(r), ix(c)]
                    // (rows and columns sharing every non-target bit) to sum_m K B K^dagger,
                    // ix: matrix bit 0 = target 0 (k = 1); bit 0 = target 1, bit 1 =
                    // target 0 (k = 2, noise_kernels.jl:145-155).  On vec(B) (J = r + D c)
                    // that is S[(r, c), (l, l')] = sum_m K[r, l] conj(K[c, l']); matrix bit
                    // b is the stored-row bit of qubit q_b (vector qubit nq + q_b), bit
                    // k + b its stored-column bit (vector qubit q_b).
                    const int q[2] = { k == 1 ? op.targets[0] : op.targets[1], op.targets[0] };
                    const size_t DD = D * D;
                    std::vector<double> sup(2 * DD * DD, 0.0);
                    for (int m = 0; m < op.n_kraus; ++m) {
                        const double* K = op.matrix + 2 * (size_t)m * DD;
                        for (size_t Jo = 0; Jo < DD; ++Jo)
                            for (size_t Ji = 0; Ji < DD; ++Ji) {
                                const size_t r = Jo % D, c = Jo / D, l = Ji % D, lp = Ji / D;
                                const double ar = K[2 * (r + D * l)], ai = K[2 * (r + D * l) + 1];
                                const double br = K[2 * (c + D * lp)], bi = -K[2 * (c + D * lp) + 1];
                                sup[2 * (Jo + DD * Ji)] += ar * br - ai * bi;
                                sup[2 * (Jo + DD * Ji) + 1] += ar * bi + ai * br;
                            }
                    }
                    int t[4];
                    uint64_t rowbits = 0;
                    for (int b = 0; b < k; ++b) {
                        t[b] = nq + q[b];
                        t[k + b] = q[b];
                        rowbits |= 1ull << (nq - 1 - q[b]);
                    }
                    if (int rc = make_spec(vq, 2 * k, t, 0, nullptr, nullptr, sup.data(), false, false, false, e.g))
                        return rc;
                    e.kmask = (((1ull << nq) - 1) & ~rowbits);
                    e.kshift = nq;
                    // exact zeros only: a rounding-level link keeps the op a full barrier
                    e.kpres = true;
                    for (size_t Jo = 0; Jo < DD && e.kpres; ++Jo)
                        for (size_t Ji = 0; Ji < DD; ++Ji)
                            if ((Jo % D == Jo / D) != (Ji % D == Ji / D) &&
                                (sup[2 * (Jo + DD * Ji)] != 0.0 || sup[2 * (Jo + DD * Ji) + 1] != 0.0)) {
                                e.kpres = false;
                                break;
                            }
                }
                out.push_back(std::move(e));
            } else {
                // physical channel: superoperator sum_m K_m (column bits) (x) conj(K_m) (row bits)
                if (k > 5) return fail(BSIM_ENOTSUP, "physical Kraus channels act on at most 5 qubits");
                if (op.n_kraus < 1) return fail(BSIM_EINVAL, "Kraus channel without matrices");
                int D = 1 << k, DD = D * D;
                // channel matrix bit b -> qubit: k=1 (q); k=2 t1 = MSB (noise_kernels.jl:145-155);
                // k>2 sorted endian targets (largest qubit number first)
                int qb[8];
                if (k == 1) qb[0] = op.targets[0];
                else if (k == 2) { qb[0] = op.targets[1]; qb[1] = op.targets[0]; }
                else {
                    for (int b = 0; b < k; ++b) qb[b] = op.targets[b];
                    std::sort(qb, qb + k, [](int a, int b) { return a > b; });
                }
                std::vector<double> sup(2 * (size_t)DD * DD, 0.0);
                for (int m = 0; m < op.n_kraus; ++m) {
                    const double* K = op.matrix + 2 * (size_t)m * DD;
                    for (int Jc = 0; Jc < DD; ++Jc)
                        for (int Jr = 0; Jr < DD; ++Jr) {
                            int r = Jr % D, c = Jr / D, rp = Jc % D, cp = Jc / D;
                            // Sup[(r,c),(r',c')] += conj(K[r,r']) K[c,c']
                            double ar = K[2 * (r + D * rp)], ai = -K[2 * (r + D * rp) + 1];
                            double br = K[2 * (c + D * cp)], bi = K[2 * (c + D * cp) + 1];
                            size_t idx = 2 * ((size_t)Jr + (size_t)DD * Jc);
                            sup[idx] += ar * br - ai * bi;
                            sup[idx + 1] += ar * bi + ai * br;
                        }
                }
                int t[10];
                for (int b = 0; b < k; ++b) {
                    t[b] = qb[b] + nq;  // row bit of the channel's matrix bit b
                    t[k + b] = qb[b];   // column bit
                }
                EOp e;
                e.kind = E_DENSE;
                if (int rc = make_spec(vq, 2 * k, t, 0, nullptr, nullptr, sup.data(), false, false, false, e.g))
                    return rc;
                e.tile4 = k == 2;
                out.push_back(std::move(e));
            }
        } else if (op.type == BSIM_OP_SUPEROP) {
            // rho[r, c] <- sum S[(r, c), (r', c')] rho[r', c'], J = r + 2^k c: the
            // row index r of the physical matrix lives on vector qubits t, its
            // column index c on vector qubits nq + t (storage is rho^T)
            if (kind != BSIM_DM) return fail(BSIM_EINVAL, "superoperators need a density matrix");
            if (nc) return fail(BSIM_EINVAL, "superoperators take no controls");
            if (k > 5) return fail(BSIM_ENOTSUP, "superoperators act on at most 5 qubits");
            int t[10];
            for (int b = 0; b < k; ++b) {
                t[b] = op.targets[b];
                t[k + b] = op.targets[b] + nq;
            }
            EOp e;
            e.kind = E_DENSE;
            if (int rc = make_spec(vq, 2 * k, t, 0, nullptr, nullptr, op.matrix, false, false, false, e.g)) return rc;
            e.tile4 = k == 2;
            out.push_back(std::move(e));
        } else {
            return fail(BSIM_EINVAL, "unknown op type");
        }
    }
    return BSIM_OK;
}

// ---------------------------------------------------------------------------
// planner
// ---------------------------------------------------------------------------

static inline int popc(uint64_t x) { return __builtin_popcountll(x); }

static uint64_t target_mask(const EOp& e) {
    uint64_t t = 0;
    for (int b = 0; b < e.g.k; ++b) t |= 1ull << e.g.tb[b];
    return t;
}

static bool is_pm1(double2 z, double s) { return z.x == s && z.y == 0.0; }

// A diagonal whose entries are +1 except -1 where two bits are both set (a
// CZ in either spelling: 2-bit diag(1,1,1,-1), or Z controlled on value 1),
// or a single-bit diag(1,-1): sign = parity(i & (i >> pd) & pm).
static bool sign_pair(const EOp& e, uint64_t& pm, int& pd) {
    const GateSpec& g = e.g;
    if (g.k == 1 && is_pm1(g.m[0], 1.0) && is_pm1(g.m[1], -1.0)) {
        if (g.cmask == 0) {
            pm = 1ull << g.tb[0];
            pd = 0;
            return true;
        }
        if (popc(g.cmask) == 1 && g.cval == g.cmask) {
            int c = __builtin_ctzll(g.cmask), t = g.tb[0];
            int lo = c < t ? c : t, hi = c < t ? t : c;
            pm = 1ull << lo;
            pd = hi - lo;
            return true;
        }
        return false;
    }
    if (g.k == 2 && g.cmask == 0 && is_pm1(g.m[0], 1.0) && is_pm1(g.m[1], 1.0) && is_pm1(g.m[2], 1.0) &&
        is_pm1(g.m[3], -1.0)) {
        int lo = g.tb[0] < g.tb[1] ? g.tb[0] : g.tb[1], hi = g.tb[0] < g.tb[1] ? g.tb[1] : g.tb[0];
        pm = 1ull << lo;
        pd = hi - lo;
        return true;
    }
    return false;
}

// FP64 operations per amplitude of one op inside a tile sweep
static int op_cost(const EOp& e) {
    if (e.kind == E_DENSE) return e.g.k == 1 ? 8 : (e.g.k == 2 ? 32 : 128);
    if (e.kind == E_DIAG) {
        uint64_t pm;
        int pd;
        return sign_pair(e, pm, pd) ? 1 : 4;
    }
    return 0;
}

// ---- fusion of a sweep's ops on at most two register bits ----------------
//
// A run of 1- and 2-register ops (dense or diagonal, controls included when
// they are register bits too) that together touch at most two register bits
// is one 4x4 (or 2x2) unitary: the product of the ops' embeddings.  Applied
// as one dense op it does the arithmetic of the two 1-qubit gates it
// replaces, and absorbs the CZ-type signs and diagonals between them.

struct SweepBlock {
    uint32_t regs = 0;     // register bits (bit r = register r)
    std::vector<int> ops;  // eop indices, in program order
};

// register bits an op reads, writes or conditions on; *outside: whether it
// also touches vector bits that are not registers of the sweep
template <class Reg>
static uint32_t op_regs(const EOp& e, Reg& reg, bool* outside) {
    uint32_t m = 0;
    *outside = false;
    auto add = [&](int vbit) {
        const int r = reg(vbit);
        if (r < 0) *outside = true;
        else m |= 1u << r;
    };
    for (int b = 0; b < e.g.k; ++b) add(e.g.tb[b]);
    for (uint64_t c = e.g.cmask; c; c &= c - 1) add(__builtin_ctzll(c));
    return m;
}

// Dense 1-bit op whose matrix is exactly X (NOT, CNOT, Toffoli targets): in
// 3-register sweeps a tile op that swaps registers instead of multiplying.
static bool is_x(const EOp& e) {
    if (e.kind != E_DENSE || e.g.k != 1) return false;
    const double2* m = e.g.m.data();
    return m[0].x == 0.0 && m[0].y == 0.0 && m[1].x == 1.0 && m[1].y == 0.0 && m[2].x == 1.0 && m[2].y == 0.0 &&
           m[3].x == 0.0 && m[3].y == 0.0;
}

// VALU instructions per thread of an op as a separate sweep op on a
// 16-amplitude sub-cube (the fusion rule compares these): dense 1-bit 128
// FP64 + dispatch, 64 selects more with controls; dense 2-bit 256 FP64; a
// sign op ~80 bit operations; a general diagonal 64 FP64 + its index.
constexpr int COST_DENSE1 = 138, COST_DENSE2 = 266;
static int sweep_cost(const EOp& e) {
    uint64_t pm;
    int pd;
    if (e.kind == E_DIAG) return sign_pair(e, pm, pd) ? 90 : 130;
    const int ctl = e.g.cmask ? 64 : 0;
    return (e.g.k == 1 ? COST_DENSE1 : COST_DENSE2) + ctl;
}

// Embedding of op e into the block basis: block index bit p <-> register R[p].
template <class Reg>
static void embed(const EOp& e, Reg& reg, const int* R, int nb, std::vector<std::complex<double>>& E) {
    const int D = 1 << nb;
    E.assign((size_t)D * D, 0.0);
    int tpos[8], cpos[8], cbit[8], nc = 0;
    for (int b = 0; b < e.g.k; ++b) {
        const int r = reg(e.g.tb[b]);
        for (int p = 0; p < nb; ++p)
            if (R[p] == r) tpos[b] = p;
    }
    for (uint64_t c = e.g.cmask; c; c &= c - 1) {
        const int v = __builtin_ctzll(c), r = reg(v);
        for (int p = 0; p < nb; ++p)
            if (R[p] == r) cpos[nc] = p;
        cbit[nc++] = (int)((e.g.cval >> v) & 1);
    }
    uint32_t tmask = 0;
    for (int b = 0; b < e.g.k; ++b) tmask |= 1u << tpos[b];
    for (int c = 0; c < D; ++c)
        for (int r = 0; r < D; ++r) {
            if ((r & ~tmask) != (c & ~tmask)) continue;  // bits the op leaves alone
            bool on = true;
            for (int i = 0; i < nc; ++i) on = on && (((r >> cpos[i]) & 1) == cbit[i]);
            if (!on) {
                if (r == c) E[r + (size_t)D * c] = 1.0;
                continue;
            }
            int rs = 0, cs = 0;
            for (int b = 0; b < e.g.k; ++b) {
                rs |= ((r >> tpos[b]) & 1) << b;
                cs |= ((c >> tpos[b]) & 1) << b;
            }
            double2 z;
            if (e.kind == E_DIAG) {
                if (rs != cs) continue;
                z = e.g.m[rs];
            } else {
                z = e.g.m[rs + (1 << e.g.k) * cs];
            }
            E[r + (size_t)D * c] = std::complex<double>(z.x, z.y);
        }
}

template <class Reg, class Emit>
static void fuse_sweep(const std::vector<EOp>& eops, const std::vector<int>& take, Reg& reg, Emit& emit_one, Plan& plan) {
    std::vector<SweepBlock> pend;
    auto flush = [&](SweepBlock& b) {
        const int nb = popc(b.regs);
        int cost = 0;
        for (int i : b.ops) cost += sweep_cost(eops[i]);
        const int fused = nb == 1 ? COST_DENSE1 : COST_DENSE2;
        if (b.ops.size() < 2 || cost <= fused) {
            for (int i : b.ops) emit_one(i);
            return;
        }
        int R[2], n = 0;
        for (int r = 0; r < 4; ++r)
            if (b.regs & (1u << r)) R[n++] = r;
        const int D = 1 << nb;
        std::vector<std::complex<double>> M((size_t)D * D, 0.0), E, T;
        for (int d = 0; d < D; ++d) M[d + (size_t)D * d] = 1.0;
        for (int i : b.ops) {  // M <- E_i M
            embed(eops[i], reg, R, nb, E);
            T.assign((size_t)D * D, 0.0);
            for (int c = 0; c < D; ++c)
                for (int k = 0; k < D; ++k) {
                    const std::complex<double> mk = M[k + (size_t)D * c];
                    if (mk == 0.0) continue;
                    for (int r = 0; r < D; ++r) T[r + (size_t)D * c] += E[r + (size_t)D * k] * mk;
                }
            M.swap(T);
        }
        TOp op;
        memset(&op, 0, sizeof(op));
        op.kind = nb == 1 ? 0 : 1;
        op.a = R[0];
        op.b = nb == 2 ? R[1] : 0;
        for (int j = 0; j < D * D; ++j) op.m[j] = make_double2(M[j].real(), M[j].imag());
        plan.tops.push_back(op);
        plan.top_src.push_back(b.ops.front());
    };
    for (int i : take) {
        const EOp& e = eops[i];
        bool outside = false;
        // a reference Kraus op depends on every bit: nothing moves across it
        const uint32_t R = e.kind == E_KRAUS_REF ? 0xFu : op_regs(e, reg, &outside);
        const bool fusable =
            !outside && (e.kind == E_DENSE || e.kind == E_DIAG) && popc(R) <= 2 && R != 0 && e.aslot < 0;
        uint32_t U = R;
        for (auto& b : pend)
            if (b.regs & R) U |= b.regs;
        if (fusable && popc(U) <= 2) {
            SweepBlock nbk;
            nbk.regs = U;
            std::vector<SweepBlock> keep;
            for (auto& b : pend) {
                if (b.regs & R) nbk.ops.insert(nbk.ops.end(), b.ops.begin(), b.ops.end());
                else keep.push_back(std::move(b));
            }
            nbk.ops.push_back(i);
            keep.push_back(std::move(nbk));
            pend.swap(keep);
            continue;
        }
        std::vector<SweepBlock> keep;
        for (auto& b : pend) {
            if (b.regs & R) flush(b);
            else keep.push_back(std::move(b));
        }
        pend.swap(keep);
        if (fusable) {
            SweepBlock nbk;
            nbk.regs = R;
            nbk.ops.push_back(i);
            pend.push_back(std::move(nbk));
        } else {
            emit_one(i);
        }
    }
    for (auto& b : pend) flush(b);
}

// Merging of adjacent dense ops into one op of at most kMergeBits bits: on
// vectors of at least kMergeMinBits bits, where each launch is an HBM pass
// (28 qubits, 3- and 4-target runs: 130 -> 76 ms), and 6-bit products from
// kMerge6MinBits (on small vectors the 6-bit matrix-core launch is
// latency-bound: an 8-qubit density matrix's 3-target gates took 16x longer
// merged).
constexpr int kMergeBits = 6;
constexpr int kMergeMinBits = 20;
constexpr int kMerge6MinBits = 22;

static bool mergeable(const EOp& e) {
    return (e.kind == E_DENSE || e.kind == E_DIAG) && e.g.cmask == 0 && e.nparams == 0 && e.aslot < 0;
}

// The group's product as one dense op on the union of its bits (ascending):
// op j of the group acts on union bits pos(tb), M <- Op_j M.
static EOp merge_group(const std::vector<EOp>& eops, const std::vector<size_t>& group, uint64_t U) {
    int ub[kMergeBits], ku = 0;
    for (int bit = 0; bit < 64; ++bit)
        if ((U >> bit) & 1) ub[ku++] = bit;
    const int D = 1 << ku;
    std::vector<std::complex<double>> M((size_t)D * D, 0.0), T((size_t)D * D);
    for (int i = 0; i < D; ++i) M[i + (size_t)D * i] = 1.0;
    for (size_t gi : group) {
        const GateSpec& g = eops[gi].g;
        int pos[16];
        uint64_t om = 0;  // union bits the op acts on
        for (int b = 0; b < g.k; ++b) {
            for (int u = 0; u < ku; ++u)
                if (ub[u] == g.tb[b]) pos[b] = u;
            om |= 1ull << pos[b];
        }
        const int d = 1 << g.k;
        auto sub = [&](int x) {  // the op's matrix index of union index x
            int s = 0;
            for (int b = 0; b < g.k; ++b) s |= ((x >> pos[b]) & 1) << b;
            return s;
        };
        // T = Op M: Op[r][c] is nonzero only where r and c agree off the op's bits
        for (int r = 0; r < D; ++r)
            for (int c = 0; c < D; ++c) {
                std::complex<double> acc = 0.0;
                const int rs = sub(r);
                if (g.diag) {
                    acc = std::complex<double>(g.m[rs].x, g.m[rs].y) * M[r + (size_t)D * c];
                } else {
                    for (int cs = 0; cs < d; ++cs) {
                        int k2 = r & ~(int)om;
                        for (int b = 0; b < g.k; ++b) k2 |= ((cs >> b) & 1) << pos[b];
                        const double2 o = g.m[rs + (size_t)d * cs];
                        acc += std::complex<double>(o.x, o.y) * M[k2 + (size_t)D * c];
                    }
                }
                T[r + (size_t)D * c] = acc;
            }
        M.swap(T);
    }
    EOp e;
    e.kind = E_DENSE;
    e.g.k = ku;
    for (int u = 0; u < ku; ++u) e.g.tb[u] = ub[u];
    e.g.m.resize((size_t)D * D);
    for (size_t i = 0; i < M.size(); ++i) e.g.m[i] = make_double2(M[i].real(), M[i].imag());
    return e;
}

// Ops held back while a pass (or a sweep) takes later ops ahead of them.  A
// later op may move ahead only if it commutes with every held-back op: its
// bits are disjoint from theirs, and -- since a reference Kraus op's
// predicate reads every other row/column pair -- a Kraus op passes only
// pair-keeping ops (pair_diag), and only pair-keeping ops pass a Kraus op.
struct HeldBack {
    uint64_t bits = 0;
    bool kraus = false;   // a kpres Kraus op is held back
    bool mixing = false;  // an op that is not pair_diag is held back
    bool blocks(const EOp& e) const {
        if (e.touched() & bits) return true;
        if (e.kind == E_KRAUS_REF && mixing) return true;
        return kraus && !e.pair_diag();
    }
    void add(const EOp& e) {
        bits |= e.touched();
        kraus |= e.kind == E_KRAUS_REF;
        mixing |= !e.pair_diag();
    }
};

// ---- register chains -------------------------------------------------------
//
// A sign op of a sweep factors, for register j of a sub-cube with base g,
// into (-1)^P(g) (one sign per thread), the uniform pattern jmask, and a
// per-thread Z on each register bit r, Z_r^t with t = parity(g & sgm[r]).
// When jmask is affine in the register bits (no CZ between two register
// bits; its linear terms are uniform Z_r), every factor acts on one register
// bit or on the whole sub-cube, so the sweep splits into one chain per
// register bit -- its 1-register dense ops and Z factors in order -- and the
// chains commute with each other; a 2-register dense op closes the chains of
// its two bits.  A chain
//   A = Z^t_end U_n Z^t_(n-1) ... U_1 Z^t_0
// is multiplied per thread (kinds 8, 9: a sign flip of a row or column and
// 32 FP64 instructions per further U) and applied once (kind 10: 128), where
// the unchained sweep applies every U to the sub-cube (128 each) and every
// sign op (about 100 VALU instructions each).  The per-thread sign P of all
// sign ops rides on one chain (kind 11).
using cd = std::complex<double>;

namespace {
struct ChainItem {
    bool u;          // a uniform 2x2 (column-major) or a per-thread Z
    cd m[4];
    uint64_t mask;   // Z: t = parity(gbase & mask)
};
}  // namespace

static void mat2_mul(const cd* a, const cd* b, cd* c) {  // c = a b, column-major
    cd t[4];
    t[0] = a[0] * b[0] + a[2] * b[1];
    t[1] = a[1] * b[0] + a[3] * b[1];
    t[2] = a[0] * b[2] + a[2] * b[3];
    t[3] = a[1] * b[2] + a[3] * b[3];
    for (int i = 0; i < 4; ++i) c[i] = t[i];
}

static TOp chain_top(int kind, int a, const cd* m, uint64_t mask) {
    TOp op;
    memset(&op, 0, sizeof(op));
    op.kind = kind;
    op.a = a;
    op.cmask = mask;
    if (m)
        for (int i = 0; i < 4; ++i) op.m[i] = make_double2(m[i].real(), m[i].imag());
    return op;
}

// VALU instructions per thread of a sweep's tile ops (the chain rule's cost
// comparison): dense 1- and 2-register ops, sign ops, chain ops.
static int sweep_valu(const TOp* t, size_t n) {
    int c = 0;
    for (size_t i = 0; i < n; ++i) {
        switch (t[i].kind) {
            case 0: c += 150; break;
            case 1: c += 290; break;
            case 3: c += 100; break;
            case 8: c += 20; break;
            case 9: c += 45; break;
            case 10: c += 145; break;
            case 11: c += 20; break;
            case 12: c += 150 + 45 * (t[i].nd - 1); break;
            default: c += 150; break;
        }
    }
    return c;
}

// Rewrite one sweep's tile ops [in, in + n) into chains; false (out
// untouched) when the sweep does not qualify or the chains cost more.
static bool chain_sweep(const TOp* in, size_t n, std::vector<TOp>& out) {
    std::vector<ChainItem> ch[4];
    std::vector<TOp> res;
    uint64_t rsgm[4] = { 0, 0, 0, 0 };
    // the per-thread sign of all sign ops: pair masks by pair distance (a
    // sign op takes four distances; more go to further sign ops)
    std::vector<std::pair<int, uint64_t>> pp;
    bool any_sign = false;
    const cd Zm[4] = { 1.0, 0.0, 0.0, -1.0 };
    auto addP = [&](uint64_t pm, int pd) {
        for (auto& q : pp)
            if (q.first == pd) {
                q.second ^= pm;
                return;
            }
        pp.push_back({ pd, pm });
    };
    // emit the chain of register r; at_end: a chain without a dense op (Z
    // factors only) joins the sweep's closing sign op, otherwise (a 2-register
    // op follows on r) it is a sign op of its own here
    auto flush = [&](int r, bool at_end) {
        std::vector<ChainItem> v;  // normalized: adjacent U multiplied, adjacent Z combined
        for (const ChainItem& it : ch[r]) {
            if (it.u) {
                if (!v.empty() && v.back().u) mat2_mul(it.m, v.back().m, v.back().m);
                else v.push_back(it);
            } else {
                if (!v.empty() && !v.back().u) {
                    v.back().mask ^= it.mask;
                    if (!v.back().mask) v.pop_back();
                } else if (it.mask) {
                    v.push_back(it);
                }
                // a zero mask between two dense ops lets them multiply
                if (v.size() >= 2 && v.back().u && v[v.size() - 2].u) {
                    mat2_mul(v.back().m, v[v.size() - 2].m, v[v.size() - 2].m);
                    v.pop_back();
                }
            }
        }
        ch[r].clear();
        bool has_u = false;
        for (const ChainItem& it : v) has_u = has_u || it.u;
        if (!has_u) {
            if (v.empty()) return;
            if (at_end) {
                rsgm[r] ^= v[0].mask;
            } else {
                TOp s;
                memset(&s, 0, sizeof(s));
                s.kind = 3;
                s.split_only = 1;
                s.sgm[r] = v[0].mask;
                res.push_back(s);
            }
            return;
        }
        size_t i = 0;
        uint64_t z0 = 0;
        if (!v[0].u) z0 = v[i++].mask;
        if (v.size() - i == 1 && !z0) {  // one dense op, no Z: a plain dense op
            TOp op = chain_top(0, r, v[i].m, 0);
            res.push_back(op);
            return;
        }
        int nu = 0;
        for (const ChainItem& it : v) nu += it.u ? 1 : 0;
        if (nu <= 4) {  // the whole chain in one op (kind 12)
            TOp op = chain_top(12, r, v[i].m, z0);
            op.nd = nu;
            int k = 0;
            for (++i; i < v.size(); ++i) {
                if (!v[i].u) {
                    op.sgm[i + 1 < v.size() ? k : 3] = v[i].mask;  // between U_(k+1) and U_(k+2), or after the last
                    continue;
                }
                ++k;
                for (int e = 0; e < 4; ++e) op.m[4 * k + e] = make_double2(v[i].m[e].real(), v[i].m[e].imag());
            }
            res.push_back(op);
            return;
        }
        res.push_back(chain_top(8, r, v[i].m, z0));
        ++i;
        uint64_t zend = 0;
        while (i < v.size()) {
            const uint64_t z = v[i++].mask;  // v alternates: a Z, then a U or the end
            if (i < v.size()) {
                res.push_back(chain_top(9, r, v[i].m, z));
                ++i;
            } else {
                zend = z;
            }
        }
        res.push_back(chain_top(10, r, nullptr, zend));
    };
    for (size_t o = 0; o < n; ++o) {
        const TOp& t = in[o];
        if (t.nparam) return false;
        if (t.kind == 0 && t.cmask == 0) {
            ChainItem it{ true, {}, 0 };
            for (int i = 0; i < 4; ++i) it.m[i] = cd(t.m[i].x, t.m[i].y);
            ch[t.a].push_back(it);
        } else if (t.kind == 1 && t.cmask == 0) {
            flush(t.a, false);
            flush(t.b, false);
            res.push_back(t);
        } else if (t.kind == 3) {
            any_sign = true;
            // jmask affine in the register bits: bit j = XOR of c_r over the set bits r of j
            const uint32_t jm = t.jmask;
            if (jm & 1u) return false;
            int c[4];
            for (int r = 0; r < 4; ++r) c[r] = (jm >> (1u << r)) & 1u;
            for (int j = 0; j < 16; ++j) {
                int b = 0;
                for (int r = 0; r < 4; ++r)
                    if ((j >> r) & 1) b ^= c[r];
                if (((jm >> j) & 1u) != (uint32_t)b) return false;
            }
            for (int r = 0; r < 4; ++r) {
                if (c[r]) ch[r].push_back(ChainItem{ true, { Zm[0], Zm[1], Zm[2], Zm[3] }, 0 });
                if (t.sgm[r]) ch[r].push_back(ChainItem{ false, {}, t.sgm[r] });
            }
            for (int k = 0; k < t.npairs; ++k)
                if (t.pm[k]) addP(t.pm[k], t.pd[k]);
        } else {
            return false;
        }
    }
    if (!any_sign) return false;  // nothing to gain: the planner's own fusion stands
    for (int r = 0; r < 4; ++r) flush(r, true);
    {
        std::vector<std::pair<int, uint64_t>> live;
        for (const auto& q : pp)
            if (q.second) live.push_back(q);
        pp.swap(live);
    }
    // the first four pair distances ride on the first chain; the sign is a
    // factor of the whole sub-cube, so it commutes with every op of the sweep
    size_t pk = 0;
    if (!pp.empty()) {
        for (size_t i = 0; i < res.size(); ++i) {
            if (res[i].kind == 12) {
                for (; pk < pp.size() && pk < 4; ++pk) {
                    res[i].pd[pk] = pp[pk].first;
                    res[i].pm[pk] = pp[pk].second;
                }
                res[i].npairs = (int)pk;
                break;
            } else if (res[i].kind == 8) {
                TOp p;
                memset(&p, 0, sizeof(p));
                p.kind = 11;
                for (; pk < pp.size() && pk < 4; ++pk) {
                    p.nd |= (pp[pk].first & 0xFF) << (8 * pk);
                    reinterpret_cast<uint64_t*>(p.m)[pk] = pp[pk].second;
                }
                p.cmask = (uint64_t)pk << 32;
                res.insert(res.begin() + (long)i + 1, p);
                break;
            }
        }
    }
    // what no chain carries: sign ops at the end of the sweep (split form
    // only), the Z factors of registers without a dense op in the first
    bool first = true;
    while (pk < pp.size() || (first && (rsgm[0] || rsgm[1] || rsgm[2] || rsgm[3]))) {
        TOp s;
        memset(&s, 0, sizeof(s));
        s.kind = 3;
        s.split_only = 1;
        int k = 0;
        for (; pk < pp.size() && k < 4; ++pk, ++k) {
            s.pd[k] = pp[pk].first;
            s.pm[k] = pp[pk].second;
        }
        s.npairs = k;
        if (first)
            for (int r = 0; r < 4; ++r) s.sgm[r] = rsgm[r];
        first = false;
        res.push_back(s);
    }
    if (sweep_valu(res.data(), res.size()) >= sweep_valu(in, n)) return false;
    out.insert(out.end(), res.begin(), res.end());
    return true;
}

// Chains for every sweep of pass P when the pass runs on a lean streaming
// kernel (12 tile bits, 4 register bits, only uncontrolled dense 1-/2-register
// ops and sign ops, plus reference Kraus ops in sweeps left as they are).
static void chain_pass(Plan& plan, const Pass& P) {
    const int ob = plan.sweeps[P.sweep_begin].op_begin;
    const int oe = plan.sweeps[P.sweep_begin + P.nsweeps - 1].op_end;
    for (int o = ob; o < oe; ++o) {
        const TOp& t = plan.tops[o];
        if (t.nparam) return;
        if (t.kind == 4) {
            o += t.nd;
            continue;
        }
        if (t.kind != 3 && ((t.kind != 0 && t.kind != 1) || t.cmask != 0)) return;
    }
    std::vector<TOp> tops(plan.tops.begin(), plan.tops.begin() + ob);
    std::vector<int> src(plan.top_src.begin(), plan.top_src.begin() + ob);
    bool changed = false;
    for (int s = P.sweep_begin; s < P.sweep_begin + P.nsweeps; ++s) {
        Sweep& sw = plan.sweeps[s];
        const int b = (int)tops.size();
        if (chain_sweep(plan.tops.data() + sw.op_begin, (size_t)(sw.op_end - sw.op_begin), tops)) {
            changed = true;
            src.insert(src.end(), tops.size() - b, plan.top_src[sw.op_begin]);
        } else {
            tops.insert(tops.end(), plan.tops.begin() + sw.op_begin, plan.tops.begin() + sw.op_end);
            src.insert(src.end(), plan.top_src.begin() + sw.op_begin, plan.top_src.begin() + sw.op_end);
        }
        sw.op_begin = b;
        sw.op_end = (int)tops.size();
    }
    if (!changed) return;
    plan.tops.swap(tops);
    plan.top_src.swap(src);
}

static int make_plan_ops(int vq, const std::vector<EOp>& eops, const TileConfig& cfg, Plan& plan) {
    plan = Plan();
    const size_t n = eops.size();
    int T = std::min(cfg.tile_bits, vq);
    // keep two free tile bits so that any 2-target gate fits a pass
    int L = std::max(0, std::min(cfg.low_bits, T - 2));
    const uint64_t all = vq >= 64 ? ~0ull : ((1ull << vq) - 1);
    if (T < 4) {
        for (size_t i = 0; i < n; ++i) plan.steps.push_back({ false, -1, (int)i });
        return BSIM_OK;
    }
    std::vector<char> done(n, 0);
    size_t head = 0;
    std::vector<int> inc;
    while (true) {
        while (head < n && done[head]) ++head;
        if (head >= n) break;
        if (!eops[head].fusable()) {
            // a dense op on 3+ bits: take the adjacent dense ops that either
            // fit inside its bits or are themselves 3+-bit ops, while the
            // union stays within kMergeBits, into one product
            const int max_bits = std::min(cfg.merge_bits, vq >= kMerge6MinBits ? kMergeBits : kMergeBits - 1);
            if (cfg.merge_dense && max_bits >= 3 && vq >= kMergeMinBits && eops[head].kind == E_DENSE &&
                mergeable(eops[head]) &&
                eops[head].g.k <= max_bits) {
                std::vector<size_t> group{ head };
                uint64_t U = target_mask(eops[head]);
                for (size_t j = head + 1; j < n; ++j) {
                    if (done[j]) continue;  // taken ahead of head by an earlier pass
                    const EOp& e = eops[j];
                    if (!mergeable(e)) break;
                    const uint64_t t = target_mask(e);
                    if (popc(U | t) > max_bits || (e.fusable() && (t & ~U))) break;
                    group.push_back(j);
                    U |= t;
                }
                if (group.size() > 1) {
                    plan.merged.push_back(merge_group(eops, group, U));
                    Plan::Step st{ false, -1, (int)head };
                    st.merged = (int)plan.merged.size() - 1;
                    plan.steps.push_back(st);
                    for (size_t j : group) done[j] = 1;
                    continue;
                }
            }
            plan.steps.push_back({ false, -1, (int)head });
            done[head] = 1;
            continue;
        }
        // ---- pass: choose the tile bits and the ops it absorbs (the low bits
        // give way when the first op needs more than T - L bits: a 2-qubit
        // reference Kraus op takes 4)
        uint64_t S = (L > 0) ? ((1ull << L) - 1) : 0;
        if ((eops[head].kind == E_DENSE || eops[head].kind == E_KRAUS_REF) && popc(S | target_mask(eops[head])) > T)
            S = 0;
        HeldBack held;
        int cost = 0;
        inc.clear();
        for (size_t i = head; i < n; ++i) {
            if (done[i]) continue;
            const EOp& e = eops[i];
            int c = op_cost(e);
            bool over = cfg.max_cost > 0 && !inc.empty() && cost + c > cfg.max_cost;
            if (held.blocks(e)) {
                held.add(e);
            } else if (!e.fusable() || over) {
                held.add(e);
            } else if (e.kind == E_DENSE || e.kind == E_KRAUS_REF) {
                uint64_t need = target_mask(e);
                if ((need & ~S) == 0 || popc(S | need) <= T) {
                    S |= need;
                    inc.push_back((int)i);
                    cost += c;
                } else {
                    held.add(e);
                }
            } else {
                inc.push_back((int)i);
                cost += c;
            }
            if ((held.bits & all) == all) break;
        }
        if (inc.empty()) {  // cannot happen (the head op always fits); never loop without progress
            plan.steps.push_back({ false, -1, (int)head });
            done[head] = 1;
            continue;
        }
        // fill the tile with the lowest free bits (longer contiguous runs)
        for (int b = 0; b < vq && popc(S) < T; ++b) S |= 1ull << b;
        Pass P;
        memset(&P, 0, sizeof(P));
        P.T = T;
        P.lmask = S;
        int pos[64];
        for (int b = 0; b < 64; ++b) pos[b] = -1;  // -1: not a tile bit
        for (int b = 0, l = 0; b < vq; ++b)
            if (S & (1ull << b)) {
                P.lbits[l] = b;
                pos[b] = l++;
            }
        P.sweep_begin = (int)plan.sweeps.size();
        // register bits per sweep: 4 when the pass holds a 4-bit op or a
        // reference Kraus op (the Kraus tile ops exist for 4 register bits only)
        int RB = cfg.reg_bits == 3 ? 3 : 4;
        for (int i : inc)
            if (eops[i].kind == E_KRAUS_REF || (eops[i].kind == E_DENSE && eops[i].g.k > 2)) RB = 4;
        if (T < RB + 2) RB = 4;  // tiny tiles: the 4-bit layout covers them
        P.R = RB;
        const int NRJ = 1 << RB;
        // ---- sweeps: groups of ops whose dense targets fit RB register bits
        std::vector<int> rem = inc;
        std::vector<int> next;
        while (!rem.empty()) {
            uint32_t W = 0;  // local bits
            HeldBack blk;
            std::vector<int> take;
            next.clear();
            for (int i : rem) {
                const EOp& e = eops[i];
                if (blk.blocks(e)) {
                    blk.add(e);
                    next.push_back(i);
                    continue;
                }
                if (e.kind == E_DENSE || e.kind == E_KRAUS_REF) {
                    uint32_t need = 0;
                    for (int b = 0; b < e.g.k; ++b) need |= 1u << pos[e.g.tb[b]];
                    if (popc(W | need) <= RB) {
                        W |= need;
                        take.push_back(i);
                    } else {
                        blk.add(e);
                        next.push_back(i);
                    }
                } else {
                    take.push_back(i);
                }
            }
            for (int b = 0; b < T && popc(W) < RB; ++b) W |= 1u << b;
            Sweep sw;
            memset(&sw, 0, sizeof(sw));
            int nb = 0;
            for (int b = 0; b < T; ++b)
                if (W & (1u << b)) sw.sb[nb++] = b;
            sw.op_begin = (int)plan.tops.size();
            auto reg = [&](int vbit) {
                const int l = pos[vbit];
                if (l < 0) return -1;
                for (int r = 0; r < RB; ++r)
                    if (sw.sb[r] == l) return r;
                return -1;
            };
            auto emit_one = [&](int i) {
                const EOp& e = eops[i];
                // diagonal index of register j (kind 6)
                auto diag_index = [&](int j) {
                    int idx = 0;
                    for (int b = 0; b < e.g.k; ++b) idx |= ((j >> reg(e.g.tb[b])) & 1) << b;
                    return idx;
                };
                TOp op;
                memset(&op, 0, sizeof(op));
                op.cmask = e.g.cmask;
                op.cval = e.g.cval;
                uint64_t pm = 0;
                int pd = 0;
                if (e.kind == E_DIAG && e.aslot < 0 && sign_pair(e, pm, pd)) {
                    // consecutive sign ops multiply: their parities add modulo 2
                    // (pair masks of the same bit distance add, others are kept
                    // as separate terms, up to 4)
                    if (plan.tops.size() > (size_t)sw.op_begin && plan.tops.back().kind == 3) {
                        TOp& prev = plan.tops.back();
                        for (int k = 0; k < prev.npairs; ++k)
                            if (prev.pd[k] == pd) {
                                prev.pm[k] ^= pm;
                                return;
                            }
                        if (prev.npairs < 4) {
                            prev.pm[prev.npairs] = pm;
                            prev.pd[prev.npairs] = pd;
                            ++prev.npairs;
                            return;
                        }
                    }
                    op.kind = 3;
                    op.pm[0] = pm;
                    op.pd[0] = pd;
                    op.npairs = 1;
                } else if (e.kind == E_KRAUS_REF || (e.kind == E_DENSE && e.g.k == 4)) {
                    // reference Kraus block update, or (mask 0: every block) a
                    // 4-bit dense superoperator
                    plan.has_kraus = true;
                    op.kind = 4;
                    op.pm[0] = e.kind == E_KRAUS_REF ? e.kmask : 0;
                    op.pd[0] = e.kind == E_KRAUS_REF ? e.kshift : 0;
                    if (e.g.k == 2) {
                        int a = reg(e.g.tb[0]), b = reg(e.g.tb[1]);
                        auto sw2 = [](int x) { return ((x & 1) << 1) | ((x >> 1) & 1); };
                        op.a = a < b ? a : b;
                        op.b = a < b ? b : a;
                        for (int r = 0; r < 4; ++r)
                            for (int c = 0; c < 4; ++c)
                                op.m[r + 4 * c] = a < b ? e.g.m[r + 4 * c] : e.g.m[sw2(r) + 4 * sw2(c)];
                    } else {
                        // 16x16 in register order: register index j (bit r = register r)
                        // <-> matrix index J (bit b = matrix bit b on register rb[b])
                        int rb[4];
                        for (int b = 0; b < 4; ++b) rb[b] = reg(e.g.tb[b]);
                        auto J = [&](int j) {
                            int x = 0;
                            for (int b = 0; b < 4; ++b) x |= ((j >> rb[b]) & 1) << b;
                            return x;
                        };
                        op.a = 0;
                        op.b = 0;
                        op.nd = 15;
                        TOp col[16];
                        for (int c = 0; c < 16; ++c) {
                            memset(&col[c], 0, sizeof(TOp));
                            col[c].kind = 5;
                            for (int r = 0; r < 16; ++r) col[c].m[r] = e.g.m[J(r) + 16 * J(c)];
                        }
                        // A physical superoperator with entries only where the row
                        // and column registers differ by an element of D = {0, P1,
                        // P2, P1 ^ P2} (P1, P2: a pairing of the four registers, e.g.
                        // a qubit's row and column bit; Pauli channels have this
                        // form) is four 4x4 blocks on the orbits o ^ D: b = 1 +
                        // pairing, the head and 3 TOps hold the blocks (orbit k,
                        // representative the k-th register with the lowest bits of
                        // P1 and P2 clear; m[r + 4 c] couples o ^ X[r] to o ^ X[c],
                        // X = {0, P1, P2, P1 ^ P2}).
                        static const int kP1[3] = { 0x3, 0x5, 0x9 }, kP2[3] = { 0xC, 0xA, 0x6 };
                        for (int pi = 0; pi < 3 && e.kind == E_DENSE; ++pi) {
                            const int P1 = kP1[pi], P2 = kP2[pi], X[4] = { 0, P1, P2, P1 ^ P2 };
                            bool form = true;
                            for (int r = 0; r < 16 && form; ++r)
                                for (int c = 0; c < 16 && form; ++c) {
                                    const int d = r ^ c;
                                    const bool in = d == X[0] || d == X[1] || d == X[2] || d == X[3];
                                    form = in || (col[c].m[r].x == 0.0 && col[c].m[r].y == 0.0);
                                }
                            if (!form) continue;
                            const int a1 = __builtin_ctz(P1), a2 = __builtin_ctz(P2);
                            TOp blk[4];
                            int k = 0;
                            for (int o = 0; o < 16; ++o) {
                                if ((o >> a1) & 1 || (o >> a2) & 1) continue;
                                memset(&blk[k], 0, sizeof(TOp));
                                blk[k].kind = 5;
                                for (int r = 0; r < 4; ++r)
                                    for (int c = 0; c < 4; ++c) blk[k].m[r + 4 * c] = col[o ^ X[c]].m[o ^ X[r]];
                                ++k;
                            }
                            op.b = 1 + pi;
                            op.nd = 3;
                            for (int r = 0; r < 16; ++r) op.m[r] = blk[0].m[r];
                            plan.tops.push_back(op);
                            plan.top_src.push_back(i);
                            for (int q = 1; q < 4; ++q) {
                                plan.tops.push_back(blk[q]);
                                plan.top_src.push_back(i);
                            }
                            return;
                        }
                        for (int r = 0; r < 16; ++r) op.m[r] = col[0].m[r];
                        plan.tops.push_back(op);
                        plan.top_src.push_back(i);
                        for (int c = 1; c < 16; ++c) {
                            plan.tops.push_back(col[c]);
                            plan.top_src.push_back(i);
                        }
                        return;
                    }
                } else if (e.kind == E_DIAG) {
                    bool inreg = true;
                    for (int b = 0; b < e.g.k; ++b) inreg = inreg && reg(e.g.tb[b]) >= 0;
                    if (inreg) {
                        // every diagonal bit is a register bit: the factor of each
                        // register is fixed for the sweep
                        op.kind = 6;
                        for (int j = 0; j < NRJ; ++j) op.m[j] = e.g.m[diag_index(j)];
                    } else {
                        op.kind = 2;
                        op.nd = e.g.k;
                        for (int b = 0; b < e.g.k; ++b) op.dbit[b] = e.g.tb[b];
                        for (int j = 0; j < (1 << e.g.k); ++j) op.m[j] = e.g.m[j];
                    }
                } else {
                    if (e.g.k == 1) {
                        // X as swaps in 3-register sweeps (the 4-register kernel keeps
                        // its op dispatch lean: it runs X as a dense op)
                        op.kind = RB == 3 && is_x(e) && !(e.aslot >= 0 && e.nparams > 0) ? 7 : 0;
                        op.a = reg(e.g.tb[0]);
                        for (int j = 0; j < 4; ++j) op.m[j] = e.g.m[j];
                    } else {
                        op.kind = 1;
                        int a = reg(e.g.tb[0]), b = reg(e.g.tb[1]);
                        if (a < b) {
                            op.a = a;
                            op.b = b;
                            for (int j = 0; j < 16; ++j) op.m[j] = e.g.m[j];
                        } else {
                            op.a = b;
                            op.b = a;
                            auto sw2 = [](int x) { return ((x & 1) << 1) | ((x >> 1) & 1); };
                            for (int r = 0; r < 4; ++r)
                                for (int c = 0; c < 4; ++c) op.m[r + 4 * c] = e.g.m[sw2(r) + 4 * sw2(c)];
                        }
                    }
                }
                if (e.aslot >= 0 && e.nparams > 0) {
                    // fused adjoint: derivative matrices, laid out like the op's own matrix
                    op.nparam = e.nparams;
                    plan.tops.push_back(op);
                    plan.top_src.push_back(i);
                    for (int p = 0; p < e.nparams; ++p) {
                        TOp d;
                        memset(&d, 0, sizeof(d));
                        d.kind = 5;
                        d.a = e.pslot[p] >= 0 ? e.pslot[p] : e.aslot;
                        d.b = e.pslot[p] >= 0 ? e.pprm[p] : p;
                        d.pre = e.ppre[p] ? 1 : 0;
                        const int D = 1 << e.g.k;
                        if (op.kind == 2) {
                            for (int j = 0; j < D; ++j) d.m[j] = e.dm[p][j + D * j];
                        } else if (op.kind == 6) {
                            for (int j = 0; j < NRJ; ++j) d.m[j] = e.dm[p][diag_index(j) * (D + 1)];
                        } else if (op.kind == 0) {
                            for (int j = 0; j < 4; ++j) d.m[j] = e.dm[p][j];
                        } else {
                            const int a0 = reg(e.g.tb[0]), b0 = reg(e.g.tb[1]);
                            auto sw2 = [](int x) { return ((x & 1) << 1) | ((x >> 1) & 1); };
                            for (int r = 0; r < 4; ++r)
                                for (int c = 0; c < 4; ++c)
                                    d.m[r + 4 * c] = a0 < b0 ? e.dm[p][r + 4 * c] : e.dm[p][sw2(r) + 4 * sw2(c)];
                        }
                        plan.tops.push_back(d);
                        plan.top_src.push_back(i);
                    }
                    return;
                }
                plan.tops.push_back(op);
                plan.top_src.push_back(i);
            };
            fuse_sweep(eops, take, reg, emit_one, plan);
            sw.op_end = (int)plan.tops.size();
            // index tests split into a per-thread part and a part fixed by the register
            uint64_t gb[4] = { 0, 0, 0, 0 }, G = 0;
            for (int r = 0; r < RB; ++r) {
                gb[r] = 1ull << P.lbits[sw.sb[r]];
                G |= gb[r];
            }
            auto rj = [&](int j) {
                uint64_t x = 0;
                for (int r = 0; r < 4; ++r)
                    if ((j >> r) & 1) x |= gb[r];
                return x;
            };
            for (int o = sw.op_begin; o < sw.op_end; ++o) {
                TOp& op = plan.tops[o];
                if (op.kind == 4) {
                    // Q is linear over GF(2): Q(gbase | r_j) = Q(gbase) ^ sum of Q(register bits)
                    for (int r = 0; r < 4; ++r) op.sgm[r] = (gb[r] ^ (gb[r] >> op.pd[0])) & op.pm[0];
                } else if (op.kind == 3) {
                    uint32_t q = 0;
                    for (int j = 0; j < NRJ; ++j) {
                        const uint64_t x = rj(j);
                        int par = 0;
                        for (int k = 0; k < op.npairs; ++k) par ^= popc(x & (x >> op.pd[k]) & op.pm[k]) & 1;
                        q |= (uint32_t)par << j;
                    }
                    op.jmask = q;
                    for (int r = 0; r < 4; ++r) {
                        op.sgm[r] = 0;
                        for (int k = 0; k < op.npairs; ++k)
                            op.sgm[r] ^= ((gb[r] >> op.pd[k]) & op.pm[k]) ^ ((gb[r] & op.pm[k]) << op.pd[k]);
                    }
                } else if (op.cmask) {
                    op.tmask = op.cmask & ~G;
                    op.tval = op.cval & ~G;
                    uint32_t q = 0;
                    for (int j = 0; j < NRJ; ++j) q |= (uint32_t)((rj(j) & op.cmask) == (op.cval & G)) << j;
                    op.jmask = q;
                }
            }
            plan.sweeps.push_back(sw);
            plan.fused_ops += (int64_t)take.size();
            rem.swap(next);
        }
        P.nsweeps = (int)plan.sweeps.size() - P.sweep_begin;
        // the shape the streaming kernels take (stream_kind below)
        if (cfg.chains && P.T == 12 && P.R == 4 && vq >= 12 && P.nsweeps > 0) chain_pass(plan, P);
        for (int i : inc) done[i] = 1;
        plan.steps.push_back({ true, (int)plan.passes.size(), -1 });
        plan.passes.push_back(P);
    }
    // Sign ops carry a copy of their parameters in the fields every op's
    // head load reads (nd: the pair distances, a byte each; cmask: jmask and,
    // from bit 32, the pair count; m[0..3]: the pair masks, then the
    // register-bit masks), so the kernel reads a sign op in one scalar round
    // trip.  Kernels that read the named fields keep working.
    for (TOp& op : plan.tops)
        if (op.kind == 3) {
            op.nd = (op.pd[0] & 0xFF) | (op.pd[1] & 0xFF) << 8 | (op.pd[2] & 0xFF) << 16 | (op.pd[3] & 0xFF) << 24;
            op.cmask = (uint64_t)op.jmask | ((uint64_t)op.npairs << 32);
            const uint64_t w[8] = { op.pm[0], op.pm[1], op.pm[2], op.pm[3], op.sgm[0], op.sgm[1], op.sgm[2], op.sgm[3] };
            memcpy(op.m, w, sizeof(w));
        }
    // Forms the adjoint tile body has cheaper paths for, in fields its dense
    // and register-diagonal ops do not otherwise use (nd): an uncontrolled
    // dense 1-register op whose matrix is real (nd = 1: RY, H, X ...); the
    // product payloads U^dagger dU of such ops that are real to rounding
    // (nd = 1: RY's -i Y / 2), and of uncontrolled register diagonals that are
    // imaginary to rounding (nd = 2: RZ's -i Z / 2).  A payload entry below
    // 1e-15 of the largest is rounding of an exact zero; the cheap path
    // leaves it out.
    auto small = [](double x, double scale) { return std::fabs(x) <= 1e-15 * scale; };
    for (size_t o = 0; o < plan.tops.size(); ++o) {
        TOp& op = plan.tops[o];
        if ((op.kind != 0 && op.kind != 6) || op.cmask != 0) continue;
        const int nm = op.kind == 0 ? 4 : 16;  // register diagonals: unused entries are zero
        if (op.kind == 0) {
            bool real = true;
            for (int j = 0; j < 4; ++j) real = real && op.m[j].y == 0.0;
            if (real) op.nd = 1;
        }
        for (int p = 0; p < op.nparam; ++p) {
            TOp& d = plan.tops[o + 1 + p];
            double scale = 0.0;
            for (int j = 0; j < nm; ++j) scale = std::max(scale, std::max(std::fabs(d.m[j].x), std::fabs(d.m[j].y)));
            bool form = scale > 0.0;
            for (int j = 0; j < nm && form; ++j) form = small(op.kind == 0 ? d.m[j].y : d.m[j].x, scale);
            if (form) d.nd = op.kind == 0 ? 1 : 2;
        }
    }
    return BSIM_OK;
}

// CNOT(c, t) = H_t CZ(c, t) H_t.  With every other op free of controls,
// the rewrite leaves an op list the lean streaming passes take (with
// reference Kraus ops, the lean Kraus passes): the CZ is a sign op, and the
// H's join the one-qubit ops of t in register chains (a controlled op keeps
// a pass on the full kernel).  Returns false (out
// untouched) when some other op keeps a control or cannot run in a tile.
static bool cnot_rewrite(const std::vector<EOp>& in, std::vector<EOp>& out) {
    auto is_cnot = [](const EOp& e) {
        const GateSpec& g = e.g;
        return e.kind == E_DENSE && !g.diag && g.k == 1 && popc(g.cmask) == 1 && g.cval == g.cmask &&
               e.nparams == 0 && e.aslot < 0 && g.m[0].x == 0.0 && g.m[0].y == 0.0 && g.m[1].x == 1.0 &&
               g.m[1].y == 0.0 && g.m[2].x == 1.0 && g.m[2].y == 0.0 && g.m[3].x == 0.0 && g.m[3].y == 0.0;
    };
    size_t ncnot = 0;
    for (const EOp& e : in) {
        if (is_cnot(e)) {
            ++ncnot;
            continue;
        }
        uint64_t pm;
        int pd;
        if (!e.fusable()) return false;
        if (e.kind == E_KRAUS_REF) continue;  // a tile op of the Kraus passes, no controls
        if (e.kind != E_DENSE && e.kind != E_DIAG) return false;
        if (e.g.cmask && !(e.kind == E_DIAG && sign_pair(e, pm, pd))) return false;
    }
    if (!ncnot) return false;
    out.clear();
    out.reserve(in.size() + 2 * ncnot);
    const double h = 0.70710678118654752440;
    for (const EOp& e : in) {
        if (!is_cnot(e)) {
            out.push_back(e);
            continue;
        }
        EOp H;
        H.kind = E_DENSE;
        H.g.k = 1;
        H.g.tb[0] = e.g.tb[0];
        H.g.m = { make_double2(h, 0.0), make_double2(h, 0.0), make_double2(h, 0.0), make_double2(-h, 0.0) };
        EOp Z;
        Z.kind = E_DIAG;
        Z.g.k = 1;
        Z.g.tb[0] = e.g.tb[0];
        Z.g.cmask = e.g.cmask;
        Z.g.cval = e.g.cval;
        Z.g.diag = true;
        Z.g.m = { make_double2(1.0, 0.0), make_double2(-1.0, 0.0) };
        out.push_back(H);
        out.push_back(Z);
        out.push_back(H);
    }
    return true;
}

int make_plan(int vq, const std::vector<EOp>& eops, const TileConfig& cfg, Plan& plan) {
    std::vector<EOp> rw;
    if (cfg.chains && cfg.tile_bits == 12 && cfg.reg_bits == 4 && vq >= 12 && cnot_rewrite(eops, rw)) {
        if (int rc = make_plan_ops(vq, rw, cfg, plan)) return rc;
        // single steps index the rewritten list: they carry their op
        for (Plan::Step& st : plan.steps)
            if (!st.tile && st.merged < 0) {
                plan.merged.push_back(rw[st.op]);
                st.merged = (int)plan.merged.size() - 1;
            }
        return BSIM_OK;
    }
    return make_plan_ops(vq, eops, cfg, plan);
}

}  // namespace bsim

#include "tile_kernel.h"

namespace bsim {

// One launch = one pass: block b runs tile b.  k_tile: 4 register bits per
// sweep (256 threads); k_tile_r3: 3 (512 threads, at most 128 VGPRs so that
// four waves fit on a SIMD); k_tile_kraus: 4, with the reference Kraus ops of
// density-matrix plans.
template <int R, bool KRAUS>
__device__ __forceinline__ void tile_pass(double2* __restrict__ sv, int vq, const Pass* __restrict__ pass,
                                          const Sweep* __restrict__ sweeps, const TOp* __restrict__ tops) {
    extern __shared__ double2 tile[];
    __shared__ uint64_t scat_lo[tile_threads<R>()];  // scatter of the low log2(TB) local bits
    __shared__ uint64_t scat_hi[64];                 // scatter of the others
    tile_body<R, KRAUS>(sv, vq, pass, sweeps, tops, blockIdx.x, false, tile, scat_lo, scat_hi);
}
__global__ __launch_bounds__(256) void k_tile(double2* __restrict__ sv, int vq, const Pass* __restrict__ pass,
                                              const Sweep* __restrict__ sweeps, const TOp* __restrict__ tops) {
    tile_pass<4, false>(sv, vq, pass, sweeps, tops);
}
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void k_tile_r3(
    double2* __restrict__ sv, int vq, const Pass* __restrict__ pass, const Sweep* __restrict__ sweeps,
    const TOp* __restrict__ tops) {
    tile_pass<3, false>(sv, vq, pass, sweeps, tops);
}
__global__ __launch_bounds__(256) void k_tile_kraus(double2* __restrict__ sv, int vq, const Pass* __restrict__ pass,
                                                    const Sweep* __restrict__ sweeps, const TOp* __restrict__ tops) {
    tile_pass<4, true>(sv, vq, pass, sweeps, tops);
}

// The streaming pass (12-bit tiles, 4 register bits): persistent workgroups,
// two per CU, taking tiles by ticket with the next tile prefetched.
template <bool LEAN, bool KRAUS>
__device__ __forceinline__ void stream_pass(double2* __restrict__ sv, int vq, const Pass* __restrict__ pass,
                                            const Sweep* __restrict__ sweeps, const TOp* __restrict__ tops,
                                            uint64_t ntiles, unsigned long long* __restrict__ ctr) {
    extern __shared__ double2 tile[];
    __shared__ uint64_t scat_lo[tile_threads<4>()];
    __shared__ uint64_t scat_hi[16];
    __shared__ uint64_t s_next[2];
    stream_body<LEAN, KRAUS>(sv, vq, pass, sweeps, tops, ntiles, ctr, tile, scat_lo, scat_hi, s_next);
}
__global__ __launch_bounds__(256, 2) void k_tile_stream(double2* __restrict__ sv, int vq,
                                                        const Pass* __restrict__ pass,
                                                        const Sweep* __restrict__ sweeps,
                                                        const TOp* __restrict__ tops, uint64_t ntiles,
                                                        unsigned long long* __restrict__ ctr) {
    stream_pass<false, false>(sv, vq, pass, sweeps, tops, ntiles, ctr);
}
__global__ __launch_bounds__(256, 2) void k_tile_stream_lean(double2* __restrict__ sv, int vq,
                                                             const Pass* __restrict__ pass,
                                                             const Sweep* __restrict__ sweeps,
                                                             const TOp* __restrict__ tops, uint64_t ntiles,
                                                             unsigned long long* __restrict__ ctr) {
    stream_pass<true, false>(sv, vq, pass, sweeps, tops, ntiles, ctr);
}
__global__ __launch_bounds__(256, 2) void k_tile_stream_kraus(double2* __restrict__ sv, int vq,
                                                              const Pass* __restrict__ pass,
                                                              const Sweep* __restrict__ sweeps,
                                                              const TOp* __restrict__ tops, uint64_t ntiles,
                                                              unsigned long long* __restrict__ ctr) {
    stream_pass<true, true>(sv, vq, pass, sweeps, tops, ntiles, ctr);
}

}  // namespace bsim

namespace bsim {

// ---------------------------------------------------------------------------
// execution
// ---------------------------------------------------------------------------

int launch_eop(bsim_state* h, double2* d, const EOp& e) {
    if (e.kind == E_KRAUS_REF) return launch_kraus_ref(h, d, e);
    return launch_gate(d, h->vq, e.g, h->scratch, h->stream);
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// ---- optional kernel timing (HIP events on the launching stream) ----------
struct Profiler {
    bool on = false;
    int64_t launches = 0;
    double total_ms = 0;
    std::vector<hipEvent_t> pool;  // reused start/stop pairs
    size_t used = 0;
};
static Profiler& prof() {
    static Profiler p;
    return p;
}

static int prof_mark(hipStream_t s, hipEvent_t* out) {
    Profiler& p = prof();
    if (p.used == p.pool.size()) {
        hipEvent_t e;
        BSIM_HIP(hipEventCreate(&e));
        p.pool.push_back(e);
    }
    *out = p.pool[p.used++];
    BSIM_HIP(hipEventRecord(*out, s));
    return BSIM_OK;
}

// after the stream has drained: fold the recorded pairs into the totals
static int prof_collect() {
    Profiler& p = prof();
    for (size_t i = 0; i + 1 < p.used; i += 2) {
        float ms = 0;
        BSIM_HIP(hipEventElapsedTime(&ms, p.pool[i], p.pool[i + 1]));
        p.total_ms += ms;
        p.launches += 1;
    }
    p.used = 0;
    return BSIM_OK;
}

size_t plan_table_bytes(const Plan& plan) {
    return align256(sizeof(Pass) * plan.passes.size()) + align256(sizeof(Sweep) * plan.sweeps.size()) +
           align256(sizeof(TOp) * plan.tops.size()) + 256;
}

int upload_plan(const Plan& plan, char* base, hipStream_t s) {
    if (plan.passes.empty()) return BSIM_OK;
    const size_t off_s = align256(sizeof(Pass) * plan.passes.size());
    const size_t off_o = off_s + align256(sizeof(Sweep) * plan.sweeps.size());
    BSIM_HIP(hipMemcpyAsync(base, plan.passes.data(), sizeof(Pass) * plan.passes.size(), hipMemcpyHostToDevice, s));
    BSIM_HIP(hipMemcpyAsync(base + off_s, plan.sweeps.data(), sizeof(Sweep) * plan.sweeps.size(), hipMemcpyHostToDevice,
                            s));
    if (!plan.tops.empty())
        BSIM_HIP(hipMemcpyAsync(base + off_o, plan.tops.data(), sizeof(TOp) * plan.tops.size(), hipMemcpyHostToDevice,
                                s));
    return BSIM_OK;
}

// The streaming kernels take 12-bit, 4-register-bit passes: the lean one
// those whose ops are all uncontrolled dense 1-/2-register ops and sign ops,
// the lean Kraus one the same plus reference Kraus ops (density-matrix
// plans), the full one any other pass without Kraus ops.
enum class StreamKind { none, full, lean, lean_kraus };

static StreamKind stream_kind(const Plan& plan, const Pass& P, int vq) {
    if (P.R != 4 || P.T != 12 || vq < 12) return StreamKind::none;
    bool lean = true;
    for (int s = P.sweep_begin; s < P.sweep_begin + P.nsweeps; ++s)
        for (int o = plan.sweeps[s].op_begin; o < plan.sweeps[s].op_end; ++o) {
            const TOp& t = plan.tops[o];
            if (t.kind == 4) {
                o += t.nd;  // a 4-bit Kraus op's matrices follow its head
                continue;
            }
            if (t.kind >= 8) continue;  // register chains: lean op sets only
            if (t.kind != 3 && ((t.kind != 0 && t.kind != 1) || t.cmask != 0)) lean = false;
        }
    if (plan.has_kraus) return lean ? StreamKind::lean_kraus : StreamKind::none;
    return lean ? StreamKind::lean : StreamKind::full;
}

// Workgroup slots of a streaming kernel on the current device: resident
// workgroups per CU (two: 69.6 KiB of LDS and up to 256 VGPRs each) times
// the CU count.
static const void* stream_kernel(StreamKind k) {
    return k == StreamKind::lean ? (const void*)k_tile_stream_lean
           : k == StreamKind::lean_kraus ? (const void*)k_tile_stream_kraus
                                          : (const void*)k_tile_stream;
}

static int stream_slots(StreamKind kind) {
    static std::atomic<int> cache[64][4];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int s = cache[dev][(int)kind].load();
    if (s) return s;
    const void* k = stream_kernel(kind);
    int per_cu = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, k, 256, sizeof(double2) * (size_t)((1 << 12) + (1 << 8)));
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    s = per_cu * device_cu_count();
    cache[dev][(int)kind].store(s);
    return s;
}

int launch_plan(bsim_state* h, double2* d, const std::vector<EOp>& eops, const Plan& plan, const char* base,
                bool timed) {
    const size_t off_s = align256(sizeof(Pass) * plan.passes.size());
    const size_t off_o = off_s + align256(sizeof(Sweep) * plan.sweeps.size());
    for (const Plan::Step& st : plan.steps) {
        if (st.tile) {
            const Pass& P = plan.passes[st.pass];
            unsigned grid = (unsigned)(1ull << (h->vq - P.T));
            size_t shmem = sizeof(double2) * (size_t)((1 << P.T) + ((1 << P.T) >> 4));
            hipEvent_t e0, e1;
            if (timed && prof().on)
                if (int rc = prof_mark(h->stream, &e0)) return rc;
            const bool r3 = P.R == 3;
            const StreamKind sk = stream_kind(plan, P, h->vq);
            if (sk != StreamKind::none) {
                const uint64_t ntiles = grid;
                const unsigned g = (unsigned)std::min<uint64_t>(ntiles, (uint64_t)stream_slots(sk));
                auto kern = sk == StreamKind::lean ? k_tile_stream_lean
                            : sk == StreamKind::lean_kraus ? k_tile_stream_kraus
                                                           : k_tile_stream;
                hipLaunchKernelGGL(kern, dim3(g), dim3(256), shmem, h->stream, d, h->vq,
                                   (const Pass*)base + st.pass, (const Sweep*)(base + off_s),
                                   (const TOp*)(base + off_o), ntiles, h->tile_ctr);
            } else {
                hipLaunchKernelGGL(r3 ? k_tile_r3 : (plan.has_kraus ? k_tile_kraus : k_tile), dim3(grid),
                                   dim3(r3 ? 512 : 256), shmem, h->stream, d, h->vq, (const Pass*)base + st.pass,
                                   (const Sweep*)(base + off_s), (const TOp*)(base + off_o));
            }
            BSIM_HIP(hipGetLastError());
            if (timed && prof().on)
                if (int rc = prof_mark(h->stream, &e1)) return rc;
        } else {
            if (int rc = launch_eop(h, d, st.merged >= 0 ? plan.merged[st.merged] : eops[st.op])) return rc;
        }
    }
    return BSIM_OK;
}

int run_plan(bsim_state* h, double2* d, const std::vector<EOp>& eops, const Plan& plan) {
    if (!plan.passes.empty()) {
        if (int rc = ensure_ws(h->ops, plan_table_bytes(plan))) return rc;
        if (int rc = upload_plan(plan, (char*)h->ops.ptr, h->stream)) return rc;
    }
    return launch_plan(h, d, eops, plan, (const char*)h->ops.ptr, true);
}

static std::atomic<bool> g_shmem_set{ false };

int prepare_tile_kernel() {
    if (g_shmem_set) return BSIM_OK;
    // allow > 64 KiB of dynamic LDS: a padded 13-bit tile (the kernel's
    // static scatter tables take another 2.5 KiB of the 160 KiB)
    for (const void* k : { (const void*)k_tile, (const void*)k_tile_r3, (const void*)k_tile_kraus,
                           (const void*)k_tile_stream, (const void*)k_tile_stream_lean,
                           (const void*)k_tile_stream_kraus })
        BSIM_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(sizeof(double2) * ((1 << 13) + (1 << 9)))));
    g_shmem_set = true;
    return BSIM_OK;
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_set_tile_config(int tile_bits, int low_bits) {
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    TileConfig c = g_cfg;
    if (tile_bits) {
        if (tile_bits < 6 || tile_bits > 13) return fail(BSIM_EINVAL, "tile_bits must be 6..13");
        c.tile_bits = tile_bits;
    }
    if (low_bits) {
        if (low_bits < 0 || low_bits > c.tile_bits - 2) return fail(BSIM_EINVAL, "low_bits must be 0..tile_bits-2");
        c.low_bits = low_bits;
    }
    if (c.low_bits > c.tile_bits - 2) c.low_bits = c.tile_bits - 2;
    g_cfg = c;
    return BSIM_OK;
}

int bsim_get_tile_config(int* tile_bits, int* low_bits) {
    const TileConfig c = tile_config();
    if (tile_bits) *tile_bits = c.tile_bits;
    if (low_bits) *low_bits = c.low_bits;
    return BSIM_OK;
}

int bsim_set_sweep_bits(int bits, int* old) {
    if (bits != 0 && bits != 3 && bits != 4) return fail(BSIM_EINVAL, "sweep register bits must be 3 or 4");
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    if (old) *old = g_cfg.reg_bits;
    if (bits) g_cfg.reg_bits = bits;
    return BSIM_OK;
}

int bsim_set_merge_width(int bits, int* old) {
    if (bits < -1 || bits == 1 || bits == 2 || bits > 6)
        return fail(BSIM_EINVAL, "merge width must be 0 (off) or 3..6 (-1: read only)");
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    if (old) *old = g_cfg.merge_bits;
    if (bits >= 0) g_cfg.merge_bits = bits;
    return BSIM_OK;
}

int bsim_set_fusion_limit(int max_cost, int* old) {
    if (max_cost < 0) return fail(BSIM_EINVAL, "max_cost must be >= 0");
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    if (old) *old = g_cfg.max_cost;
    g_cfg.max_cost = max_cost;
    return BSIM_OK;
}

int bsim_set_profiling(int enable) {
    prof().on = enable != 0;
    return BSIM_OK;
}

int bsim_kernel_stats(int reset, int64_t* launches, double* total_ms) {
    if (launches) *launches = prof().launches;
    if (total_ms) *total_ms = prof().total_ms;
    if (reset) {
        prof().launches = 0;
        prof().total_ms = 0;
    }
    return BSIM_OK;
}

int bsim_plan_stats(int vector_qubits, int is_dm, const bsim_op* ops, int n_ops, int64_t* stats) {
    if (!stats || (n_ops > 0 && !ops)) return fail(BSIM_EINVAL, "null argument");
    int nq = is_dm ? vector_qubits / 2 : vector_qubits;
    std::vector<EOp> eops;
    if (int rc = lower_ops(nq, is_dm ? BSIM_DM : BSIM_SV, ops, n_ops, eops, false)) return rc;
    Plan plan;
    if (int rc = make_plan(is_dm ? 2 * nq : nq, eops, tile_config(), plan)) return rc;
    int64_t singles = 0;
    for (auto& s : plan.steps) singles += s.tile ? 0 : 1;
    stats[0] = (int64_t)plan.passes.size();
    stats[1] = singles;
    stats[2] = (int64_t)plan.sweeps.size();
    stats[3] = plan.fused_ops;
    return BSIM_OK;
}

int bsim_evolve(bsim_state* h, const bsim_op* ops, int n_ops, int mode) {
    if (int rc = check_state(h)) return rc;
    if (n_ops < 0 || (n_ops > 0 && !ops)) return fail(BSIM_EINVAL, "bad op list");
    std::vector<EOp> eops;
    if (int rc = lower_ops(h->nq, h->kind, ops, n_ops, eops, false)) return rc;
    h->alias_ready = false;
    if (mode == 1) {
        for (const EOp& e : eops)
            if (int rc = launch_eop(h, h->d, e)) return rc;
    } else {
        if (int rc = prepare_tile_kernel()) return rc;
        Plan plan;
        if (int rc = make_plan(h->vq, eops, tile_config(), plan)) return rc;
        if (int rc = run_plan(h, h->d, eops, plan)) return rc;
    }
    BSIM_HIP(hipStreamSynchronize(h->stream));
    if (prof().on)
        if (int rc = prof_collect()) return rc;
    return BSIM_OK;
}

}  // extern "C"
