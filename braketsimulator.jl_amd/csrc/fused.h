// Fused evolution: instruction stream -> elementary vector ops -> planned
// HBM passes (LDS tiles) -> register sweeps.  Shared with the adjoint sweep.
#pragma once

#include <functional>
#include <vector>

#include "common.h"
#include "gates.h"

namespace bsim {

enum EKind { E_DENSE = 0, E_DIAG = 1, E_KRAUS_REF = 2 };

// One elementary op on the 2^vq vector.
struct EOp {
    int kind = E_DENSE;
    GateSpec g;                  // dense / diag description (vector bits)
    // Kraus (reference semantics): channel qubits + matrices, applied by its own kernel
    int kq = 0;
    int ktargets[8] = {0};
    int nkraus = 0;
    std::vector<double2> kraus;
    // ... and, for kq <= 2, the same map as a predicated superoperator for the
    // tile passes: g (dense on the 2 kq row/column bits) is applied where the
    // stored row and column agree on every other bit,
    // ((i ^ (i >> kshift)) & kmask) == 0.
    uint64_t kmask = 0;
    int kshift = 0;
    // kpres: the superoperator links no block entry whose stored row and
    // column agree on the channel qubits with one where they differ.  Such a
    // predicated op commutes with every diagonal op and every other kpres op
    // on other bits (both keep each other's predicate); without it the op
    // commutes with nothing, since its predicate reads every other bit.
    bool kpres = false;
    // dense on 4 bits that may run as a tile op (the 16x16 superoperator of a
    // 2-qubit physical channel); other 3+-bit dense ops launch on their own
    bool tile4 = false;
    // differentiation (set by the adjoint path)
    int nparams = 0;
    int pidx[3] = { -1, -1, -1 };
    std::vector<double2> dm[3];
    int aslot = -1;              // fused adjoint: partial-sum slot of the op's parameters
    // fused adjoint, a run of one-qubit ops merged into one op: product p
    // belongs to parameter pprm[p] of slot pslot[p] (pslot[p] < 0: parameter
    // p of aslot) and is taken before the op's update when ppre[p]
    int pslot[3] = { -1, -1, -1 };
    int pprm[3] = { 0, 0, 0 };
    bool ppre[3] = { false, false, false };
    uint64_t touched() const;    // every vector bit the op reads, writes or depends on
                                 // (a kpres Kraus op: its row/column bits; its predicate
                                 // is accounted for by pair_diag)
    bool pair_diag() const;      // keeps every row/column pair predicate: diagonal ops, kpres Kraus
    bool fusable() const;        // fits a tile sweep (dense k <= 2, diag k <= 4, reference Kraus kq <= 2)
};

// Translate bsim_op instructions into elementary ops on the vector.
// DM gates become U on column bits + conj(U) on row bits; physical Kraus
// becomes one superoperator; PhaseFlip a 2-bit diagonal.
int lower_ops(int nq, int kind, const bsim_op* ops, int n_ops, std::vector<EOp>& out, bool keep_params);

// Device-side tile program ------------------------------------------------

// 64-byte aligned (tables start at 256-byte offsets): an op's head load
// (kind .. nd, cmask, cval, m[0..4]) reads exactly its first two lines.
struct alignas(64) TOp {
    // 0 dense1, 1 dense2, 2 diag, 3 sign parity, 4 predicated superoperator,
    // 5 its matrix, 6 diagonal on register bits only (m[j] = the factor of
    // register j, so every thread reads the same entries), 7 X on register a
    // (swaps, 3-register sweeps only; m holds the X matrix).
    // Register chains (12-bit, 4-register passes of the lean streaming
    // kernels): each thread keeps a 2x2 matrix A and, with
    // t(mask) = parity(gbase & mask) read from its sub-cube base,
    //   8  A <- U Z^t(cmask)             (U = m[0..3], column-major)
    //   9  A <- U Z^t(cmask) A
    //   10 A <- Z^t(cmask) A, then A is applied to register a
    //   11 A <- (-1)^P(gbase) A, P the sign op's pair form (its packed
    //      copy: nd = pair distances, cmask >> 32 = pair count, m[0..1] =
    //      pair masks)
    //   12 a whole chain of nd = 1..4 matrices in one op, applied to
    //      register a: A = Z^t(sgm[3]) U_nd ... Z^t(sgm[0]) U_1 Z^t(cmask)
    //      (U_k = m[4k-4..4k-1]), times (-1)^P(gbase) with P = pairs
    //      (pm, pd, npairs) when npairs > 0
    int32_t kind;
    int32_t a, b;      // dense: register bits (a < b for dense2)
    int32_t nd;        // diag: bit count; kind 4: TOps that follow with the matrix (0 or 15)
    uint64_t cmask, cval;
    // the matrix right after the head fields: an op's head load (kind .. nd,
    // cmask, m[0..3]) touches the first two 64-byte lines of its record
    double2 m[16];
    int32_t dbit[4];   // diag: vector bit of diagonal index bit i
    // sign: amplitude i is negated when sum_k popcount(i & (i >> pd[k]) & pm[k]) is odd
    uint64_t pm[4];
    int32_t pd[4];
    int32_t npairs;
    // sign op left by the register-chain rewrite (the per-thread sign and
    // the Z factors of registers without a dense op): defined by its split
    // form below only, its pairs hold just the per-thread part
    int32_t split_only;
    // Split of the index tests over a sweep (computed by the planner): a
    // register j of a sub-cube has index gbase | r_j, gbase the sub-cube's
    // base and r_j the global bits of j's register bits.
    //   dense and kinds 6, 7 with controls: on_j = ((gbase & tmask) == tval) && bit j of jmask;
    //   sign: negate_j = bit j of (jmask ^ P(gbase) * 0xFFFF ^ sum_i parity(gbase & sgm[i]) * pattern_i),
    //         P(x) = sum_k parity(x & (x >> pd[k]) & pm[k]), pattern_i = registers with bit i set;
    //   kind 4: on_j = (Q(gbase) ^ sum over the register bits i of j of sgm[i]) == 0,
    //         Q(x) = (x ^ (x >> pd[0])) & pm[0] (linear, so Q(gbase | r_j) splits).
    //   Kind 4 on 2 registers (a < b) keeps its 4x4 matrix in m; on all 4
    //   registers its 16x16 matrix (register order) is column c = m of the
    //   c-th TOp from this one (the 15 that follow are kind 5).
    uint64_t tmask, tval;
    uint64_t sgm[4];
    uint32_t jmask;
    // fused adjoint: the op's product matrices follow as `nparam` kind-5
    // TOps (m = the EOp's dm in this op's layout, a = partial-sum slot,
    // b = parameter, pre = 1: taken before the op's update).
    // Pauli-coded products of an uncontrolled dense 1-register op (kind 0,
    // cmask 0, cval != 0): payload -i s P with P = X, Y or Z, read from the
    // head alone: cval = post | pre << 32, a word = row << 2 | P (P = 1 X,
    // 2 Y, 3 Z; 0 no product; row = slot * PM + parameter), m[4] = (s of the
    // pre product, s of the post product); the kind-5 TOps follow as well
    int32_t nparam;
    int32_t pre;
};

struct Sweep {
    int32_t sb[4];     // local bit indices of the pass's R register bits, ascending
    int32_t op_begin, op_end;
};

struct Pass {
    int32_t T;                // tile bits
    int32_t nsweeps;
    int32_t sweep_begin;
    int32_t R;                // register bits per sweep (3 or 4)
    int32_t lbits[16];        // vector bit of local bit i (ascending)
    uint64_t lmask;
};

struct Plan {
    // sequence of steps: a step is either a tile pass or a single op launch
    struct Step {
        bool tile;
        int pass;              // index into passes
        int op;                // index into eops (single launch)
        int merged = -1;       // >= 0: single launch of merged[merged] instead
    };
    std::vector<Step> steps;
    std::vector<Pass> passes;
    std::vector<Sweep> sweeps;
    std::vector<TOp> tops;
    std::vector<int> top_src;  // eop index each TOp came from
    // runs of adjacent dense ops around a 3+-bit one, multiplied into one
    // op on at most 6 bits (one HBM pass on the matrix cores instead of
    // several)
    std::vector<EOp> merged;
    int64_t fused_ops = 0;
    bool has_kraus = false;    // holds kind-4 ops: launch k_tile_kraus
};

struct TileConfig {
    int tile_bits = 12;
    int low_bits = 3;  // 8 contiguous amplitudes (128 B) per run: fewer passes than 4, same coalescing
    // register bits per sweep: 4 (16-amplitude sub-cubes, 256 threads, two
    // waves per SIMD) or 3 (8 amplitudes, 512 threads, four waves: the same
    // speed on the 30-qubit brick-wall, more sweeps); a pass with a 4-bit op
    // or a reference Kraus op always uses 4
    int reg_bits = 4;
    // FP64 work budget of one pass, in operations per amplitude (dense 1q
    // = 8, dense 2q = 32, diagonal = 4, sign = 1); 0 = unlimited.  Bounds
    // the fusion width so that a pass stays HBM-bound.
    int max_cost = 0;
    // merge runs of adjacent dense ops that do not fit a tile pass (3+
    // target bits) with their neighbours into one op of at most merge_bits
    // bits (0: no merging; 3..6)
    bool merge_dense = true;
    int merge_bits = 6;
    // rewrite the sweeps of 12-bit, 4-register passes into register chains
    // (tile op kinds 8..11, run by the lean streaming kernels only); callers
    // whose kernels take the full op set (the batch kernel) turn it off
    bool chains = true;
};
// A snapshot of the library-wide configuration (taken under a lock), and
// its setter; every entry point plans with one snapshot.
TileConfig tile_config();
void set_tile_config(const TileConfig& c);

int make_plan(int vq, const std::vector<EOp>& eops, const TileConfig& cfg, Plan& plan);

// Launch a plan on d (ops buffer reused).  `eops` provides single launches.
int run_plan(bsim_state* h, double2* d, const std::vector<EOp>& eops, const Plan& plan);
// The same in two parts: the plan's device tables (passes | sweeps | ops,
// plan_table_bytes bytes), then the launches reading them (`timed`: tile
// passes are bracketed by the profiling events when profiling is on).
size_t plan_table_bytes(const Plan& plan);
int upload_plan(const Plan& plan, char* dst, hipStream_t s);
int launch_plan(bsim_state* h, double2* d, const std::vector<EOp>& eops, const Plan& plan, const char* tables,
                bool timed);

// One-time launch attributes of the tile kernel (dynamic LDS above 64 KiB).
int prepare_tile_kernel();

// Single elementary op launch (per-gate path).
int launch_eop(bsim_state* h, double2* d, const EOp& e);

// Kraus channel with reference (block-diagonal) semantics.
int launch_kraus_ref(bsim_state* h, double2* d, const EOp& e);

// Backward list of the state-vector adjoint: back[i] = U^dagger of
// eops[L - 1 - i] with dm[p] = U^dagger dU_p (the product matrix against
// the updated vectors), parametrized ops numbered (aslot) in forward order,
// and its plan (11 tile bits, 3 register bits).  nslots: parametrized ops;
// pm: their largest parameter count.
// `base` (optional) is the configuration snapshot whose other fields the
// plan keeps; by default the library configuration.  Runs of uncontrolled
// one-qubit ops on one qubit are merged into one op of `back` (their
// products taken before or after it); `src` (optional) receives the eops
// index of each op of `back` (-1: a merged run).
int adjoint_back_plan(int vq, const std::vector<EOp>& eops, std::vector<EOp>& back, Plan& plan, int& nslots,
                      int& pm, const TileConfig* base = nullptr, std::vector<int>* src = nullptr);

// Pauli form of a product payload (2 x 2, column-major): m = -i s P with
// P = X (1), Y (2), Z (3) and s real, to rounding; 0 when it has none.
int pauli_form(const double2* m, double& s);

// Bytes of the per-tile partial sums adjoint_tile_run needs for `plan`.
size_t adjoint_tile_partial_bytes(const Plan& plan, int vq, int nslots, int pm);

// Fused backward sweep of the state-vector adjoint: runs `plan` (a plan of
// the reversed U^dagger ops, 11 tile bits, 3 register bits; parametrized ops
// carry aslot and their derivative payloads) on (psi, lam); a non-tile step
// calls single(op index).  Writes out3[slot * 3 + p] = Re <lam|dU_p|psi> of
// every slot (pm: the largest parameter count of an op).
int adjoint_tile_run(bsim_state* h, double2* psi, double2* lam, const Plan& plan, int nslots, int pm, double* out3,
                     const std::function<int(int)>& single);

}  // namespace bsim
