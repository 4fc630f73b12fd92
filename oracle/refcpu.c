/*
 * refcpu.c -- CPU restatement of BraketSimulator.jl's evolution hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP library is
 * checked against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path (libbsim) never does.
 *
 * Every function restates one reference loop, cited as file:line of the
 * BraketSimulator.jl sources (v0.0.8).  The loops keep the reference's
 * index arithmetic (pad_bit / flip_bit, endian qubit = n-1-q), its matrix
 * conventions (2-qubit dense kernel: t1 is the LSB of the 4x4 index;
 * k-qubit fallback: sorted endian targets, given order ignored; 2-qubit
 * Kraus: t1 is the MSB) and its evaluation order for sums of products.
 * Complex products are written out as Julia's Base `*` for Complex{Float64}
 * (no fused multiply-add); build with -ffp-contract=off.
 *
 * Storage: complex128 interleaved (re, im) -- byte-identical to Julia's
 * Vector{ComplexF64}.  Density matrices are column-major N x N, viewed as a
 * 4^n vector whose low n bits are the row index and high n bits the column.
 *
 * OpenMP parallelises the same loops the reference threads
 * (Threads.@threads); every element is written by one iteration, so the
 * results do not depend on the thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return r;
}
static inline cplx cadd(cplx a, cplx b) { cplx r = { a.re + b.re, a.im + b.im }; return r; }
static inline cplx cconj(cplx a) { cplx r = { a.re, -a.im }; return r; }

/* gate_kernels.jl:34-38 */
static inline int64_t pad_bit(int64_t ix, int bit) {
    int64_t left = (ix >> bit) << bit;
    int64_t right = ix - left;
    return (left << 1) ^ right;
}
/* gate_kernels.jl:78-84 (bits applied in the given order, not adjusted) */
static inline int64_t pad_bits(int64_t ix, const int* bits, int nb) {
    for (int i = 0; i < nb; ++i) ix = pad_bit(ix, bits[i]);
    return ix;
}
/* gate_kernels.jl:119-121 */
static inline int64_t flip_bit(int64_t ix, int bit) { return ix ^ ((int64_t)1 << bit); }

static void sort_ints(int* a, int n) {
    for (int i = 1; i < n; ++i) {
        int v = a[i], j = i - 1;
        while (j >= 0 && a[j] > v) { a[j + 1] = a[j]; --j; }
        a[j + 1] = v;
    }
}

static inline cplx ld(const double* m, int i) { cplx r = { m[2 * i], m[2 * i + 1] }; return r; }

/* ------------------------------------------------------------------ */
/* State-vector gate kernels                                           */
/* ------------------------------------------------------------------ */

/* gate_kernels.jl:489-537.  m = column-major 2x2 (g_00, g_10, g_01, g_11). */
void ref_apply_1q(double* sv_, int nq, int qubit, const double* m) {
    cplx* sv = (cplx*)sv_;
    int e = nq - 1 - qubit;                       /* endian_qubits, :193 */
    int64_t n_tasks = ((int64_t)1 << nq) >> 1;
    int64_t flipper = (int64_t)1 << e;
    cplx g00 = ld(m, 0), g10 = ld(m, 1), g01 = ld(m, 2), g11 = ld(m, 3);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < n_tasks; ++t) {
        int64_t lo = pad_bit(t, e), hi = lo + flipper;
        cplx a = sv[lo], b = sv[hi];
        sv[lo] = cadd(cmul(g00, a), cmul(g01, b));
        sv[hi] = cadd(cmul(g10, a), cmul(g11, b));
    }
}

/* gate_kernels.jl:540-576.  4x4 column-major; amplitude order
 * (ix_00, flip t1, flip t2, ix_11), so t1 is bit 0 of the matrix index. */
void ref_apply_2q(double* sv_, int nq, int t1, int t2, const double* m) {
    cplx* sv = (cplx*)sv_;
    int e1 = nq - 1 - t1, e2 = nq - 1 - t2;
    int st[2] = { e1 < e2 ? e1 : e2, e1 < e2 ? e2 : e1 };
    int64_t n_tasks = ((int64_t)1 << nq) >> 2;
    cplx g[16];
    for (int i = 0; i < 16; ++i) g[i] = ld(m, i);
#pragma omp parallel for schedule(static)
    for (int64_t ix = 0; ix < n_tasks; ++ix) {
        int64_t i00 = pad_bits(ix, st, 2);
        int64_t i10 = flip_bit(i00, e2);
        int64_t i01 = flip_bit(i00, e1);
        int64_t i11 = flip_bit(i10, e1);
        int64_t id[4] = { i00, i01, i10, i11 };
        cplx a[4] = { sv[i00], sv[i01], sv[i10], sv[i11] };
        cplx r[4];
        for (int row = 0; row < 4; ++row) {
            cplx acc = cmul(g[row], a[0]);        /* g[row, col] at row + 4 col */
            for (int col = 1; col < 4; ++col) acc = cadd(acc, cmul(g[row + 4 * col], a[col]));
            r[row] = acc;
        }
        for (int k = 0; k < 4; ++k) sv[id[k]] = r[k];
    }
}

/* gate_kernels.jl:579-606: one control, one target, 2x2 on the c_bit half. */
void ref_apply_ctrl_1c1t(double* sv_, int nq, int c_bit, int control, int target, const double* m) {
    cplx* sv = (cplx*)sv_;
    int ec = nq - 1 - control, et = nq - 1 - target;
    int small_t = ec < et ? ec : et, big_t = ec < et ? et : ec;
    cplx g00 = ld(m, 0), g10 = ld(m, 1), g01 = ld(m, 2), g11 = ld(m, 3);
    int64_t n = ((int64_t)1 << nq) >> 2;
#pragma omp parallel for schedule(static)
    for (int64_t ix = 0; ix < n; ++ix) {
        int64_t i00 = pad_bit(pad_bit(ix, small_t), big_t);
        int64_t i10 = flip_bit(i00, ec);
        int64_t i01 = flip_bit(i00, et);
        int64_t i11 = flip_bit(i01, ec);
        int64_t lo = c_bit ? i10 : i00, hi = c_bit ? i11 : i01;
        cplx a = sv[lo], b = sv[hi];
        sv[lo] = cadd(cmul(g00, a), cmul(g01, b));
        sv[hi] = cadd(cmul(g10, a), cmul(g11, b));
    }
}

/* gate_kernels.jl:608-636: one control, two targets (CSwap); order
 * (00, flip t1, flip t2, 11), t1 = bit 0 of the 4x4 index. */
void ref_apply_ctrl_1c2t(double* sv_, int nq, int c_bit, int control, int t1, int t2, const double* m) {
    cplx* sv = (cplx*)sv_;
    int ec = nq - 1 - control, e1 = nq - 1 - t1, e2 = nq - 1 - t2;
    int s[3] = { ec, e1, e2 };
    sort_ints(s, 3);
    cplx g[16];
    for (int i = 0; i < 16; ++i) g[i] = ld(m, i);
    int64_t n = ((int64_t)1 << nq) >> 3;
#pragma omp parallel for schedule(static)
    for (int64_t ix = 0; ix < n; ++ix) {
        int64_t i00 = pad_bits(ix, s, 3);
        if (c_bit) i00 = flip_bit(i00, ec);
        int64_t i10 = flip_bit(i00, e2);
        int64_t i01 = flip_bit(i00, e1);
        int64_t i11 = flip_bit(i01, e2);
        int64_t id[4] = { i00, i01, i10, i11 };
        cplx a[4] = { sv[i00], sv[i01], sv[i10], sv[i11] };
        cplx r[4];
        for (int row = 0; row < 4; ++row) {
            cplx acc = cmul(g[row], a[0]);
            for (int col = 1; col < 4; ++col) acc = cadd(acc, cmul(g[row + 4 * col], a[col]));
            r[row] = acc;
        }
        for (int k = 0; k < 4; ++k) sv[id[k]] = r[k];
    }
}

/* gate_kernels.jl:638-670: two controls, one target (CCNot). */
void ref_apply_ctrl_2c1t(double* sv_, int nq, int c1_bit, int c2_bit, int c1, int c2, int target,
                         const double* m) {
    cplx* sv = (cplx*)sv_;
    int e1 = nq - 1 - c1, e2 = nq - 1 - c2, et = nq - 1 - target;
    int s[3] = { e1, e2, et };
    sort_ints(s, 3);
    cplx g00 = ld(m, 0), g10 = ld(m, 1), g01 = ld(m, 2), g11 = ld(m, 3);
    int64_t n = ((int64_t)1 << nq) >> 3;
#pragma omp parallel for schedule(static)
    for (int64_t ix = 0; ix < n; ++ix) {
        int64_t lo = pad_bits(ix, s, 3);
        if (c1_bit) lo = flip_bit(lo, e1);
        if (c2_bit) lo = flip_bit(lo, e2);
        int64_t hi = flip_bit(lo, et);
        cplx a = sv[lo], b = sv[hi];
        sv[lo] = cadd(cmul(g00, a), cmul(g01, b));
        sv[hi] = cadd(cmul(g10, a), cmul(g11, b));
    }
}

/* gate_kernels.jl:753-775: k-qubit dense fallback.  The matrix index bit j
 * is the j-th smallest endian target (the given target order is ignored). */
void ref_apply_kq(double* sv_, int nq, int k, const int* targets, const double* m) {
    cplx* sv = (cplx*)sv_;
    int ord[16];
    for (int i = 0; i < k; ++i) ord[i] = nq - 1 - targets[i];
    sort_ints(ord, k);
    int N = 1 << k;
    int64_t n = ((int64_t)1 << nq) >> k;
    cplx* g = (cplx*)malloc(sizeof(cplx) * N * N);
    for (int i = 0; i < N * N; ++i) g[i] = ld(m, i);
#pragma omp parallel
    {
        int64_t* ixs = (int64_t*)malloc(sizeof(int64_t) * N);
        cplx* a = (cplx*)malloc(sizeof(cplx) * N);
#pragma omp for schedule(static)
        for (int64_t ix = 0; ix < n; ++ix) {
            int64_t p = pad_bits(ix, ord, k);
            for (int j = 0; j < N; ++j) {
                int64_t f = p;
                for (int b = 0; b < k; ++b) if ((j >> b) & 1) f = flip_bit(f, ord[b]);
                ixs[j] = f;
                a[j] = sv[f];
            }
            for (int row = 0; row < N; ++row) {
                cplx acc = cmul(g[row], a[0]);
                for (int col = 1; col < N; ++col) acc = cadd(acc, cmul(g[row + N * col], a[col]));
                sv[ixs[row]] = acc;
            }
        }
        free(ixs);
        free(a);
    }
    free(g);
}

/* custom_gates.jl:390-416: DoubleExcitation on (t1,t2,t3,t4); (c, s) =
 * (cos, sin) of angle*pow/2. */
void ref_apply_double_excitation(double* sv_, int nq, const int* t, double c, double s) {
    cplx* sv = (cplx*)sv_;
    int e[4], ord[4];
    for (int i = 0; i < 4; ++i) ord[i] = e[i] = nq - 1 - t[i];
    sort_ints(ord, 4);
    int64_t n = ((int64_t)1 << nq) >> 4;
#pragma omp parallel for schedule(static)
    for (int64_t ix = 0; ix < n; ++ix) {
        int64_t p = pad_bits(ix, ord, 4);
        int64_t i0011 = flip_bit(flip_bit(p, e[2]), e[3]);
        int64_t i1100 = flip_bit(flip_bit(p, e[0]), e[1]);
        cplx a = sv[i0011], b = sv[i1100];
        cplx r0 = { c * a.re - s * b.re, c * a.im - s * b.im };
        cplx r1 = { s * a.re + c * b.re, s * a.im + c * b.im };
        sv[i0011] = r0;
        sv[i1100] = r1;
    }
}

/* Not a reference loop: a dense operator on k vector bits with the
 * convention of ref_apply_kq (bit j of the matrix index is the j-th
 * smallest endian target), skipping the operator's zero entries.  Used for
 * the superoperators of physical noise channels on large density matrices,
 * where most entries are zero; pinned against the block form of
 * ref_dm_kraus_physical on small registers. */
void ref_apply_kq_sparse(double* sv_, int nq, int k, const int* targets, const double* m) {
    cplx* sv = (cplx*)sv_;
    int ord[16];
    for (int i = 0; i < k; ++i) ord[i] = nq - 1 - targets[i];
    sort_ints(ord, k);
    int N = 1 << k;
    int64_t n = ((int64_t)1 << nq) >> k;
    int* nzc = (int*)malloc(sizeof(int) * N * N);   /* per row: nonzero columns */
    cplx* nzv = (cplx*)malloc(sizeof(cplx) * N * N);
    int* cnt = (int*)calloc(N, sizeof(int));
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * N);
    for (int row = 0; row < N; ++row)
        for (int col = 0; col < N; ++col) {
            const double* e = m + 2 * (row + N * col);
            if (e[0] != 0.0 || e[1] != 0.0) {
                nzc[row * N + cnt[row]] = col;
                nzv[row * N + cnt[row]].re = e[0];
                nzv[row * N + cnt[row]].im = e[1];
                ++cnt[row];
            }
        }
    for (int j = 0; j < N; ++j) {
        int64_t f = 0;
        for (int b = 0; b < k; ++b) if ((j >> b) & 1) f |= (int64_t)1 << ord[b];
        off[j] = f;
    }
#pragma omp parallel
    {
        cplx* a = (cplx*)malloc(sizeof(cplx) * N);
#pragma omp for schedule(static)
        for (int64_t ix = 0; ix < n; ++ix) {
            int64_t p = pad_bits(ix, ord, k);
            for (int j = 0; j < N; ++j) a[j] = sv[p | off[j]];
            for (int row = 0; row < N; ++row) {
                cplx acc = { 0.0, 0.0 };
                for (int t = 0; t < cnt[row]; ++t) acc = cadd(acc, cmul(nzv[row * N + t], a[nzc[row * N + t]]));
                sv[p | off[row]] = acc;
            }
        }
        free(a);
    }
    free(nzc); free(nzv); free(cnt); free(off);
}

/* sv_simulator.jl:179-198: 2-qubit Hermitian observable; order
 * (00, flip t2, flip t1, 11), i.e. t1 is the MSB of the 4x4 index. */
void ref_apply_hermitian_2q(double* sv_, int nq, int t1, int t2, const double* m) {
    cplx* sv = (cplx*)sv_;
    int e1 = nq - 1 - t1, e2 = nq - 1 - t2;
    int st[2] = { e1 < e2 ? e1 : e2, e1 < e2 ? e2 : e1 };
    cplx g[16];
    for (int i = 0; i < 16; ++i) g[i] = ld(m, i);
    int64_t n = ((int64_t)1 << nq) >> 2;
#pragma omp parallel for schedule(static)
    for (int64_t ix = 0; ix < n; ++ix) {
        int64_t i00 = pad_bits(ix, st, 2);
        int64_t i10 = flip_bit(i00, e2);
        int64_t i01 = flip_bit(i00, e1);
        int64_t i11 = flip_bit(i10, e1);
        int64_t id[4] = { i00, i10, i01, i11 };
        cplx a[4] = { sv[i00], sv[i10], sv[i01], sv[i11] };
        for (int row = 0; row < 4; ++row) {
            cplx acc = cmul(g[row], a[0]);
            for (int col = 1; col < 4; ++col) acc = cadd(acc, cmul(g[row + 4 * col], a[col]));
            sv[id[row]] = acc;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Density-matrix noise kernels                                        */
/* ------------------------------------------------------------------ */

/* noise_kernels.jl:7-23: PhaseFlip, elementwise over all 4^n entries. */
void ref_dm_phaseflip(double* dm_, int nq, int qubit, double p) {
    cplx* dm = (cplx*)dm_;
    int64_t N = (int64_t)1 << nq;
    int e = nq - 1 - qubit;
#pragma omp parallel for schedule(static)
    for (int64_t idx = 0; idx < N * N; ++idx) {
        int64_t ix = idx % N, jx = idx / N;
        double iv = 1.0 - 2.0 * (double)((ix >> e) & 1);
        double jv = 1.0 - 2.0 * (double)((jx >> e) & 1);
        double f = (1.0 - p) + iv * jv * p;
        dm[ix + jx * N].re *= f;
        dm[ix + jx * N].im *= f;
    }
}

/* B <- sum_m K_m B K_m^dagger for a d x d block (column-major). */
static void kraus_block(const cplx* K, int nK, int d, const cplx* B, cplx* out, cplx* tmp) {
    for (int i = 0; i < d * d; ++i) { out[i].re = 0; out[i].im = 0; }
    for (int mi = 0; mi < nK; ++mi) {
        const cplx* k = K + (size_t)mi * d * d;
        /* tmp = K * B */
        for (int r = 0; r < d; ++r)
            for (int c = 0; c < d; ++c) {
                cplx acc = cmul(k[r], B[c * d]);
                for (int l = 1; l < d; ++l) acc = cadd(acc, cmul(k[r + l * d], B[l + c * d]));
                tmp[r + c * d] = acc;
            }
        /* out += tmp * K^dagger ;  K^dagger[l, c] = conj(K[c, l]) */
        for (int r = 0; r < d; ++r)
            for (int c = 0; c < d; ++c) {
                cplx acc = cmul(tmp[r], cconj(k[c]));
                for (int l = 1; l < d; ++l) acc = cadd(acc, cmul(tmp[r + l * d], cconj(k[c + l * d])));
                out[r + c * d] = cadd(out[r + c * d], acc);
            }
    }
}

/* noise_kernels.jl:105-196: Kraus channels, reference ("block-diagonal")
 * semantics: only the 2^k x 2^k blocks whose row and column share every
 * non-target bit are updated.  Block index order:
 *   k=1: (lower, higher)                            (:113-121)
 *   k=2: (00, flip t2, flip t1, 11) -> t1 is the MSB (:144-155)
 *   k>2: sorted endian targets, bit j = j-th smallest (:175-186)
 * K: nK matrices, each d x d column-major interleaved. */
void ref_dm_kraus_ref(double* dm_, int nq, int k, const int* targets, int nK, const double* K_) {
    cplx* dm = (cplx*)dm_;
    int64_t N = (int64_t)1 << nq;
    int d = 1 << k;
    int e[16], ord[16];
    for (int i = 0; i < k; ++i) ord[i] = e[i] = nq - 1 - targets[i];
    sort_ints(ord, k);
    const cplx* K = (const cplx*)K_;
    int64_t n = N >> k;
#pragma omp parallel
    {
        int64_t* ixs = (int64_t*)malloc(sizeof(int64_t) * d);
        cplx* B = (cplx*)malloc(sizeof(cplx) * d * d);
        cplx* O = (cplx*)malloc(sizeof(cplx) * d * d);
        cplx* T = (cplx*)malloc(sizeof(cplx) * d * d);
#pragma omp for schedule(static)
        for (int64_t ix = 0; ix < n; ++ix) {
            int64_t p = pad_bits(ix, ord, k);
            if (k == 1) {
                ixs[0] = p; ixs[1] = flip_bit(p, e[0]);
            } else if (k == 2) {
                ixs[0] = p;
                ixs[1] = flip_bit(p, e[1]);
                ixs[2] = flip_bit(p, e[0]);
                ixs[3] = flip_bit(flip_bit(p, e[0]), e[1]);
            } else {
                for (int j = 0; j < d; ++j) {
                    int64_t f = p;
                    for (int b = 0; b < k; ++b) if ((j >> b) & 1) f = flip_bit(f, ord[b]);
                    ixs[j] = f;
                }
            }
            for (int r = 0; r < d; ++r)
                for (int c = 0; c < d; ++c) B[r + c * d] = dm[ixs[r] + ixs[c] * N];
            kraus_block(K, nK, d, B, O, T);
            for (int r = 0; r < d; ++r)
                for (int c = 0; c < d; ++c) dm[ixs[r] + ixs[c] * N] = O[r + c * d];
        }
        free(ixs); free(B); free(O); free(T);
    }
}

/* Physical Kraus channel on the stored matrix S = rho^T (the layout the
 * reference's DM gate pass produces, dm_simulator.jl:130-140):
 *   S <- sum_m conj(K_m) S K_m^T   over every 2^k x 2^k block,
 * with the same block index order as ref_dm_kraus_ref.  Not a reference
 * loop: the reference updates only the block-diagonal blocks. */
void ref_dm_kraus_physical(double* dm_, int nq, int k, const int* targets, int nK, const double* K_) {
    cplx* dm = (cplx*)dm_;
    int64_t N = (int64_t)1 << nq;
    int d = 1 << k;
    int e[16], ord[16];
    for (int i = 0; i < k; ++i) ord[i] = e[i] = nq - 1 - targets[i];
    sort_ints(ord, k);
    cplx* Kc = (cplx*)malloc(sizeof(cplx) * nK * d * d);
    const cplx* K = (const cplx*)K_;
    for (int i = 0; i < nK * d * d; ++i) Kc[i] = cconj(K[i]);
    int64_t n = N >> k;
#pragma omp parallel
    {
        int64_t* rx = (int64_t*)malloc(sizeof(int64_t) * d);
        int64_t* cx = (int64_t*)malloc(sizeof(int64_t) * d);
        cplx* B = (cplx*)malloc(sizeof(cplx) * d * d);
        cplx* O = (cplx*)malloc(sizeof(cplx) * d * d);
        cplx* T = (cplx*)malloc(sizeof(cplx) * d * d);
#pragma omp for schedule(static)
        for (int64_t blk = 0; blk < n * n; ++blk) {
            int64_t bi = blk % n, bj = blk / n;
            int64_t pr = pad_bits(bi, ord, k), pc = pad_bits(bj, ord, k);
            for (int pass = 0; pass < 2; ++pass) {
                int64_t p = pass == 0 ? pr : pc;
                int64_t* o = pass == 0 ? rx : cx;
                if (k == 1) {
                    o[0] = p; o[1] = flip_bit(p, e[0]);
                } else if (k == 2) {
                    o[0] = p; o[1] = flip_bit(p, e[1]); o[2] = flip_bit(p, e[0]);
                    o[3] = flip_bit(flip_bit(p, e[0]), e[1]);
                } else {
                    for (int j = 0; j < d; ++j) {
                        int64_t f = p;
                        for (int b = 0; b < k; ++b) if ((j >> b) & 1) f = flip_bit(f, ord[b]);
                        o[j] = f;
                    }
                }
            }
            for (int r = 0; r < d; ++r)
                for (int c = 0; c < d; ++c) B[r + c * d] = dm[rx[r] + cx[c] * N];
            kraus_block(Kc, nK, d, B, O, T);   /* conj(K) B conj(K)^dagger = conj(K) B K^T */
            for (int r = 0; r < d; ++r)
                for (int c = 0; c < d; ++c) dm[rx[r] + cx[c] * N] = O[r + c * d];
        }
        free(rx); free(cx); free(B); free(O); free(T);
    }
    free(Kc);
}

/* dm_simulator.jl:190-224: Hermitian observable on a DM copy: for every
 * pair of slim indices, block(rows jxs, cols ixs) <- O^T * block.  The
 * target bits are the sorted endian targets (bit j = j-th smallest). */
void ref_dm_apply_hermitian(double* dm_, int nq, int k, const int* targets, const double* O_) {
    cplx* dm = (cplx*)dm_;
    int64_t N = (int64_t)1 << nq;
    int d = 1 << k;
    int ord[16];
    for (int i = 0; i < k; ++i) ord[i] = nq - 1 - targets[i];
    sort_ints(ord, k);
    const cplx* O = (const cplx*)O_;
    int64_t slim = N >> k;
#pragma omp parallel
    {
        int64_t* ixs = (int64_t*)malloc(sizeof(int64_t) * d);
        int64_t* jxs = (int64_t*)malloc(sizeof(int64_t) * d);
        cplx* E = (cplx*)malloc(sizeof(cplx) * d * d);
#pragma omp for schedule(static)
        for (int64_t raw = 0; raw < slim * slim; ++raw) {
            int64_t ix = raw / slim, jx = raw % slim;
            int64_t pi = pad_bits(ix, ord, k), pj = pad_bits(jx, ord, k);
            for (int j = 0; j < d; ++j) {
                int64_t fi = pi, fj = pj;
                for (int b = 0; b < k; ++b)
                    if ((j >> b) & 1) { fi = flip_bit(fi, ord[b]); fj = flip_bit(fj, ord[b]); }
                ixs[j] = fi; jxs[j] = fj;
            }
            for (int r = 0; r < d; ++r)
                for (int c = 0; c < d; ++c) E[r + c * d] = dm[jxs[r] + ixs[c] * N];
            for (int r = 0; r < d; ++r)
                for (int c = 0; c < d; ++c) {
                    /* (O^T E)[r, c] = sum_l O[l, r] E[l, c] */
                    cplx acc = cmul(O[0 + r * d], E[0 + c * d]);
                    for (int l = 1; l < d; ++l) acc = cadd(acc, cmul(O[l + r * d], E[l + c * d]));
                    dm[jxs[r] + ixs[c] * N] = acc;
                }
        }
        free(ixs); free(jxs); free(E);
    }
}

/* ------------------------------------------------------------------ */
/* Reductions                                                          */
/* ------------------------------------------------------------------ */

/* sv_simulator.jl:271 */
void ref_sv_probabilities(const double* sv_, int nq, double* out) {
    const cplx* sv = (const cplx*)sv_;
    int64_t N = (int64_t)1 << nq;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) out[i] = sv[i].re * sv[i].re + sv[i].im * sv[i].im;
}

/* dm_simulator.jl:262 */
void ref_dm_probabilities(const double* dm_, int nq, double* out) {
    const cplx* dm = (const cplx*)dm_;
    int64_t N = (int64_t)1 << nq;
    for (int64_t i = 0; i < N; ++i) out[i] = dm[i + i * N].re;
}

/* sum_i conj(a_i) b_i  (LinearAlgebra.dot, used by sv_simulator.jl:247) */
void ref_dot(const double* a_, const double* b_, int64_t n, double* out2) {
    const cplx* a = (const cplx*)a_;
    const cplx* b = (const cplx*)b_;
    double re = 0, im = 0;
#pragma omp parallel for reduction(+ : re, im) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        re += a[i].re * b[i].re + a[i].im * b[i].im;
        im += a[i].re * b[i].im - a[i].im * b[i].re;
    }
    out2[0] = re;
    out2[1] = im;
}

/* real(sum(diag(dm)))  (dm_simulator.jl:259) */
double ref_dm_trace_re(const double* dm_, int nq) {
    const cplx* dm = (const cplx*)dm_;
    int64_t N = (int64_t)1 << nq;
    double s = 0;
    for (int64_t i = 0; i < N; ++i) s += dm[i + i * N].re;
    return s;
}

/* result_types.jl:96-113: marginal probabilities over `targets`. */
void ref_marginal_probability(const double* probs, int nq, int nt, const int* targets, double* out) {
    int unused[64], nu = 0;
    for (int q = 0; q < nq; ++q) {
        int used = 0;
        for (int t = 0; t < nt; ++t) used |= targets[t] == q;
        if (!used) unused[nu++] = nq - 1 - q;
    }
    sort_ints(unused, nu);
    int64_t nout = (int64_t)1 << nt;
    int64_t ncomb = (int64_t)1 << nu;
#pragma omp parallel for schedule(static)
    for (int64_t ix = 0; ix < nout; ++ix) {
        int64_t p = pad_bits(ix, unused, nu);
        double s = 0;
        for (int64_t c = 0; c < ncomb; ++c) {
            int64_t f = p;
            for (int b = 0; b < nu; ++b) if ((c >> b) & 1) f = flip_bit(f, unused[b]);
            s += probs[f];
        }
        out[ix] = s;
    }
}

/* ------------------------------------------------------------------ */
/* Sampling                                                            */
/* ------------------------------------------------------------------ */

/* result_types.jl:8-54 (make_alias_table!), 1-based indices kept as the
 * reference stores them.  larges/smalls: length-n workspaces. */
void ref_make_alias_table(const double* w, int64_t n, double wsum, double* ap, int64_t* alias,
                          int64_t* larges, int64_t* smalls) {
    double ac = (double)n / wsum;
    for (int64_t i = 0; i < n; ++i) ap[i] = w[i] * ac;
    int64_t kl = 0, ks = 0;
    for (int64_t i = 1; i <= n; ++i) {
        double ai = ap[i - 1];
        if (ai > 1.0) larges[kl++] = i;
        else if (ai < 1.0) smalls[ks++] = i;
    }
    while (kl > 0 && ks > 0) {
        int64_t s = smalls[--ks];
        int64_t l = larges[--kl];
        alias[s - 1] = l;
        double al = ap[l - 1] = (ap[l - 1] - 1.0) + ap[s - 1];
        if (al > 1.0) larges[kl++] = l;
        else smalls[ks++] = l;
    }
    for (int64_t i = 0; i < ks; ++i) ap[smalls[i] - 1] = 1.0;
}

/* result_types.jl:69-72 with the draws supplied: idx1 is the 1-based
 * amplitude draw, u the uniform in [0,1).  Returns 0-based shots. */
void ref_alias_lookup(const double* ap, const int64_t* alias, int64_t nshots, const int64_t* idx1,
                      const double* u, int64_t* out) {
    for (int64_t s = 0; s < nshots; ++s) {
        int64_t i = idx1[s];
        out[s] = u[s] < ap[i - 1] ? i - 1 : alias[i - 1] - 1;
    }
}

/* result_types.jl:73-78, the n >= 30 branch: one StatsBase.sample(rng,
 * weights) per shot.  StatsBase.jl 0.34 (Project.toml:45, not vendored)
 * publishes sample(rng, wv::AbstractWeights) as a linear scan:
 * t = rand(rng) * sum(wv); i = 1; cw = wv[1]; while cw < t && i < n:
 * i += 1, cw += wv[i]; return i.  Here the uniforms u are supplied and the
 * shots are returned 0-based.  Sequential by construction (a running sum). */
void ref_direct_sample(const double* w, int64_t n, double wsum, int64_t nshots, const double* u, int64_t* out) {
    for (int64_t s = 0; s < nshots; ++s) {
        const double t = u[s] * wsum;
        int64_t i = 1;
        double cw = w[0];
        while (cw < t && i < n) {
            ++i;
            cw += w[i - 1];
        }
        out[s] = i - 1;
    }
}

/* BraketSimulator.jl:88-95: shot index -> bits, qubit 0 first. */
void ref_index_to_endian_bits(int64_t ix, int nq, int32_t* bits) {
    for (int q = 0; q < nq; ++q) bits[nq - 1 - q] = (int32_t)((ix >> q) & 1);
}

/* ------------------------------------------------------------------ */
/* Checkers for the device sampling path (not reference loops: the    */
/* reference's sampler is unseeded, src/result_types.jl:60).           */
/* ------------------------------------------------------------------ */

/* Philox-4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3" (SC'11): 10 rounds, multipliers 0xD2511F53 /
 * 0xCD9E8D57, Weyl key increments 0x9E3779B9 / 0xBB67AE85. */
void ref_philox4x32_10(const uint32_t in[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Shot s: index = top nq bits of (x1:x0), u = (x3:x2 >> 11) * 2^-53. */
void ref_shot_draws(uint64_t seed, int nq, int64_t first, int64_t n, int64_t* idx, double* u) {
    for (int64_t s = 0; s < n; ++s) {
        uint64_t ctr = (uint64_t)(first + s);
        uint32_t in[4] = { (uint32_t)ctr, (uint32_t)(ctr >> 32), 0, 0 }, o[4];
        ref_philox4x32_10(in, (uint32_t)seed, (uint32_t)(seed >> 32), o);
        uint64_t hi = ((uint64_t)o[1] << 32) | o[0];
        idx[s] = (int64_t)(hi >> (64 - nq));
        u[s] = (double)((((uint64_t)o[3] << 32) | o[2]) >> 11) * 0x1.0p-53;
    }
}

/* Exact-integer sweep alias table (the device's parallel build), restated
 * sequentially: W_i = floor(p_i * n * 2^32 / wsum), remainder to the first
 * largest weight, then the sweep closed form on inclusive prefix sums of
 * light deficits D and heavy excesses E.  alias is 0-based. */
void ref_alias_exact(const double* p, int64_t n, double wsum, double* ap, int32_t* alias) {
    const uint64_t C = (uint64_t)1 << 32;
    uint64_t* W = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint64_t* D = (uint64_t*)malloc(sizeof(uint64_t) * n);
    uint64_t* E = (uint64_t*)malloc(sizeof(uint64_t) * n);
    double scale = (double)n * (double)C / wsum;
    uint64_t tot = 0, mx = 0;
    int64_t arg = 0;
    for (int64_t i = 0; i < n; ++i) {
        double x = p[i] * scale;
        W[i] = x > 0 ? (uint64_t)x : 0;
        tot += W[i];
        if (W[i] > mx) { mx = W[i]; arg = i; }
    }
    W[arg] += (uint64_t)n * C - tot;
    uint64_t d = 0, e = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (W[i] <= C) d += C - W[i]; else e += W[i] - C;
        D[i] = d; E[i] = e;
    }
    for (int64_t i = 0; i < n; ++i) {
        uint64_t w = W[i];
        if (w <= C) {
            ap[i] = (double)w / (double)C;
            if (w == C) { alias[i] = (int32_t)i; continue; }
            uint64_t dex = D[i] - (C - w);
            int64_t lo = 0, hi = n;                  /* first k with E[k] > dex */
            while (lo < hi) { int64_t m = lo + (hi - lo) / 2; if (E[m] > dex) hi = m; else lo = m + 1; }
            alias[i] = (int32_t)(lo < n ? lo : i);
        } else {
            uint64_t ei = E[i];
            int64_t lo = 0, hi = n;                  /* first k with D[k] >= ei */
            while (lo < hi) { int64_t m = lo + (hi - lo) / 2; if (D[m] >= ei) hi = m; else lo = m + 1; }
            uint64_t r = lo < n ? C + ei - D[lo] : C;
            int64_t l2 = i + 1, h2 = n;              /* next heavy: first k > i with E[k] > ei */
            while (l2 < h2) { int64_t m = l2 + (h2 - l2) / 2; if (E[m] > ei) h2 = m; else l2 = m + 1; }
            int64_t nx = l2;
            if (nx >= n) { r = C; nx = i; }
            ap[i] = (double)r / (double)C;
            alias[i] = (int32_t)nx;
        }
    }
    free(W); free(D); free(E);
}

/* SplitMix64 (Steele, Lea, Flood, OOPSLA'14): the workload angle stream. */
void ref_splitmix64_doubles(uint64_t seed, int64_t n, double* out) {
    uint64_t x = seed;
    for (int64_t i = 0; i < n; ++i) {
        x += 0x9E3779B97F4A7C15ull;
        uint64_t z = x;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        out[i] = (double)(z >> 11) * 0x1.0p-53;
    }
}

/* Batched-execution shots (no counterpart in the reference, whose batch
 * calls the alias sampler per circuit, src/BraketSimulator.jl:347-395):
 * inverse CDF over the exact integer weights W_i = floor(p_i * 2^n * 2^32)
 * with p_i = re^2 + im^2 of the final amplitudes; shot s draws
 * Philox-4x32-10(counter (s, 0, 0, 0), key seed), r = (x1:x0),
 * t = floor(r * sum(W) / 2^64), and returns the first i with
 * W_0 + ... + W_i > t. */
void ref_cdf_shots(const double* amp, int nq, int64_t shots, uint64_t seed, int64_t* out) {
    const int64_t n = (int64_t)1 << nq;
    uint64_t* P = (uint64_t*)malloc(sizeof(uint64_t) * n);
    const double scale = (double)n * 4294967296.0;
    uint64_t run = 0;
    for (int64_t i = 0; i < n; ++i) {
        double re = amp[2 * i], im = amp[2 * i + 1];
        double re2 = re * re, im2 = im * im;
        double p = re2 + im2;
        double x = p * scale;
        run += x > 0 ? (uint64_t)x : 0;
        P[i] = run;
    }
    for (int64_t s = 0; s < shots; ++s) {
        uint32_t in[4] = { (uint32_t)s, (uint32_t)((uint64_t)s >> 32), 0, 0 }, o[4];
        ref_philox4x32_10(in, (uint32_t)seed, (uint32_t)(seed >> 32), o);
        uint64_t r = ((uint64_t)o[1] << 32) | o[0];
        uint64_t t = (uint64_t)(((unsigned __int128)r * run) >> 64);
        int64_t lo = 0, hi = n - 1;
        while (lo < hi) { int64_t m = (lo + hi) >> 1; if (P[m] > t) hi = m; else lo = m + 1; }
        out[s] = lo;
    }
    free(P);
}
