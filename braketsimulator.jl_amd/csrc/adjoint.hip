// Adjoint differentiation of <H> for a Pauli-sum H (new: the reference has
// the AdjointGradient type only, src/results.jl:90-160, and rejects the
// pragma, src/pragmas.jl:67-68).
//
//   forward:  psi = U_L ... U_1 |0>          (fused tile passes)
//   lambda = H psi,  <H> = Re <psi|lambda>
//   backward, k = L..1:
//       psi    <- U_k^dagger psi             (now psi_{k-1})
//       g_p    += 2 Re <lambda| dU_k/dtheta_p |psi>
//       lambda <- U_k^dagger lambda
// One kernel per op reads (psi, lambda) once, un-applies U_k to both and
// reduces the <lambda|dU|psi> products of the op's parameters in the same
// pass (wavefront shuffle tree -> LDS -> per-block partials -> one ordered
// sum), so the gradient is deterministic.
#include <atomic>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fused.h"
#include "reduce.h"

namespace bsim {

constexpr int AB = 256;  // adjoint block

__device__ __forceinline__ double wave_sum_adj(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <int P>
__device__ __forceinline__ void block_sum_adj(double (&v)[P > 0 ? P : 1], double* lds, double* out) {
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < P; ++i) {
        double s = wave_sum_adj(v[i]);
        if (lane == 0) lds[i * 8 + wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < P; ++i) {
            double s = 0;
            for (int w = 0; w < AB / 64; ++w) s += lds[i * 8 + w];
            out[i] = s;
        }
    }
}

struct AdjArgs {
    int nins;
    int ins[16];
    uint64_t cval;
    int64_t ngroups;
    int t0, t1;
};

template <int K>
struct AdjMats {
    double2 udag[1 << (2 * K)];
    double2 d[3][1 << (2 * K)];
};

// dense K = 1, 2 (with controls)
template <int K, int P>
__global__ __launch_bounds__(AB) void k_adj_dense(double2* __restrict__ psi, double2* __restrict__ lam, AdjArgs a,
                                                  AdjMats<K> M, double* __restrict__ partials) {
    constexpr int D = 1 << K;
    __shared__ double lds[3 * 8];
    double acc[P > 0 ? P : 1];
#pragma unroll
    for (int p = 0; p < (P > 0 ? P : 1); ++p) acc[p] = 0;
    const uint64_t b0 = 1ull << a.t0, b1 = K > 1 ? (1ull << a.t1) : 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.ngroups; g += stride) {
        uint64_t base = (uint64_t)g;
        for (int i = 0; i < a.nins; ++i) base = pad_bit(base, a.ins[i]);
        base |= a.cval;
        uint64_t off[D];
        double2 x[D], l[D];
#pragma unroll
        for (int j = 0; j < D; ++j) {
            off[j] = base | ((j & 1) ? b0 : 0) | ((j & 2) ? b1 : 0);
            x[j] = psi[off[j]];
            l[j] = lam[off[j]];
        }
        double2 xn[D], ln[D];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            double2 s = cmul(M.udag[r], x[0]), t = cmul(M.udag[r], l[0]);
#pragma unroll
            for (int c = 1; c < D; ++c) {
                s = cfma(M.udag[r + D * c], x[c], s);
                t = cfma(M.udag[r + D * c], l[c], t);
            }
            xn[r] = s;
            ln[r] = t;
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
#pragma unroll
            for (int r = 0; r < D; ++r) {
                double2 mu = cmul(M.d[p][r], xn[0]);
#pragma unroll
                for (int c = 1; c < D; ++c) mu = cfma(M.d[p][r + D * c], xn[c], mu);
                acc[p] += l[r].x * mu.x + l[r].y * mu.y;  // Re(conj(lambda) mu)
            }
        }
#pragma unroll
        for (int j = 0; j < D; ++j) {
            psi[off[j]] = xn[j];
            lam[off[j]] = ln[j];
        }
    }
    if (P > 0) block_sum_adj<P>(acc, lds, partials + (size_t)blockIdx.x * 3);
}

// diagonal, k <= 4 target bits anywhere, with controls
template <int P>
__global__ __launch_bounds__(AB) void k_adj_diag(double2* __restrict__ psi, double2* __restrict__ lam, AdjArgs a,
                                                 int k, int4 tb, AdjMats<2> M, double* __restrict__ partials) {
    __shared__ double lds[3 * 8];
    double acc[P > 0 ? P : 1];
#pragma unroll
    for (int p = 0; p < (P > 0 ? P : 1); ++p) acc[p] = 0;
    const int t[4] = { tb.x, tb.y, tb.z, tb.w };
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.ngroups; g += stride) {
        uint64_t idx = (uint64_t)g;
        for (int i = 0; i < a.nins; ++i) idx = pad_bit(idx, a.ins[i]);
        idx |= a.cval;
        int j = 0;
        for (int b = 0; b < k; ++b) j |= (int)((idx >> t[b]) & 1) << b;
        const double2 x = cmul(M.udag[j], psi[idx]);
        const double2 l = lam[idx];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const double2 mu = cmul(M.d[p][j], x);
            acc[p] += l.x * mu.x + l.y * mu.y;
        }
        psi[idx] = x;
        lam[idx] = cmul(M.udag[j], l);
    }
    if (P > 0) block_sum_adj<P>(acc, lds, partials + (size_t)blockIdx.x * 3);
}

// lam = H psi by gathers, one per word (accumulate: added to lam)
__global__ __launch_bounds__(AB) void k_pauli_sum(const double2* __restrict__ psi, double2* __restrict__ lam,
                                                  int64_t len, const PWord* __restrict__ words, int nw,
                                                  int accumulate) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        double2 acc = accumulate ? lam[i] : make_double2(0, 0);
        for (int w = 0; w < nw; ++w) acc = cfma(pphase((uint64_t)i, words[w]), psi[(uint64_t)i ^ words[w].f], acc);
        lam[i] = acc;
    }
}

// lam = H psi as tile passes.  A word maps amplitude i ^ f to i with a phase
// that depends on i: phase(i) = (-1)^(n0 ^ parity(i & m)) (cr + i ci) with
// m = y ^ z, n0 = parity(y) and cr + i ci = c i^ny (pphase: the parity of
// popcount(~i & y) + popcount(i & z)).  Words are grouped by f.  A pass
// holds one tile of psi (PS_T local bits) in LDS and adds, for every group
// whose f lies in the tile's bits, the group's phase sum times psi[i ^ f]
// read from LDS; the first pass's tile is the low PS_T bits (contiguous),
// later ones bits 0..2 and PS_T - 3 bits chosen for the groups left.
// The first pass writes lam (and, when asked, the working copy of psi),
// later ones add to it.  Groups that no tile covers are added by
// k_pauli_sum.  Traffic per pass: psi read, lam written (and read after the
// first): 2 or 3 x 16 B x 2^n, against one gather of psi per word.
constexpr int PS_T = 11, PS_TB = 256, PS_R = (1 << PS_T) / PS_TB;  // 8 amplitudes per thread

struct PSWord {
    uint64_t m;
    double cr, ci;
    int32_t n0, pad_;
};
struct PSGroup {
    uint32_t fl;  // flip mask in local index bits
    int32_t w0, w1;
    int32_t uniform;  // every word has m = 0: the phase sum is the same for all amplitudes
};
struct PSPass {
    int32_t lbits[PS_T];  // vector bit of local bit b, ascending
    uint64_t lmask;
    int32_t g0, g1;
};

__global__ __launch_bounds__(PS_TB) void k_pauli_tiles(const double2* __restrict__ psi, double2* __restrict__ lam,
                                                       double2* __restrict__ work, int vq, int first,
                                                       const PSPass* __restrict__ pass,
                                                       const PSGroup* __restrict__ groups,
                                                       const PSWord* __restrict__ words,
                                                       double* __restrict__ dot) {
    extern __shared__ double2 ps_tile[];  // 2^PS_T amplitudes
    const int t = threadIdx.x;
    uint64_t lo = 0;  // local bits 0..7 of the thread's amplitudes
    for (int b = 0; b < 8; ++b)
        if ((t >> b) & 1) lo |= 1ull << pass->lbits[b];
    uint64_t base = 0;
    {
        const uint64_t lm = pass->lmask, w = blockIdx.x;
        for (int bit = 0, k = 0; bit < vq; ++bit)
            if (!((lm >> bit) & 1)) base |= ((w >> k++) & 1) << bit;
    }
    const uint64_t gt = base | lo;
    uint64_t sh[PS_R];  // the same for every thread: scalar
#pragma unroll
    for (int r = 0; r < PS_R; ++r) {
        uint64_t h = 0;
#pragma unroll
        for (int b = 8; b < PS_T; ++b)
            if ((r >> (b - 8)) & 1) h |= 1ull << pass->lbits[b];
        sh[r] = h;
    }
    uint64_t g[PS_R];
    double2 acc[PS_R];
#pragma unroll
    for (int r = 0; r < PS_R; ++r) {
        g[r] = gt | sh[r];
        const double2 v = psi[g[r]];
        ps_tile[t + (r << 8)] = v;
        if (work) work[g[r]] = v;
        acc[r] = first ? make_double2(0.0, 0.0) : lam[g[r]];
    }
    __syncthreads();
    for (int k = pass->g0; k < pass->g1; ++k) {
        const PSGroup gr = groups[k];
        if (gr.uniform) {
            double sr = 0.0, si = 0.0;
            for (int w = gr.w0; w < gr.w1; ++w) {
                const double sg = words[w].n0 ? -1.0 : 1.0;
                sr += sg * words[w].cr;
                si += sg * words[w].ci;
            }
            const double2 sp = make_double2(sr, si);
#pragma unroll
            for (int r = 0; r < PS_R; ++r) acc[r] = cfma(sp, ps_tile[(t + (r << 8)) ^ gr.fl], acc[r]);
        } else {
            // parity(g_r & m) = parity(gt & m) ^ parity(sh[r] & m): a sign per
            // thread and word times a scalar per word and row
            double2 sp[PS_R];
#pragma unroll
            for (int r = 0; r < PS_R; ++r) sp[r] = make_double2(0.0, 0.0);
            for (int w = gr.w0; w < gr.w1; ++w) {
                const PSWord wd = words[w];
                const double sg = ((__popcll(gt & wd.m) + wd.n0) & 1) ? -1.0 : 1.0;
                if (wd.ci == 0.0) {
#pragma unroll
                    for (int r = 0; r < PS_R; ++r) {
                        const double u = (__popcll(sh[r] & wd.m) & 1) ? -wd.cr : wd.cr;
                        sp[r].x = fma(sg, u, sp[r].x);
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < PS_R; ++r) {
                        const bool hn = (__popcll(sh[r] & wd.m) & 1) != 0;
                        sp[r].x = fma(sg, hn ? -wd.cr : wd.cr, sp[r].x);
                        sp[r].y = fma(sg, hn ? -wd.ci : wd.ci, sp[r].y);
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < PS_R; ++r) acc[r] = cfma(sp[r], ps_tile[(t + (r << 8)) ^ gr.fl], acc[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < PS_R; ++r) lam[g[r]] = acc[r];
    if (dot) {
        // the last pass: <psi|lam> over the tile, dot[2 b .. 2 b + 1] (fixed order)
        __shared__ double red[2][PS_TB / 64];
        double dr = 0.0, di = 0.0;
#pragma unroll
        for (int r = 0; r < PS_R; ++r) {
            const double2 x = ps_tile[t + (r << 8)];
            dr += x.x * acc[r].x + x.y * acc[r].y;
            di += x.x * acc[r].y - x.y * acc[r].x;
        }
        dr = wave_sum_adj(dr);
        di = wave_sum_adj(di);
        if ((t & 63) == 0) {
            red[0][t >> 6] = dr;
            red[1][t >> 6] = di;
        }
        __syncthreads();
        if (t < 2) {
            double v = 0.0;
            for (int w = 0; w < PS_TB / 64; ++w) v += red[t][w];
            dot[2 * (size_t)blockIdx.x + t] = v;
        }
    }
}

// out[i] = sum over b of part[2 b + i], i = 0, 1 (one block each, fixed order)
__global__ __launch_bounds__(256) void k_ps_dot_finish(const double* __restrict__ part, int64_t nb,
                                                       double* __restrict__ out) {
    __shared__ double sh[256];
    double v = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += 256) v += part[2 * b + blockIdx.x];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// The passes of lam = H psi over a 2^vq vector (vq >= PS_T): groups of words
// by flip mask, each group in the first pass whose tile holds its f; `rest`
// gets the words of groups no tile holds (at most PS_T - 3 bits of f above
// bit 2 fit one).
struct PauliTilePlan {
    std::vector<PSPass> passes;
    std::vector<PSGroup> groups;
    std::vector<PSWord> words;
    std::vector<PWord> rest;
};

static void pauli_tile_plan(int vq, const std::vector<PWord>& in, PauliTilePlan& out) {
    std::vector<uint64_t> fs;
    std::vector<std::vector<int>> members;
    for (int w = 0; w < (int)in.size(); ++w) {
        size_t k = 0;
        while (k < fs.size() && fs[k] != in[w].f) ++k;
        if (k == fs.size()) {
            fs.push_back(in[w].f);
            members.emplace_back();
        }
        members[k].push_back(w);
    }
    std::vector<int> left(fs.size());
    for (size_t k = 0; k < fs.size(); ++k) left[k] = (int)k;
    uint64_t tm = (1ull << PS_T) - 1;  // the first tile: the low PS_T bits
    const uint64_t all = vq >= 64 ? ~0ull : (1ull << vq) - 1;
    while (!left.empty()) {
        PSPass P;
        memset(&P, 0, sizeof(P));
        P.lmask = tm;
        for (int bit = 0, b = 0; bit < vq; ++bit)
            if ((tm >> bit) & 1) P.lbits[b++] = bit;
        P.g0 = (int)out.groups.size();
        std::vector<int> keep;
        for (int k : left) {
            if (fs[k] & ~tm) {
                keep.push_back(k);
                continue;
            }
            PSGroup G;
            G.fl = 0;
            for (int b = 0; b < PS_T; ++b)
                if ((fs[k] >> P.lbits[b]) & 1) G.fl |= 1u << b;
            G.w0 = (int)out.words.size();
            G.uniform = 1;
            for (int w : members[k]) {
                const PWord& x = in[w];
                PSWord y;
                y.m = x.y ^ x.z;
                y.n0 = __builtin_popcountll(x.y) & 1;
                y.pad_ = 0;
                const double c = x.c;
                switch (x.ny & 3) {
                    case 0: y.cr = c; y.ci = 0.0; break;
                    case 1: y.cr = 0.0; y.ci = c; break;
                    case 2: y.cr = -c; y.ci = 0.0; break;
                    default: y.cr = 0.0; y.ci = -c; break;
                }
                G.uniform = G.uniform && y.m == 0;
                out.words.push_back(y);
            }
            G.w1 = (int)out.words.size();
            out.groups.push_back(G);
        }
        P.g1 = (int)out.groups.size();
        if (P.g1 == P.g0) break;  // no group fits the tile
        out.passes.push_back(P);
        left.swap(keep);
        // the next tile: bits 0..2 (coalesced runs) and the flip bits of the
        // groups left, as many whole groups as fit, then the lowest free bits
        tm = 7;
        for (int k : left) {
            const uint64_t need = fs[k] & ~7ull;
            if (__builtin_popcountll(tm | need) <= PS_T) tm |= need;
        }
        for (int bit = 3; bit < vq && __builtin_popcountll(tm) < PS_T; ++bit) tm |= 1ull << bit;
        tm &= all;
    }
    for (int k : left)
        for (int w : members[k]) out.rest.push_back(in[w]);
}

// One block per parameter slot: partials[slot][block][3] -> out[slot][3].
// Each thread sums a fixed strided subset, then a fixed tree: deterministic.
__global__ __launch_bounds__(256) void k_adj_finish(const double* __restrict__ partials, int nblocks,
                                                    double* __restrict__ out) {
    __shared__ double sh[3][256];
    const double* p = partials + (size_t)blockIdx.x * nblocks * 3;
    double s[3] = { 0, 0, 0 };
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int q = 0; q < 3; ++q) s[q] += p[(size_t)b * 3 + q];
    for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3) out[(size_t)blockIdx.x * 3 + threadIdx.x] = sh[threadIdx.x][0];
}

static std::atomic<bool> g_ps_shmem{ false };

// lam = H psi on h's vector with the plan `pt`: its tables go to
// dev + off_*, the gather words to dwords; `work` (optional) receives a
// copy of psi.
static int pauli_sum_passes(bsim_state* h, const PauliTilePlan& pt, char* dev, size_t off_pass, size_t off_grp,
                            size_t off_wd, PWord* dwords, double2* lam, double2* work, double* dot2) {
    const int64_t len = h->len;
    // <psi|lam> from the last tile pass when no gather follows it
    const bool fused_dot = !pt.passes.empty() && pt.rest.empty();
    double* part = nullptr;
    if (fused_dot) {
        if (int rc = ensure_ws(h->red, sizeof(double) * 2 * ((size_t)1 << (h->vq - PS_T)))) return rc;
        part = (double*)h->red.ptr;
    }
    if (!pt.passes.empty()) {
        if (!g_ps_shmem) {
            BSIM_HIP(hipFuncSetAttribute((const void*)k_pauli_tiles, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)(sizeof(double2) << PS_T)));
            g_ps_shmem = true;
        }
        BSIM_HIP(hipMemcpyAsync(dev + off_pass, pt.passes.data(), sizeof(PSPass) * pt.passes.size(),
                                hipMemcpyHostToDevice, h->stream));
        BSIM_HIP(hipMemcpyAsync(dev + off_grp, pt.groups.data(), sizeof(PSGroup) * pt.groups.size(),
                                hipMemcpyHostToDevice, h->stream));
        BSIM_HIP(hipMemcpyAsync(dev + off_wd, pt.words.data(), sizeof(PSWord) * pt.words.size(),
                                hipMemcpyHostToDevice, h->stream));
        const unsigned grid = (unsigned)(1ull << (h->vq - PS_T));
        for (size_t p = 0; p < pt.passes.size(); ++p) {
            const bool last = p + 1 == pt.passes.size();
            hipLaunchKernelGGL(k_pauli_tiles, dim3(grid), dim3(PS_TB), sizeof(double2) << PS_T, h->stream, h->d, lam,
                               p == 0 ? work : nullptr, h->vq, p == 0 ? 1 : 0, (const PSPass*)(dev + off_pass) + p,
                               (const PSGroup*)(dev + off_grp), (const PSWord*)(dev + off_wd),
                               fused_dot && last ? part : nullptr);
            BSIM_HIP(hipGetLastError());
        }
        if (fused_dot) {
            hipLaunchKernelGGL(k_ps_dot_finish, dim3(2), dim3(256), 0, h->stream, (const double*)part, (int64_t)grid,
                               dot2);
            BSIM_HIP(hipGetLastError());
            return BSIM_OK;
        }
    } else if (work) {
        BSIM_HIP(hipMemcpyAsync(work, h->d, sizeof(double2) * (size_t)len, hipMemcpyDeviceToDevice, h->stream));
    }
    if (!pt.rest.empty() || pt.passes.empty()) {
        BSIM_HIP(hipMemcpyAsync(dwords, pt.rest.data(), sizeof(PWord) * pt.rest.size(), hipMemcpyHostToDevice,
                                h->stream));
        const int64_t g = grid_for(len, AB, (int64_t)device_cu_count() * 32);
        hipLaunchKernelGGL(k_pauli_sum, dim3((unsigned)g), dim3(AB), 0, h->stream, h->d, lam, len, dwords,
                           (int)pt.rest.size(), pt.passes.empty() ? 0 : 1);
        BSIM_HIP(hipGetLastError());
    }
    return dot_device(h->d, lam, len, h->red, h->stream, dot2);
}

static void conj_transpose(const std::vector<double2>& m, int D, double2* out) {
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) out[r + D * c] = make_double2(m[c + D * r].x, -m[c + D * r].y);
}

static AdjArgs adj_args(int vq, const GateSpec& g, bool include_targets) {
    AdjArgs a;
    memset(&a, 0, sizeof(a));
    uint64_t mask = g.cmask;
    if (include_targets)
        for (int b = 0; b < g.k; ++b) mask |= 1ull << g.tb[b];
    int n = 0;
    for (int bit = 0; bit < vq; ++bit)
        if (mask & (1ull << bit)) a.ins[n++] = bit;
    a.nins = n;
    a.cval = g.cval;
    a.ngroups = (int64_t)1 << (vq - n);
    a.t0 = g.k > 0 ? g.tb[0] : 0;
    a.t1 = g.k > 1 ? g.tb[1] : 0;
    return a;
}

// One backward step for eops[k]; the op's per-block parameter products land
// in partials[block][0..P) (the caller's slot of the partials buffer).
static int adjoint_step(bsim_state* h, double2* psi, double2* lam, const EOp& e, double* partials, int grid) {
    const GateSpec& g = e.g;
    const int P = e.nparams;
    const int vq = h->vq;
    if (e.kind == E_DENSE && g.k <= 2) {
        AdjArgs a = adj_args(vq, g, true);
        const int D = 1 << g.k;
        if (g.k == 1) {
            AdjMats<1> M;
            memset(&M, 0, sizeof(M));
            conj_transpose(g.m, D, M.udag);
            for (int p = 0; p < P; ++p)
                for (int x = 0; x < D * D; ++x) M.d[p][x] = e.dm[p][x];
            switch (P) {
                case 0: hipLaunchKernelGGL((k_adj_dense<1, 0>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, M, partials); break;
                case 1: hipLaunchKernelGGL((k_adj_dense<1, 1>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, M, partials); break;
                case 2: hipLaunchKernelGGL((k_adj_dense<1, 2>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, M, partials); break;
                default: hipLaunchKernelGGL((k_adj_dense<1, 3>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, M, partials); break;
            }
        } else {
            AdjMats<2> M;
            memset(&M, 0, sizeof(M));
            conj_transpose(g.m, D, M.udag);
            for (int p = 0; p < P; ++p)
                for (int x = 0; x < D * D; ++x) M.d[p][x] = e.dm[p][x];
            switch (P) {
                case 0: hipLaunchKernelGGL((k_adj_dense<2, 0>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, M, partials); break;
                case 1: hipLaunchKernelGGL((k_adj_dense<2, 1>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, M, partials); break;
                case 2: hipLaunchKernelGGL((k_adj_dense<2, 2>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, M, partials); break;
                default: hipLaunchKernelGGL((k_adj_dense<2, 3>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, M, partials); break;
            }
        }
    } else if (e.kind == E_DIAG && g.k <= 4) {
        AdjArgs a = adj_args(vq, g, false);
        AdjMats<2> M;  // 16 diagonal entries fit the 4x4 slots
        memset(&M, 0, sizeof(M));
        const int D = 1 << g.k;
        for (int j = 0; j < D; ++j) M.udag[j] = make_double2(g.m[j].x, -g.m[j].y);
        for (int p = 0; p < P; ++p)
            for (int j = 0; j < D; ++j) M.d[p][j] = e.dm[p][j + D * j];  // dU of a diagonal gate is diagonal
        int4 tb = make_int4(g.tb[0], g.k > 1 ? g.tb[1] : 0, g.k > 2 ? g.tb[2] : 0, g.k > 3 ? g.tb[3] : 0);
        switch (P) {
            case 0: hipLaunchKernelGGL((k_adj_diag<0>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, g.k, tb, M, partials); break;
            case 1: hipLaunchKernelGGL((k_adj_diag<1>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, g.k, tb, M, partials); break;
            case 2: hipLaunchKernelGGL((k_adj_diag<2>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, g.k, tb, M, partials); break;
            default: hipLaunchKernelGGL((k_adj_diag<3>), dim3(grid), dim3(AB), 0, h->stream, psi, lam, a, g.k, tb, M, partials); break;
        }
    } else {
        if (P > 0) return fail(BSIM_ENOTSUP, "adjoint: parametrized gates act on at most 2 dense or 4 diagonal targets");
        // un-apply a parameter-free gate to both vectors with the per-gate kernels
        GateSpec inv = g;
        if (g.diag) {
            for (auto& z : inv.m) z.y = -z.y;
        } else {
            conj_transpose(g.m, 1 << g.k, inv.m.data());
        }
        if (int rc = launch_gate(psi, vq, inv, h->scratch, h->stream)) return rc;
        if (int rc = launch_gate(lam, vq, inv, h->scratch, h->stream)) return rc;
        // the staged matrix is reused by the second launch: wait before it changes
        BSIM_HIP(hipStreamSynchronize(h->stream));
        return BSIM_OK;
    }
    BSIM_HIP(hipGetLastError());
    return BSIM_OK;
}


// ---- density matrices: adjoint with checkpointed recomputation -------------
//
// The forward map of a noisy circuit is not invertible, so the backward sweep
// cannot uncompute rho.  The forward pass keeps rho every `s` ops
// (s = ceil(sqrt(D)) by default, D ops); the backward sweep rebuilds the
// states of one segment from its checkpoint, then walks the segment in
// reverse: for a parameterized gate U(theta) with derivative dU,
//   d<H>/dtheta = 2 Re tr(Lambda dU rho U^dagger),
// and Lambda <- E^dagger(Lambda) for every op E (U^dagger . U for gates, the
// same channel with K^dagger for Kraus maps; PhaseFlip is self-adjoint).
// Lambda starts as H.  Storage of every operator is the transpose (the
// density-matrix convention), so tr(Lambda X) = conj(vec Lambda) . vec X for
// Hermitian Lambda.

// Lambda = H: stored element (r, c) = H[c][r] = sum_w [r == c ^ f_w] phase_w(c)
__global__ __launch_bounds__(AB) void k_dm_pauli_op(double2* __restrict__ lam, int nq, const PWord* __restrict__ words,
                                                    int nw) {
    const int64_t N = (int64_t)1 << nq, len = N * N;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += stride) {
        const uint64_t r = (uint64_t)i & (uint64_t)(N - 1), c = (uint64_t)i >> nq;
        double2 acc = make_double2(0, 0);
        for (int w = 0; w < nw; ++w)
            if (r == (c ^ words[w].f)) acc = cadd(acc, pphase(c, words[w]));
        lam[i] = acc;
    }
}

static std::atomic<int> g_checkpoint_every{ 0 };  // 0: ceil(sqrt(ops))
static std::atomic<int> g_adj_fusion{ 1 };        // state vectors: fused tile backward sweep

struct DmParam {
    int pidx;
    GateSpec dcol, urow, proj;  // dU on the column bits, conj(U) on the row bits, control projector
    bool ctl;
    // 1-qubit gate without controls: dU (column bit) x conj(U) (row bit) as
    // one 4x4 map on the two bits, dotted with Lambda in one pass
    bool one_pass = false;
    GateSpec sop;
};
struct DmStep {
    std::vector<EOp> fwd, adj;
    std::vector<DmParam> params;
};

static int dm_steps(int nq, const bsim_op* ops, int n_ops, int n_params, std::vector<DmStep>& st) {
    st.assign(n_ops, DmStep());
    const int vq = 2 * nq;
    std::vector<double> buf;
    for (int i = 0; i < n_ops; ++i) {
        const bsim_op& op = ops[i];
        if (int rc = lower_ops(nq, BSIM_DM, &op, 1, st[i].fwd, false)) return rc;
        bsim_op a = op;
        a.n_params = 0;
        const int k = op.n_targets, D = 1 << k;
        if (op.type == BSIM_OP_GATE) {
            buf.resize(2 * (size_t)D * D);
            for (int r = 0; r < D; ++r)
                for (int c = 0; c < D; ++c) {
                    buf[2 * (r + (size_t)D * c)] = op.matrix[2 * (c + (size_t)D * r)];
                    buf[2 * (r + (size_t)D * c) + 1] = -op.matrix[2 * (c + (size_t)D * r) + 1];
                }
            a.matrix = buf.data();
        } else if (op.type == BSIM_OP_DIAG) {
            buf.resize(2 * (size_t)D);
            for (int j = 0; j < D; ++j) {
                buf[2 * j] = op.matrix[2 * j];
                buf[2 * j + 1] = -op.matrix[2 * j + 1];
            }
            a.matrix = buf.data();
        } else if (op.type == BSIM_OP_KRAUS_REF || op.type == BSIM_OP_KRAUS_PHYS) {
            buf.resize(2 * (size_t)op.n_kraus * D * D);
            for (int m = 0; m < op.n_kraus; ++m) {
                const double* K = op.matrix + 2 * (size_t)m * D * D;
                double* A = buf.data() + 2 * (size_t)m * D * D;
                for (int r = 0; r < D; ++r)
                    for (int c = 0; c < D; ++c) {
                        A[2 * (r + (size_t)D * c)] = K[2 * (c + (size_t)D * r)];
                        A[2 * (r + (size_t)D * c) + 1] = -K[2 * (c + (size_t)D * r) + 1];
                    }
            }
            a.matrix = buf.data();
        } else if (op.type == BSIM_OP_SUPEROP) {  // adjoint map: S^dagger
            const size_t DD = (size_t)D * D;
            buf.resize(2 * DD * DD);
            for (size_t r = 0; r < DD; ++r)
                for (size_t c = 0; c < DD; ++c) {
                    buf[2 * (r + DD * c)] = op.matrix[2 * (c + DD * r)];
                    buf[2 * (r + DD * c) + 1] = -op.matrix[2 * (c + DD * r) + 1];
                }
            a.matrix = buf.data();
        }  // PhaseFlip: elementwise real factors, its own adjoint
        if (int rc = lower_ops(nq, BSIM_DM, &a, 1, st[i].adj, false)) return rc;
        if (op.n_params < 0 || op.n_params > 3) return fail(BSIM_EINVAL, "op: n_params must be 0..3");
        if (op.n_params > 0 && op.type != BSIM_OP_GATE)
            return fail(BSIM_EINVAL, "parameterized ops must be dense gates (BSIM_OP_GATE)");
        for (int j = 0; j < op.n_params; ++j) {
            DmParam P;
            P.pidx = op.param_index[j];
            if (P.pidx < 0 || P.pidx >= n_params) return fail(BSIM_EINVAL, "param_index out of range");
            if (!op.dmatrix[j]) return fail(BSIM_EINVAL, "op: dmatrix is null");
            int tc[8], tr[8], cc[8], cr[8];
            for (int b = 0; b < k; ++b) {
                tc[b] = op.targets[b];
                tr[b] = op.targets[b] + nq;
            }
            for (int b = 0; b < op.n_controls; ++b) {
                cc[b] = op.controls[b];
                cr[b] = op.controls[b] + nq;
            }
            if (int rc = make_spec(vq, k, tc, op.n_controls, cc, op.control_values, op.dmatrix[j], false, false, false,
                                   P.dcol))
                return rc;
            if (int rc = make_spec(vq, k, tr, op.n_controls, cr, op.control_values, op.matrix, false, true, false, P.urow))
                return rc;
            P.ctl = op.n_controls > 0;
            if (k == 1 && !P.ctl) {
                // S[(a_c + 2 a_r), (b_c + 2 b_r)] = dU[a_c, b_c] conj(U[a_r, b_r]); matrix bit 0 is
                // the column bit (vector qubit q), bit 1 the row bit (vector qubit nq + q)
                double S[32];
                const double* dU = op.dmatrix[j];
                const double* U = op.matrix;
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c) {
                        const int ac = r & 1, ar = r >> 1, bc = c & 1, br = c >> 1;
                        const double dx = dU[2 * (ac + 2 * bc)], dy = dU[2 * (ac + 2 * bc) + 1];
                        const double ux = U[2 * (ar + 2 * br)], uy = -U[2 * (ar + 2 * br) + 1];
                        S[2 * (r + 4 * c)] = dx * ux - dy * uy;
                        S[2 * (r + 4 * c) + 1] = dx * uy + dy * ux;
                    }
                const int t2[2] = { op.targets[0], op.targets[0] + nq };
                if (int rc = make_spec(vq, 2, t2, 0, nullptr, nullptr, S, false, false, false, P.sop)) return rc;
                P.one_pass = true;
            }
            if (P.ctl) {  // dU of a controlled gate vanishes off the control subspace
                const int nc = op.n_controls, DC = 1 << nc;
                std::vector<double> d(2 * (size_t)DC, 0.0);
                for (int jj = 0; jj < DC; ++jj) {
                    bool on = true;
                    for (int b = 0; b < nc; ++b) on = on && (((jj >> b) & 1) == (op.control_values[b] ? 1 : 0));
                    d[2 * jj] = on ? 1.0 : 0.0;
                }
                if (int rc = make_spec(vq, nc, cc, 0, nullptr, nullptr, d.data(), true, false, false, P.proj)) return rc;
            }
            st[i].params.push_back(std::move(P));
        }
    }
    return BSIM_OK;
}

// Re sum conj(Lambda) (S rho) over the 4-element groups of S's two bits:
// 2 Re tr(Lambda dU rho U^dagger) / 2 for a 1-qubit gate in one pass.
constexpr int DM_GRID = 1024;
__global__ __launch_bounds__(AB) void k_dm_grad(const double2* __restrict__ rho, const double2* __restrict__ lam,
                                                AdjArgs a, AdjMats<2> M, double* __restrict__ partials) {
    __shared__ double lds[3 * 8];
    double acc[1] = { 0.0 };
    const uint64_t b0 = 1ull << a.t0, b1 = 1ull << a.t1;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < a.ngroups; g += stride) {
        uint64_t base = (uint64_t)g;
        for (int i = 0; i < a.nins; ++i) base = pad_bit(base, a.ins[i]);
        base |= a.cval;
        double2 x[4], l[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t off = base | ((j & 1) ? b0 : 0) | ((j & 2) ? b1 : 0);
            x[j] = rho[off];
            l[j] = lam[off];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            double2 mu = cmul(M.udag[r], x[0]);
#pragma unroll
            for (int c = 1; c < 4; ++c) mu = cfma(M.udag[r + 4 * c], x[c], mu);
            acc[0] += l[r].x * mu.x + l[r].y * mu.y;
        }
    }
    block_sum_adj<1>(acc, lds, partials + (size_t)blockIdx.x * 3);
}

__global__ __launch_bounds__(256) void k_dm_grad_finish(const double* __restrict__ partials, int nblocks,
                                                        double* __restrict__ out) {
    __shared__ double sh[256];
    double v = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) v += partials[(size_t)b * 3];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sh[0];
        out[1] = 0.0;
    }
}

static int apply_eops(bsim_state* h, double2* d, const std::vector<EOp>& e) {
    for (const EOp& x : e)
        if (int rc = launch_eop(h, d, x)) return rc;
    return BSIM_OK;
}

static int dm_adjoint(bsim_state* h, const bsim_op* ops, int n_ops, const std::vector<PWord>& words, int n_params,
                      double* grad, double* expval) {
    std::vector<DmStep> st;
    if (int rc = dm_steps(h->nq, ops, n_ops, n_params, st)) return rc;
    const int64_t len = h->len;
    const size_t bytes = sizeof(double2) * (size_t)len;
    const int D = n_ops;
    const int every = g_checkpoint_every.load();
    int s = every > 0 ? every : 1;
    if (every <= 0)
        while ((int64_t)s * s < D) ++s;
    const int nseg = D > 0 ? (D + s - 1) / s : 0;
    int nslots = 0;
    for (auto& x : st) nslots += (int)x.params.size();
    // arena: checkpoints (nseg) | segment states (s) | Lambda | sigma
    char* arena = nullptr;
    const size_t nbuf = (size_t)nseg + (size_t)s + 2;
    hipError_t e = hipMalloc((void**)&arena, bytes * nbuf);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(density-matrix adjoint checkpoints)");
    struct Guard {
        char* p;
        ~Guard() { (void)hipFree(p); }
    } guard{ arena };
    auto buf = [&](size_t i) { return (double2*)(arena + bytes * i); };
    double2* lam = buf(nbuf - 2);
    double2* sig = buf(nbuf - 1);
    if (int rc = ensure_ws(h->res, sizeof(double) * (2 * (size_t)nslots + 8) + sizeof(PWord) * words.size())) return rc;
    double* res = (double*)h->res.ptr;
    PWord* dwords = (PWord*)(res + 2 * (size_t)nslots + 8);
    // forward, fused per segment, with a checkpoint before each segment
    if (int rc = bsim_state_reset(h)) return rc;
    if (int rc = prepare_tile_kernel()) return rc;
    for (int c = 0; c < nseg; ++c) {
        BSIM_HIP(hipMemcpyAsync(buf(c), h->d, bytes, hipMemcpyDeviceToDevice, h->stream));
        std::vector<EOp> seg;
        for (int k = c * s; k < std::min(D, (c + 1) * s); ++k) seg.insert(seg.end(), st[k].fwd.begin(), st[k].fwd.end());
        Plan plan;
        if (int rc = make_plan(h->vq, seg, tile_config(), plan)) return rc;
        if (int rc = run_plan(h, h->d, seg, plan)) return rc;
    }
    // Lambda = H, <H> = tr(H rho)
    BSIM_HIP(hipMemcpyAsync(dwords, words.data(), sizeof(PWord) * words.size(), hipMemcpyHostToDevice, h->stream));
    const int64_t g = grid_for(len, AB, (int64_t)device_cu_count() * 32);
    hipLaunchKernelGGL(k_dm_pauli_op, dim3((unsigned)g), dim3(AB), 0, h->stream, lam, h->nq, (const PWord*)dwords,
                       (int)words.size());
    BSIM_HIP(hipGetLastError());
    if (int rc = dot_device(lam, h->d, len, h->red, h->stream, res + 2 * (size_t)nslots)) return rc;
    // backward, segment by segment (one-pass products use their own partials)
    // (sized for dot_device too, so that neither reallocates the other's buffer)
    if (int rc = ensure_ws(h->red, sizeof(double) * std::max<size_t>(3 * (size_t)DM_GRID, 4096 * 2 + 64))) return rc;
    double* gpart = (double*)h->red.ptr;
    int slot = nslots;
    std::vector<int> slot_pidx(nslots);
    auto run_fused = [&](double2* d, std::vector<EOp>& list) -> int {
        if (list.empty()) return BSIM_OK;
        Plan plan;
        if (int rc = make_plan(h->vq, list, tile_config(), plan)) return rc;
        if (int rc = run_plan(h, d, list, plan)) return rc;
        list.clear();
        return BSIM_OK;
    };
    std::vector<EOp> pend;
    for (int c = nseg - 1; c >= 0; --c) {
        const int a = c * s, b = std::min(D, (c + 1) * s);
        // states before the segment's parametrized ops: the ops between two of
        // them run as fused passes on a working copy (sig), each state kept by a copy
        std::vector<double2*> before((size_t)(b - a), nullptr);
        BSIM_HIP(hipMemcpyAsync(sig, buf(c), bytes, hipMemcpyDeviceToDevice, h->stream));
        int used = 0;
        for (int k = a; k < b; ++k) {
            if (!st[k].params.empty()) {
                if (int rc = run_fused(sig, pend)) return rc;
                before[k - a] = buf(nseg + used++);
                BSIM_HIP(hipMemcpyAsync(before[k - a], sig, bytes, hipMemcpyDeviceToDevice, h->stream));
            }
            pend.insert(pend.end(), st[k].fwd.begin(), st[k].fwd.end());
        }
        pend.clear();  // the ops after the last parametrized one are not needed
        // Lambda backwards through the segment, fused between parametrized ops
        for (int k = b - 1; k >= a; --k) {
            if (!st[k].params.empty())
                if (int rc = run_fused(lam, pend)) return rc;
            for (int j = (int)st[k].params.size() - 1; j >= 0; --j) {
                const DmParam& P = st[k].params[j];
                if (P.one_pass) {
                    AdjArgs aa = adj_args(h->vq, P.sop, true);
                    AdjMats<2> M;
                    memset(&M, 0, sizeof(M));
                    for (int x = 0; x < 16; ++x) M.udag[x] = P.sop.m[x];
                    hipLaunchKernelGGL(k_dm_grad, dim3(DM_GRID), dim3(AB), 0, h->stream, (const double2*)before[k - a],
                                       (const double2*)lam, aa, M, gpart);
                    BSIM_HIP(hipGetLastError());
                    --slot;
                    slot_pidx[slot] = P.pidx;
                    hipLaunchKernelGGL(k_dm_grad_finish, dim3(1), dim3(256), 0, h->stream, (const double*)gpart, DM_GRID,
                                       res + 2 * (size_t)slot);
                    BSIM_HIP(hipGetLastError());
                    continue;
                }
                BSIM_HIP(hipMemcpyAsync(sig, before[k - a], bytes, hipMemcpyDeviceToDevice, h->stream));
                if (int rc = launch_gate(sig, h->vq, P.dcol, h->scratch, h->stream)) return rc;
                if (P.ctl)
                    if (int rc = launch_gate(sig, h->vq, P.proj, h->scratch, h->stream)) return rc;
                if (int rc = launch_gate(sig, h->vq, P.urow, h->scratch, h->stream)) return rc;
                --slot;
                slot_pidx[slot] = P.pidx;
                if (int rc = dot_device(lam, sig, len, h->red, h->stream, res + 2 * (size_t)slot)) return rc;
            }
            pend.insert(pend.end(), st[k].adj.begin(), st[k].adj.end());
        }
        if (int rc = run_fused(lam, pend)) return rc;
    }
    std::vector<double> hres(2 * (size_t)nslots + 2);
    BSIM_HIP(hipMemcpyAsync(hres.data(), res, sizeof(double) * hres.size(), hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    for (int p = 0; p < n_params; ++p) grad[p] = 0.0;
    for (int q = 0; q < nslots; ++q) grad[slot_pidx[q]] += 2.0 * hres[2 * (size_t)q];
    if (expval) *expval = hres[2 * (size_t)nslots];
    h->alias_ready = false;
    return BSIM_OK;
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_adjoint_gradient(bsim_state* h, const bsim_op* ops, int n_ops, int n_words, const char* paulis,
                          const double* coeffs, int n_params, double* grad, double* expval) {
    if (int rc = check_state(h)) return rc;
    if (n_ops < 0 || (n_ops > 0 && !ops) || n_words < 1 || !paulis || !coeffs || n_params < 0 ||
        (n_params > 0 && !grad))
        return fail(BSIM_EINVAL, "bad adjoint arguments");
    if (h->kind == BSIM_DM) {
        std::vector<PWord> words(n_words);
        for (int w = 0; w < n_words; ++w) {
            PauliMask pm;
            if (int rc = parse_pauli(h->nq, paulis + (size_t)w * h->nq, pm)) return rc;
            words[w] = { pm.x | pm.y, pm.y, pm.z, pm.ny, 0, coeffs[w] };
        }
        return dm_adjoint(h, ops, n_ops, words, n_params, grad, expval);
    }
    std::vector<EOp> eops;
    if (int rc = lower_ops(h->nq, h->kind, ops, n_ops, eops, true)) return rc;
    for (const EOp& e : eops)
        for (int p = 0; p < e.nparams; ++p)
            if (e.pidx[p] < 0 || e.pidx[p] >= n_params) return fail(BSIM_EINVAL, "param_index out of range");
    std::vector<PWord> words(n_words);
    for (int w = 0; w < n_words; ++w) {
        PauliMask pm;
        if (int rc = parse_pauli(h->nq, paulis + (size_t)w * h->nq, pm)) return rc;
        words[w] = { pm.x | pm.y, pm.y, pm.z, pm.ny, 0, coeffs[w] };
    }
    const int64_t len = h->len;
    // buffers: lambda and a working copy of psi (the handle keeps psi_L)
    if (int rc = ensure_ws(h->lam, sizeof(double2) * (size_t)len * 2)) return rc;
    double2* lam = (double2*)h->lam.ptr;
    double2* work = lam + len;
    const int grid = 1024;
    int nslots = 0;
    for (const EOp& e : eops) nslots += e.nparams > 0 ? 1 : 0;
    // partials: [slot][block][3]; sized so that the <H> dot product (which
    // uses the same workspace first: up to 4096 x 2 doubles, or 2 per tile of
    // the last lambda pass) never regrows it after `partials` is taken
    const size_t dot_doubles = h->vq >= PS_T ? (size_t)2 << (h->vq - PS_T) : 0;
    size_t red_doubles = std::max({ (size_t)(nslots + 1) * grid * 3, (size_t)4096 * 2, dot_doubles }) + 64;
    if (int rc = ensure_ws(h->red, sizeof(double) * red_doubles)) return rc;
    double* partials = (double*)h->red.ptr;
    // lambda = H psi: tile passes (vectors of at least PS_T bits), the groups
    // no tile holds by gathers
    PauliTilePlan pt;
    if (h->vq >= PS_T) pauli_tile_plan(h->vq, words, pt);
    else pt.rest = words;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t res_bytes = al(sizeof(double) * (3 * (size_t)nslots + 8));
    const size_t off_pass = res_bytes + al(sizeof(PWord) * (size_t)n_words);
    const size_t off_grp = off_pass + al(sizeof(PSPass) * pt.passes.size());
    const size_t off_wd = off_grp + al(sizeof(PSGroup) * pt.groups.size());
    if (int rc = ensure_ws(h->res, off_wd + al(sizeof(PSWord) * pt.words.size()))) return rc;
    double* res = (double*)h->res.ptr;
    PWord* dwords = (PWord*)((char*)h->res.ptr + res_bytes);
    // forward
    if (int rc = bsim_state_reset(h)) return rc;
    Plan plan;
    if (int rc = make_plan(h->vq, eops, tile_config(), plan)) return rc;
    if (int rc = prepare_tile_kernel()) return rc;
    if (int rc = run_plan(h, h->d, eops, plan)) return rc;
    // lambda = H psi (the first tile pass also writes the working copy of
    // psi, the last one <H>)
    if (int rc = pauli_sum_passes(h, pt, (char*)h->res.ptr, off_pass, off_grp, off_wd, dwords, lam, work,
                                  res + 3 * (size_t)nslots))
        return rc;
    // backward
    int slot = nslots;
    // tile passes over the reversed list of U^dagger ops (both vectors in LDS,
    // parameter products reduced inside the pass), unless their per-tile
    // partial sums would outgrow max(4 GiB, the state's size): then one
    // kernel per op
    std::vector<EOp> back;
    std::vector<int> back_src;  // eops index of each op of `back` (-1: merged run)
    Plan bplan;
    int ns = 0, pm = 0;  // ns == nslots
    bool fused = g_adj_fusion && h->vq >= 5;
    if (fused) {
        if (int rc = adjoint_back_plan(h->vq, eops, back, bplan, ns, pm, nullptr, &back_src)) return rc;
        const size_t limit = std::max<size_t>((size_t)4 << 30, sizeof(double2) * (size_t)len);
        fused = adjoint_tile_partial_bytes(bplan, h->vq, nslots, pm) <= limit;
    }
    if (fused) {
        auto single = [&](int i) -> int {
            if (back_src[i] < 0) return fail(BSIM_EINVAL, "adjoint: a merged run outside the tile passes");
            const EOp& e = eops[back_src[i]];
            if (e.nparams > 0)
                return fail(BSIM_ENOTSUP, "adjoint: parametrized gates act on at most 2 dense or 4 diagonal targets");
            return adjoint_step(h, work, lam, e, nullptr, grid);  // parameter-free: no partials
        };
        if (int rc = adjoint_tile_run(h, work, lam, bplan, nslots, pm, res, single)) return rc;
    } else {
        for (int k = (int)eops.size() - 1; k >= 0; --k) {
            const EOp& e = eops[k];
            double* r = partials;  // unused by parameter-free ops
            if (e.nparams > 0) r = partials + (size_t)(--slot) * grid * 3;
            if (int rc = adjoint_step(h, work, lam, e, r, grid)) return rc;
        }
        if (nslots > 0) {
            hipLaunchKernelGGL(k_adj_finish, dim3(nslots), dim3(256), 0, h->stream, (const double*)partials, grid, res);
            BSIM_HIP(hipGetLastError());
        }
    }
    std::vector<double> hres(3 * (size_t)nslots + 2);
    BSIM_HIP(hipMemcpyAsync(hres.data(), res, sizeof(double) * hres.size(), hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    for (int p = 0; p < n_params; ++p) grad[p] = 0.0;
    slot = 0;
    for (const EOp& e : eops) {
        if (e.nparams == 0) continue;
        for (int p = 0; p < e.nparams; ++p) grad[e.pidx[p]] += 2.0 * hres[3 * (size_t)slot + p];
        ++slot;
    }
    if (expval) *expval = hres[3 * (size_t)nslots];
    h->alias_ready = false;
    return BSIM_OK;
}

int bsim_set_adjoint_fusion(int on, int* old) {
    if (old) *old = g_adj_fusion;
    g_adj_fusion = on ? 1 : 0;
    return BSIM_OK;
}

int bsim_set_checkpoint_interval(int ops, int* old) {
    if (ops < 0) return fail(BSIM_EINVAL, "checkpoint interval must be >= 0");
    if (old) *old = g_checkpoint_every;
    g_checkpoint_every = ops;
    return BSIM_OK;
}

}  // extern "C"
