// Parallel exact alias-table build (the construction is described in
// sample.hip).
//
//   k_quant_tiles    W_i = floor(p_i * scale) and, per tile of 8192 weights,
//                    light deficit / heavy excess sums, counts, the sum of W
//                    and the first largest weight.
//   k_tiles_finish   one block: the rounding remainder goes to the first
//                    largest weight (its tile sums are reclassified), then an
//                    exclusive scan of the tile sums.
//   k_compact        per tile, row by row (wave scans): the inclusive prefix
//                    of every weight; lights and heavies are compacted in
//                    index order with their prefix (lights also keep W).
//   k_partition      coarse split points of the two sweeps below (one
//                    bisection every MERGE_G items).
//   k_merge<0>       light l gets alias = first heavy whose prefix excess
//                    exceeds the deficits of the lights before l; both
//                    sequences are non-decreasing, so a wavefront of 64
//                    consecutive lights finds its answers in one window of
//                    heavies staged in LDS (a merge of the two sequences).
//   k_merge<1>       heavy h gets the first light whose prefix deficit reaches
//                    h's prefix excess: the same merge with ties the other way.
#pragma once

#include "common.h"

namespace bsim {

constexpr uint64_t ALIAS_C = 1ull << 32;
constexpr int AT_BLOCK = 256;
constexpr int AT_ROWS = 32;
constexpr int AT_TILE = AT_ROWS * AT_BLOCK;  // weights per tile
constexpr int MERGE_G = 1024;                // A items per coarse split point

struct TileAcc {
    uint64_t d, e, nl, nh;  // light deficits, heavy excesses, light / heavy counts
    uint64_t wsum, mx;      // sum of W, largest W
    int64_t arg, pad_;      // first index of the largest W
};

struct Acc4 {
    uint64_t d, e, nl, nh;
};

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int o) {
    return ((uint64_t)(uint32_t)__shfl_xor((int)(v >> 32), o, 64) << 32) | (uint32_t)__shfl_xor((int)v, o, 64);
}
__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, int o) {
    return ((uint64_t)(uint32_t)__shfl_up((int)(v >> 32), o, 64) << 32) | (uint32_t)__shfl_up((int)v, o, 64);
}
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
    return ((uint64_t)(uint32_t)__shfl((int)(v >> 32), src, 64) << 32) | (uint32_t)__shfl((int)v, src, 64);
}
__device__ __forceinline__ uint64_t wave_scan_u64(uint64_t v, int lane) {  // inclusive
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t y = shfl_up_u64(v, o);
        if (lane >= o) v += y;
    }
    return v;
}
// (larger w, then smaller index) wins
__device__ __forceinline__ void max_arg(uint64_t& mx, int64_t& arg, uint64_t w, int64_t i) {
    if (w > mx || (w == mx && i < arg)) {
        mx = w;
        arg = i;
    }
}

template <class Src>
__global__ __launch_bounds__(AT_BLOCK) void k_quant_tiles(Src p, int64_t n,
                                                          const double* __restrict__ wsum_dev, double wsum_host,
                                                          uint64_t* __restrict__ W, TileAcc* __restrict__ ta) {
    __shared__ uint64_t sh[4][7];
    const double wsum = wsum_host > 0 ? wsum_host : wsum_dev[0];
    const double scale = (double)n * (double)ALIAS_C / wsum;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t base = (int64_t)blockIdx.x * AT_TILE;
    uint64_t d = 0, e = 0, nl = 0, nh = 0, ws = 0, mx = 0;
    int64_t arg = INT64_MAX;
    for (int r = 0; r < AT_ROWS; ++r) {
        const int64_t i = base + r * AT_BLOCK + t;
        if (i < n) {
            const double x = p(i) * scale;
            const uint64_t w = x > 0 ? (uint64_t)x : 0;
            W[i] = w;
            if (w <= ALIAS_C) {
                d += ALIAS_C - w;
                ++nl;
            } else {
                e += w - ALIAS_C;
                ++nh;
            }
            ws += w;
            max_arg(mx, arg, w, i);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        d += shfl_xor_u64(d, o);
        e += shfl_xor_u64(e, o);
        nl += shfl_xor_u64(nl, o);
        nh += shfl_xor_u64(nh, o);
        ws += shfl_xor_u64(ws, o);
        max_arg(mx, arg, shfl_xor_u64(mx, o), (int64_t)shfl_xor_u64((uint64_t)arg, o));
    }
    if (lane == 0) {
        sh[wave][0] = d;
        sh[wave][1] = e;
        sh[wave][2] = nl;
        sh[wave][3] = nh;
        sh[wave][4] = ws;
        sh[wave][5] = mx;
        sh[wave][6] = (uint64_t)arg;
    }
    __syncthreads();
    if (t == 0) {
        TileAcc a = { 0, 0, 0, 0, 0, 0, INT64_MAX, 0 };
        for (int w = 0; w < 4; ++w) {
            a.d += sh[w][0];
            a.e += sh[w][1];
            a.nl += sh[w][2];
            a.nh += sh[w][3];
            a.wsum += sh[w][4];
            max_arg(a.mx, a.arg, sh[w][5], (int64_t)sh[w][6]);
        }
        ta[blockIdx.x] = a;
    }
}

// One block of 1024 threads.
__global__ __launch_bounds__(1024) void k_tiles_finish(TileAcc* __restrict__ ta, int64_t ntiles, int64_t n,
                                                       uint64_t* __restrict__ W, Acc4* __restrict__ total) {
    __shared__ uint64_t sh[16][5];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    // ---- sum of W and the first largest weight
    uint64_t ws = 0, mx = 0;
    int64_t arg = INT64_MAX;
    for (int64_t k = t; k < ntiles; k += 1024) {
        ws += ta[k].wsum;
        max_arg(mx, arg, ta[k].mx, ta[k].arg);
    }
    for (int o = 32; o > 0; o >>= 1) {
        ws += shfl_xor_u64(ws, o);
        max_arg(mx, arg, shfl_xor_u64(mx, o), (int64_t)shfl_xor_u64((uint64_t)arg, o));
    }
    if (lane == 0) {
        sh[wave][0] = ws;
        sh[wave][1] = mx;
        sh[wave][2] = (uint64_t)arg;
    }
    __syncthreads();
    if (t == 0) {
        uint64_t S = 0, M = 0;
        int64_t A = INT64_MAX;
        for (int w = 0; w < 16; ++w) {
            S += sh[w][0];
            max_arg(M, A, sh[w][1], (int64_t)sh[w][2]);
        }
        // the remainder (modular: either sign) goes to the first largest weight
        const uint64_t w0 = W[A], w1 = w0 + ((uint64_t)n * ALIAS_C - S);
        W[A] = w1;
        TileAcc& a = ta[A / AT_TILE];
        if (w0 <= ALIAS_C) {
            a.d -= ALIAS_C - w0;
            --a.nl;
        } else {
            a.e -= w0 - ALIAS_C;
            --a.nh;
        }
        if (w1 <= ALIAS_C) {
            a.d += ALIAS_C - w1;
            ++a.nl;
        } else {
            a.e += w1 - ALIAS_C;
            ++a.nh;
        }
    }
    __threadfence_block();
    __syncthreads();
    // ---- exclusive scan of the tile sums: thread t owns a run of tiles
    const int64_t per = (ntiles + 1023) / 1024, k0 = (int64_t)t * per;
    uint64_t v[4] = { 0, 0, 0, 0 };
    for (int64_t k = k0; k < k0 + per && k < ntiles; ++k) {
        v[0] += ta[k].d;
        v[1] += ta[k].e;
        v[2] += ta[k].nl;
        v[3] += ta[k].nh;
    }
    uint64_t inc[4];
    for (int q = 0; q < 4; ++q) inc[q] = wave_scan_u64(v[q], lane);
    if (lane == 63)
        for (int q = 0; q < 4; ++q) sh[wave][q] = inc[q];
    __syncthreads();
    uint64_t off[4] = { 0, 0, 0, 0 };
    for (int w = 0; w < wave; ++w)
        for (int q = 0; q < 4; ++q) off[q] += sh[w][q];
    uint64_t o[4];
    for (int q = 0; q < 4; ++q) o[q] = off[q] + inc[q] - v[q];
    for (int64_t k = k0; k < k0 + per && k < ntiles; ++k) {
        const TileAcc a = ta[k];
        ta[k].d = o[0];
        ta[k].e = o[1];
        ta[k].nl = o[2];
        ta[k].nh = o[3];
        o[0] += a.d;
        o[1] += a.e;
        o[2] += a.nl;
        o[3] += a.nh;
    }
    if (t == 1023) *total = { off[0] + inc[0], off[1] + inc[1], off[2] + inc[2], off[3] + inc[3] };
}

__global__ __launch_bounds__(AT_BLOCK) void k_compact(const uint64_t* __restrict__ W, int64_t n,
                                                      const TileAcc* __restrict__ ta, int32_t* __restrict__ Lidx,
                                                      uint64_t* __restrict__ Lval, uint64_t* __restrict__ Lw,
                                                      int32_t* __restrict__ Hidx, uint64_t* __restrict__ Hval) {
    __shared__ uint64_t sh[2][4][2];  // [row parity][wave][light, heavy]
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t base = (int64_t)blockIdx.x * AT_TILE;
    const TileAcc a = ta[blockIdx.x];
    uint64_t od = a.d, oe = a.e, onl = a.nl, onh = a.nh;
    // per row: light deficit + (light count << 48) and heavy excess + (heavy
    // count << 48) in one 64-bit scan each (a row's deficits stay below 2^40)
    constexpr int CS = 48;
    constexpr uint64_t LOW = (1ull << CS) - 1;
    for (int r = 0; r < AT_ROWS; ++r) {
        const int64_t i = base + r * AT_BLOCK + t;
        const bool valid = i < n;
        const uint64_t w = valid ? W[i] : 0;
        const bool isL = valid && w <= ALIAS_C, isH = valid && w > ALIAS_C;
        const uint64_t vl = wave_scan_u64(isL ? (ALIAS_C - w) | (1ull << CS) : 0, lane);
        const uint64_t vh = wave_scan_u64(isH ? (w - ALIAS_C) | (1ull << CS) : 0, lane);
        const int par = r & 1;
        if (lane == 63) {
            sh[par][wave][0] = vl;
            sh[par][wave][1] = vh;
        }
        __syncthreads();
        uint64_t pl = 0, ph = 0, rl = 0, rh = 0;
        for (int ww = 0; ww < 4; ++ww) {
            if (ww < wave) {
                pl += sh[par][ww][0];
                ph += sh[par][ww][1];
            }
            rl += sh[par][ww][0];
            rh += sh[par][ww][1];
        }
        if (isL) {
            const uint64_t x = pl + vl;
            const uint64_t pos = onl + (x >> CS) - 1;
            Lidx[pos] = (int32_t)i;
            Lval[pos] = od + (x & LOW);
            Lw[pos] = w;
        } else if (isH) {
            const uint64_t x = ph + vh;
            const uint64_t pos = onh + (x >> CS) - 1;
            Hidx[pos] = (int32_t)i;
            Hval[pos] = oe + (x & LOW);
        }
        od += rl & LOW;
        oe += rh & LOW;
        onl += rl >> CS;
        onh += rh >> CS;
    }
}

// Merge of sequence A (the items being assigned) with sequence B.  MODE 0
// (lights): A = deficits before each light, B = heavy prefix excesses; an A
// item goes before the B items it is smaller than.  MODE 1 (heavies): A =
// heavy prefix excesses, B = light prefix deficits; an A item goes before the
// B items it does not exceed.
template <int MODE>
__device__ __forceinline__ uint64_t key_a(const uint64_t* __restrict__ Lval, const uint64_t* __restrict__ Lw,
                                          const uint64_t* __restrict__ Hval, int64_t i) {
    if constexpr (MODE == 0) return Lval[i] - (ALIAS_C - Lw[i]);
    else return Hval[i];
}
template <int MODE>
__device__ __forceinline__ uint64_t key_b(const uint64_t* __restrict__ Lval, const uint64_t* __restrict__ Hval,
                                          int64_t j) {
    if constexpr (MODE == 0) return Hval[j];
    else return Lval[j];
}
template <int MODE>
__device__ __forceinline__ bool a_first(uint64_t a, uint64_t b) {
    if constexpr (MODE == 0) return a < b;
    else return a <= b;
}

// r(l) = first j with a_first(A[l], B[j]) (nB if none), searched in [lo, hi)
template <int MODE>
__device__ __forceinline__ int64_t first_after(const uint64_t* __restrict__ Lval, const uint64_t* __restrict__ Hval,
                                               uint64_t a, int64_t lo, int64_t hi) {
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (a_first<MODE>(a, key_b<MODE>(Lval, Hval, mid))) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Coarse split points: rc[c] = r(c * MERGE_G) for every MERGE_G-th A item
// (rc = nB past the end).
template <int MODE>
__global__ void k_partition(const uint64_t* __restrict__ Lval, const uint64_t* __restrict__ Lw,
                            const uint64_t* __restrict__ Hval, const Acc4* __restrict__ total,
                            int64_t* __restrict__ rc, int64_t nsplits) {
    const int64_t nl = (int64_t)total->nl, nh = (int64_t)total->nh;
    const int64_t nA = MODE == 0 ? nl : nh, nB = MODE == 0 ? nh : nl;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < nsplits; c += (int64_t)gridDim.x * blockDim.x) {
        const int64_t l = c * MERGE_G;
        rc[c] = l < nA ? first_after<MODE>(Lval, Hval, key_a<MODE>(Lval, Lw, Hval, l), 0, nB) : nB;
    }
}

// One wavefront assigns the A items l0 .. l0+63 (those below lend): the
// wave's smallest answer r0 comes from the coarse splits, the 128 B items from
// r0 on are staged in LDS (`win`, coalesced), and every lane finds its own
// answer there (a global search only if it lies beyond the window).  The
// result (index, ap, alias) goes to `put`.
template <int MODE, class Put>
__device__ __forceinline__ void assign_wave(const uint64_t* __restrict__ Lval, const uint64_t* __restrict__ Lw,
                                            const int32_t* __restrict__ Lidx, const uint64_t* __restrict__ Hval,
                                            const int32_t* __restrict__ Hidx, int64_t nA, int64_t nB,
                                            const int64_t* __restrict__ rc, int64_t l0, int64_t lend,
                                            uint64_t* __restrict__ win, Put&& put) {
    constexpr int WIN = 128;
    const int lane = threadIdx.x & 63;
    const int64_t l = l0 + lane;
    const bool valid = l < lend;
    const uint64_t a = valid ? key_a<MODE>(Lval, Lw, Hval, l) : 0;
    // r0 = answer of item l0, in [rc[c], rc[c+1]]: every round the 64 lanes
    // probe 64 evenly spaced points and the interval shrinks 64-fold
    const uint64_t a0 = shfl_u64(a, 0);
    const int64_t c = l0 / MERGE_G;
    int64_t lo = rc[c], hi = rc[c + 1];
    while (lo < hi) {
        const int64_t step = (hi - lo + 63) / 64;
        const int64_t pk = lo + (int64_t)lane * step;
        const bool pred = pk < hi && a_first<MODE>(a0, key_b<MODE>(Lval, Hval, pk));
        const uint64_t m = __ballot(pred);
        if (m == 0) {  // past the last probe below hi
            lo = lo + ((hi - 1 - lo) / step) * step + 1;
        } else {  // between the last false probe and the first true one
            const int k = __builtin_ctzll(m);
            const int64_t first = lo + (int64_t)k * step;
            if (k > 0) lo = first - step + 1;
            hi = first;
        }
    }
    const int64_t r0 = hi;
    const int64_t nwin = nB - r0 < WIN ? nB - r0 : WIN;
    for (int k = lane; k < WIN; k += 64) win[k] = k < nwin ? key_b<MODE>(Lval, Hval, r0 + k) : ~0ull;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (valid) {
        int lo = 0, hi = (int)nwin;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (a_first<MODE>(a, win[mid])) hi = mid;
            else lo = mid + 1;
        }
        int64_t r = r0 + lo;
        uint64_t br = 0;  // B[r] if r < nB
        if (lo < nwin) {
            br = win[lo];
        } else if (r < nB) {  // beyond the window: gallop, then bisect
            int64_t step = 1, lo2 = r, hi2 = r;
            while (hi2 < nB && !a_first<MODE>(a, key_b<MODE>(Lval, Hval, hi2))) {
                lo2 = hi2 + 1;
                hi2 = r + step;
                step <<= 1;
            }
            if (hi2 > nB) hi2 = nB;
            r = first_after<MODE>(Lval, Hval, a, lo2, hi2);
            if (r < nB) br = key_b<MODE>(Lval, Hval, r);
        }
        if constexpr (MODE == 0) {  // light: alias = first heavy whose prefix excess exceeds a
            const int32_t idx = Lidx[l];
            const uint64_t w = Lw[l];
            put(idx, (double)w / (double)ALIAS_C, (w == ALIAS_C || r >= nB) ? idx : Hidx[r]);
        } else {  // heavy: residual from the first light whose prefix deficit reaches a
            const int32_t idx = Hidx[l];
            uint64_t rr = r < nB ? ALIAS_C + a - br : ALIAS_C;
            int32_t nx = idx;
            if (l + 1 < nA) nx = Hidx[l + 1];
            else rr = ALIAS_C;
            put(idx, (double)rr / (double)ALIAS_C, nx);
        }
    }
    __builtin_amdgcn_wave_barrier();  // the window is rewritten by the next call
}

// One wavefront per 64 consecutive A items, all waves of the grid striding
// over the list; results are written straight to the table.
template <int MODE>
__global__ __launch_bounds__(256) void k_merge(const uint64_t* __restrict__ Lval, const uint64_t* __restrict__ Lw,
                                               const int32_t* __restrict__ Lidx, const uint64_t* __restrict__ Hval,
                                               const int32_t* __restrict__ Hidx, const Acc4* __restrict__ total,
                                               const int64_t* __restrict__ rc, double* __restrict__ ap,
                                               int32_t* __restrict__ alias) {
    __shared__ uint64_t win[4][128];
    const int64_t nl = (int64_t)total->nl, nh = (int64_t)total->nh;
    const int64_t nA = MODE == 0 ? nl : nh, nB = MODE == 0 ? nh : nl;
    const int wave = threadIdx.x >> 6;
    const int64_t nwaves = (nA + 63) / 64;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / 64);
    auto put = [&](int32_t idx, double a, int32_t al) {
        ap[idx] = a;
        alias[idx] = al;
    };
    for (int64_t wv = (int64_t)blockIdx.x * (blockDim.x / 64) + wave; wv < nwaves; wv += wstride)
        assign_wave<MODE>(Lval, Lw, Lidx, Hval, Hidx, nA, nB, rc, wv * 64, nA, win[wave], put);
}

}  // namespace bsim
