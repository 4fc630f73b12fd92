// Parallel exact alias-table build (see sample.hip for the construction).
//
// Pass 1 (k_scan_tiles):   per tile of 2048 weights: light deficit sum,
//                          heavy excess sum, light count, heavy count.
// Pass 2 (k_scan_tilesums): exclusive scan of the tile sums (one block).
// Pass 3 (k_scan_apply):   inclusive prefix of every item; lights and
//                          heavies are compacted in index order with their
//                          prefix value (light: deficits so far, heavy:
//                          excesses so far).
// Pass 4 (k_alias_lights / k_alias_heavies): each thread owns a run of
//                          consecutive lights (heavies); the searches of a
//                          run move monotonically through the other list, so
//                          one galloping search per item starts where the
//                          previous one ended.
#pragma once

#include "common.h"

namespace bsim {

constexpr uint64_t ALIAS_C = 1ull << 32;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_TILE = SCAN_ITEMS * SCAN_BLOCK;
constexpr int ALIAS_RUN = 64;

struct Acc4 {
    uint64_t d, e, nl, nh;
};

__device__ __forceinline__ void classify(uint64_t w, Acc4& a) {
    if (w <= ALIAS_C) {
        a.d += ALIAS_C - w;
        a.nl += 1;
    } else {
        a.e += w - ALIAS_C;
        a.nh += 1;
    }
}

// block-wide inclusive scan of 4 lanes of uint64 held one per thread
__device__ __forceinline__ void block_scan4(Acc4& v, uint64_t (*sh)[SCAN_BLOCK]) {
    sh[0][threadIdx.x] = v.d;
    sh[1][threadIdx.x] = v.e;
    sh[2][threadIdx.x] = v.nl;
    sh[3][threadIdx.x] = v.nh;
    __syncthreads();
    for (int o = 1; o < SCAN_BLOCK; o <<= 1) {
        uint64_t a[4] = { 0, 0, 0, 0 };
        if ((int)threadIdx.x >= o)
            for (int q = 0; q < 4; ++q) a[q] = sh[q][threadIdx.x - o];
        __syncthreads();
        for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += a[q];
        __syncthreads();
    }
    v.d = sh[0][threadIdx.x];
    v.e = sh[1][threadIdx.x];
    v.nl = sh[2][threadIdx.x];
    v.nh = sh[3][threadIdx.x];
    __syncthreads();
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(const uint64_t* __restrict__ W, int64_t n,
                                                           Acc4* __restrict__ tsum) {
    __shared__ uint64_t sh[4][SCAN_BLOCK];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    Acc4 a = { 0, 0, 0, 0 };
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) classify(W[base + k], a);
    block_scan4(a, sh);
    if (threadIdx.x == SCAN_BLOCK - 1) tsum[blockIdx.x] = a;
}

// exclusive scan of the tile sums in place (one block, sequential chunks)
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tilesums(Acc4* __restrict__ tsum, int64_t ntiles,
                                                              Acc4* __restrict__ total) {
    __shared__ uint64_t sh[4][SCAN_BLOCK];
    int64_t per = (ntiles + SCAN_BLOCK - 1) / SCAN_BLOCK;
    int64_t b0 = (int64_t)threadIdx.x * per;
    Acc4 s = { 0, 0, 0, 0 };
    for (int64_t k = 0; k < per; ++k)
        if (b0 + k < ntiles) {
            Acc4 t = tsum[b0 + k];
            s.d += t.d;
            s.e += t.e;
            s.nl += t.nl;
            s.nh += t.nh;
        }
    Acc4 inc = s;
    block_scan4(inc, sh);
    Acc4 o = { inc.d - s.d, inc.e - s.e, inc.nl - s.nl, inc.nh - s.nh };
    for (int64_t k = 0; k < per; ++k)
        if (b0 + k < ntiles) {
            Acc4 t = tsum[b0 + k];
            tsum[b0 + k] = o;
            o.d += t.d;
            o.e += t.e;
            o.nl += t.nl;
            o.nh += t.nh;
        }
    if (threadIdx.x == SCAN_BLOCK - 1) *total = inc;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const uint64_t* __restrict__ W, int64_t n,
                                                           const Acc4* __restrict__ tsum, int32_t* __restrict__ Lidx,
                                                           uint64_t* __restrict__ Lval, int32_t* __restrict__ Hidx,
                                                           uint64_t* __restrict__ Hval) {
    __shared__ uint64_t sh[4][SCAN_BLOCK];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t w[SCAN_ITEMS];
    Acc4 a = { 0, 0, 0, 0 };
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        w[k] = base + k < n ? W[base + k] : ALIAS_C;
        if (base + k < n) classify(w[k], a);
    }
    Acc4 inc = a;
    block_scan4(inc, sh);
    const Acc4 t = tsum[blockIdx.x];
    Acc4 o = { t.d + inc.d - a.d, t.e + inc.e - a.e, t.nl + inc.nl - a.nl, t.nh + inc.nh - a.nh };
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        int64_t i = base + k;
        if (i >= n) break;
        if (w[k] <= ALIAS_C) {
            o.d += ALIAS_C - w[k];
            Lidx[o.nl] = (int32_t)i;
            Lval[o.nl] = o.d;
            o.nl += 1;
        } else {
            o.e += w[k] - ALIAS_C;
            Hidx[o.nh] = (int32_t)i;
            Hval[o.nh] = o.e;
            o.nh += 1;
        }
    }
}

// first k >= from in a non-decreasing a[0..n) with a[k] > v (gt) or a[k] >= v
// (!gt); n if none.  Gallops from `from`, then bisects.
template <bool GT>
__device__ __forceinline__ int64_t gallop(const uint64_t* __restrict__ a, int64_t n, int64_t from, uint64_t v) {
    auto ok = [&](int64_t k) { return GT ? (a[k] > v) : (a[k] >= v); };
    int64_t lo = from, step = 1, hi = from;
    while (hi < n && !ok(hi)) {
        lo = hi + 1;
        hi = from + step;
        step <<= 1;
    }
    if (hi > n) hi = n;
    while (lo < hi) {
        int64_t mid = lo + ((hi - lo) >> 1);
        if (ok(mid)) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Lanes of a wavefront take consecutive items (coalesced).  The first lane's
// answer, found by bisection and broadcast, is a lower bound for every other
// lane (answers are monotone in the item), so each lane gallops from it over
// a few cached entries.
__device__ __forceinline__ int64_t wave_first_lane_i64(int64_t v) {
    return ((int64_t)__shfl(int(v >> 32), 0, 64) << 32) | (uint32_t)__shfl(int(v), 0, 64);
}

__global__ void k_alias_lights(const uint64_t* __restrict__ W, const int32_t* __restrict__ Lidx,
                               const uint64_t* __restrict__ Lval, const int32_t* __restrict__ Hidx,
                               const uint64_t* __restrict__ Hval, const Acc4* __restrict__ total,
                               double* __restrict__ ap, int32_t* __restrict__ alias) {
    const int64_t nl = (int64_t)total->nl, nh = (int64_t)total->nh;
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (nl + 63) / 64;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t wv = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); wv < nwaves; wv += wstride) {
        const int64_t l = wv * 64 + lane;
        const bool valid = l < nl;
        int32_t i = 0;
        uint64_t w = ALIAS_C, dex = 0;
        if (valid) {
            i = Lidx[l];
            w = W[i];
            dex = Lval[l] - (ALIAS_C - w);  // deficits of the lights before this one
        }
        // lane 0's query is the smallest of the wave
        int64_t r0 = 0;
        if (lane == 0 && valid) r0 = gallop<true>(Hval, nh, 0, dex);
        r0 = wave_first_lane_i64(r0);
        if (valid) {
            ap[i] = (double)w / (double)ALIAS_C;
            if (w == ALIAS_C) {
                alias[i] = i;
            } else {
                const int64_t r = gallop<true>(Hval, nh, r0, dex);
                alias[i] = r < nh ? Hidx[r] : i;
            }
        }
    }
}

__global__ void k_alias_heavies(const uint64_t* __restrict__ Lval, const int32_t* __restrict__ Hidx,
                                const uint64_t* __restrict__ Hval, const Acc4* __restrict__ total,
                                double* __restrict__ ap, int32_t* __restrict__ alias) {
    const int64_t nl = (int64_t)total->nl, nh = (int64_t)total->nh;
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (nh + 63) / 64;
    const int64_t wstride = (int64_t)gridDim.x * (blockDim.x / 64);
    for (int64_t wv = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); wv < nwaves; wv += wstride) {
        const int64_t h = wv * 64 + lane;
        const bool valid = h < nh;
        int32_t i = 0;
        uint64_t e = 0;
        if (valid) {
            i = Hidx[h];
            e = Hval[h];
        }
        int64_t l0 = 0;
        if (lane == 0 && valid) l0 = gallop<false>(Lval, nl, 0, e);
        l0 = wave_first_lane_i64(l0);
        if (valid) {
            const int64_t l = gallop<false>(Lval, nl, l0, e);  // first light whose prefix deficit reaches e
            uint64_t rr = l < nl ? ALIAS_C + e - Lval[l] : ALIAS_C;
            int32_t nx = i;
            if (h + 1 < nh) nx = Hidx[h + 1];
            else rr = ALIAS_C;
            ap[i] = (double)rr / (double)ALIAS_C;
            alias[i] = nx;
        }
    }
}

}  // namespace bsim
