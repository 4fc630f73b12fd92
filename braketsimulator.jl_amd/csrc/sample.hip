// Shot sampling (src/result_types.jl:8-80): alias table built on the device
// and one table lookup per shot.
//
// Two table builds:
//  - BSIM_ALIAS_REFERENCE: make_alias_table! exactly (result_types.jl:8-54):
//    same classification order, LIFO pairing and `(ap[l] - 1) + ap[s]`
//    update, run by one device thread (the algorithm is sequential).
//  - BSIM_ALIAS_PARALLEL: an exact-integer form of the sweep construction
//    (Vose's alias method in the sweeping form used by Huebschle-Schneider &
//    Sanders, "Parallel Weighted Random Sampling", 2019).  Weights become
//    integers W_i = floor(p_i * N * C / wsum) (C = 2^32 per bucket; the
//    rounding remainder goes to the largest weight), lights (W <= C) and
//    heavies are swept in index order.  With inclusive prefix sums D (light
//    deficits C - W) and E (heavy excesses W - C):
//      light i: ap = W_i / C, alias = first heavy k with E[k] > D[i] - d_i;
//      heavy i: ap = (C + E[i] - D[k]) / C with k the first light where
//               D[k] >= E[i]; alias = the next heavy.
//    Integer sums make the table independent of the reduction order, so a
//    CPU restatement reproduces it bit for bit.
// Shots draw Philox-4x32-10 (Salmon et al., SC'11) with counter = shot
// index and key = seed: index = top n bits of (x1:x0), u = (x3:x2 >> 11)
// * 2^-53, shot = u < ap[index] ? index : alias[index].
#include <vector>

#include "alias_kernels.h"
#include "common.h"
#include "reduce.h"

namespace bsim {

// ---- reference-order build (one thread) -----------------------------------
__global__ void k_alias_reference(const double* __restrict__ w, int64_t n, double wsum, double* __restrict__ ap,
                                  int32_t* __restrict__ alias, int32_t* __restrict__ larges,
                                  int32_t* __restrict__ smalls) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double ac = (double)n / wsum;
    for (int64_t i = 0; i < n; ++i) {
        ap[i] = w[i] * ac;
        alias[i] = 0;
    }
    int64_t kl = 0, ks = 0;
    for (int64_t i = 0; i < n; ++i) {
        double ai = ap[i];
        if (ai > 1.0) larges[kl++] = (int32_t)i;
        else if (ai < 1.0) smalls[ks++] = (int32_t)i;
    }
    while (kl > 0 && ks > 0) {
        int32_t s = smalls[--ks];
        int32_t l = larges[--kl];
        alias[s] = l;
        double al = ap[l] = (ap[l] - 1.0) + ap[s];
        if (al > 1.0) larges[kl++] = l;
        else smalls[ks++] = l;
    }
    for (int64_t i = 0; i < ks; ++i) ap[smalls[i]] = 1.0;
}

// ---- parallel exact build ---------------------------------------------------
constexpr int SUM_BLOCKS = 2048;  // partial sums of k_sum_p

template <class Src>
__global__ __launch_bounds__(256) void k_sum_p(Src p, int64_t n, double* __restrict__ partials) {
    __shared__ double sh[256];
    double s = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        s += p(i);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// one 256-thread block: fixed strided sums, then a fixed tree (deterministic)
__global__ __launch_bounds__(256) void k_sum_finish(const double* __restrict__ partials, int nb,
                                                    double* __restrict__ out) {
    __shared__ double sh[256];
    double s = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partials[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// ---- lookups ---------------------------------------------------------------
__global__ void k_sample(const double* __restrict__ ap, const int32_t* __restrict__ alias, int nq, int64_t shots,
                         uint32_t k0, uint32_t k1, int64_t* __restrict__ out) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < shots; s += (int64_t)gridDim.x * blockDim.x) {
        uint32_t c[4] = { (uint32_t)s, (uint32_t)((uint64_t)s >> 32), 0u, 0u };
        philox4x32_10(c, k0, k1);
        uint64_t hi = ((uint64_t)c[1] << 32) | c[0];
        int64_t idx = (int64_t)(hi >> (64 - nq));
        double u = (double)((((uint64_t)c[3] << 32) | c[2]) >> 11) * 0x1.0p-53;
        out[s] = u < ap[idx] ? idx : (int64_t)alias[idx];
    }
}

__global__ void k_sample_draws(const double* __restrict__ ap, const int32_t* __restrict__ alias, int64_t shots,
                               const int64_t* __restrict__ idx, const double* __restrict__ u,
                               int64_t* __restrict__ out) {
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < shots; s += (int64_t)gridDim.x * blockDim.x) {
        int64_t i = idx[s];
        out[s] = u[s] < ap[i] ? i : (int64_t)alias[i];
    }
}

__global__ void k_shots_to_bits(const int64_t* __restrict__ shots, int64_t n, int nq, int nm,
                                const int* __restrict__ measured, int8_t* __restrict__ out) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n * nm; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t s = t / nm;
        int j = (int)(t % nm);
        out[t] = (int8_t)((shots[s] >> (nq - 1 - measured[j])) & 1);
    }
}

static int alias_build(bsim_state* h, int mode, double wsum) {
    const int64_t n = (int64_t)1 << h->nq;
    if (h->nq > 31) return fail(BSIM_ENOTSUP, "alias tables hold at most 2^31 entries");
    if (int rc = ensure_ws(h->ap, sizeof(double) * (size_t)n)) return rc;
    if (int rc = ensure_ws(h->alias, sizeof(int32_t) * (size_t)n)) return rc;
    // work: p (8n, later the lights' W) | W (8n) | Lval (8n) | Hval (8n) | Lidx (4n) |
    // Hidx (4n) | tile sums | merge splits (2x) | totals | scalars
    const int64_t ntiles = (n + AT_TILE - 1) / AT_TILE;
    const int64_t nsplits = (n + MERGE_G - 1) / MERGE_G + 2;
    size_t bytes = sizeof(uint64_t) * (size_t)4 * n + sizeof(int32_t) * (size_t)2 * n + 256 +
                   sizeof(TileAcc) * (size_t)ntiles + sizeof(int64_t) * 2 * (size_t)nsplits + sizeof(Acc4) * 2 +
                   sizeof(double) * (8 + SUM_BLOCKS) + 1024;
    if (mode == BSIM_ALIAS_REFERENCE) bytes = sizeof(double) * (size_t)n + sizeof(int32_t) * 2 * (size_t)n + 65536;
    if (int rc = ensure_ws(h->alias_work, bytes)) return rc;
    char* base = (char*)h->alias_work.ptr;
    double* p = (double*)base;
    // a state vector's parallel build reads the amplitudes directly
    const bool from_amps = mode == BSIM_ALIAS_PARALLEL && h->kind == BSIM_SV;
    if (!from_amps)
        if (int rc = probabilities_device(h, p)) return rc;
    double* ap = (double*)h->ap.ptr;
    int32_t* al = (int32_t*)h->alias.ptr;
    if (mode == BSIM_ALIAS_REFERENCE) {
        int32_t* larges = (int32_t*)(base + sizeof(double) * n);
        int32_t* smalls = larges + n;
        double* scal = (double*)(smalls + n);
        if (wsum <= 0) {
            const int gs = (int)grid_for(n, 256, SUM_BLOCKS);
            hipLaunchKernelGGL(k_sum_p<ProbVec>, dim3(gs), dim3(256), 0, h->stream, ProbVec{ p }, n, scal + 8);
            hipLaunchKernelGGL(k_sum_finish, dim3(1), dim3(256), 0, h->stream, scal + 8, gs, scal);
            BSIM_HIP(hipMemcpyAsync(&wsum, scal, sizeof(double), hipMemcpyDeviceToHost, h->stream));
            BSIM_HIP(hipStreamSynchronize(h->stream));
        }
        hipLaunchKernelGGL(k_alias_reference, dim3(1), dim3(64), 0, h->stream, p, n, wsum, ap, al, larges, smalls);
        BSIM_HIP(hipGetLastError());
    } else if (mode == BSIM_ALIAS_PARALLEL) {
        uint64_t* W = (uint64_t*)(base + sizeof(double) * n);
        uint64_t* Lval = W + n;
        uint64_t* Hval = Lval + n;
        int32_t* Lidx = (int32_t*)(Hval + n);
        int32_t* Hidx = Lidx + n;
        char* q = (char*)(Hidx + n);
        q = (char*)(((uintptr_t)q + 255) & ~(uintptr_t)255);
        TileAcc* ta = (TileAcc*)q;
        int64_t* splitsL = (int64_t*)(ta + ntiles);
        int64_t* splitsH = splitsL + nsplits;
        Acc4* total = (Acc4*)(splitsH + nsplits);
        double* scal = (double*)(total + 2);
        uint64_t* Lw = (uint64_t*)p;  // p is not read after the quantization
        const int gs = (int)grid_for(n, 256, SUM_BLOCKS);
        if (from_amps) {
            const ProbAmp src{ h->d };
            if (wsum <= 0) {
                hipLaunchKernelGGL(k_sum_p<ProbAmp>, dim3(gs), dim3(256), 0, h->stream, src, n, scal + 8);
                hipLaunchKernelGGL(k_sum_finish, dim3(1), dim3(256), 0, h->stream, scal + 8, gs, scal);
            }
            hipLaunchKernelGGL(k_quant_tiles<ProbAmp>, dim3((unsigned)ntiles), dim3(AT_BLOCK), 0, h->stream, src, n,
                               (const double*)scal, wsum, W, ta);
        } else {
            const ProbVec src{ p };
            if (wsum <= 0) {
                hipLaunchKernelGGL(k_sum_p<ProbVec>, dim3(gs), dim3(256), 0, h->stream, src, n, scal + 8);
                hipLaunchKernelGGL(k_sum_finish, dim3(1), dim3(256), 0, h->stream, scal + 8, gs, scal);
            }
            hipLaunchKernelGGL(k_quant_tiles<ProbVec>, dim3((unsigned)ntiles), dim3(AT_BLOCK), 0, h->stream, src, n,
                               (const double*)scal, wsum, W, ta);
        }
        hipLaunchKernelGGL(k_tiles_finish, dim3(1), dim3(1024), 0, h->stream, ta, ntiles, n, W, total);
        hipLaunchKernelGGL(k_compact, dim3((unsigned)ntiles), dim3(AT_BLOCK), 0, h->stream, W, n, (const TileAcc*)ta,
                           Lidx, Lval, Lw, Hidx, Hval);
        const unsigned gp = (unsigned)grid_for(nsplits, 256, 1 << 20);
        hipLaunchKernelGGL(k_partition<0>, dim3(gp), dim3(256), 0, h->stream, Lval, Lw, Hval, (const Acc4*)total,
                           splitsL, nsplits);
        hipLaunchKernelGGL(k_partition<1>, dim3(gp), dim3(256), 0, h->stream, Lval, Lw, Hval, (const Acc4*)total,
                           splitsH, nsplits);
        const unsigned gm = (unsigned)grid_for(n, 256, (int64_t)device_cu_count() * 256);  // one A item per lane
        hipLaunchKernelGGL(k_merge<0>, dim3(gm), dim3(256), 0, h->stream, Lval, Lw, Lidx, Hval, Hidx,
                           (const Acc4*)total, (const int64_t*)splitsL, ap, al);
        hipLaunchKernelGGL(k_merge<1>, dim3(gm), dim3(256), 0, h->stream, Lval, Lw, Lidx, Hval, Hidx,
                           (const Acc4*)total, (const int64_t*)splitsH, ap, al);
        BSIM_HIP(hipGetLastError());
    } else {
        return fail(BSIM_EINVAL, "unknown alias mode");
    }
    h->alias_ready = true;
    return BSIM_OK;
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_alias_build(bsim_state* h, int mode, double wsum) {
    if (int rc = check_state(h)) return rc;
    if (int rc = alias_build(h, mode, wsum)) return rc;
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_alias_download(bsim_state* h, double* ap, int32_t* alias) {
    if (int rc = check_state(h)) return rc;
    if (!h->alias_ready) return fail(BSIM_EINVAL, "no alias table: call bsim_alias_build after the last change");
    int64_t n = (int64_t)1 << h->nq;
    if (ap) BSIM_HIP(hipMemcpyAsync(ap, h->ap.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    if (alias) BSIM_HIP(hipMemcpyAsync(alias, h->alias.ptr, sizeof(int32_t) * n, hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_sample(bsim_state* h, int64_t shots, uint64_t seed, int64_t* out) {
    if (int rc = check_state(h)) return rc;
    if (shots < 0 || (shots > 0 && !out)) return fail(BSIM_EINVAL, "bad shots/out");
    if (!h->alias_ready) return fail(BSIM_EINVAL, "no alias table: call bsim_alias_build after the last change");
    if (shots == 0) return BSIM_OK;
    if (int rc = ensure_ws(h->scratch, sizeof(int64_t) * (size_t)shots)) return rc;
    int64_t* d = (int64_t*)h->scratch.ptr;
    int64_t g = grid_for(shots, 256, (int64_t)device_cu_count() * 16);
    hipLaunchKernelGGL(k_sample, dim3((unsigned)g), dim3(256), 0, h->stream, (const double*)h->ap.ptr,
                       (const int32_t*)h->alias.ptr, h->nq, shots, (uint32_t)seed, (uint32_t)(seed >> 32), d);
    BSIM_HIP(hipGetLastError());
    BSIM_HIP(hipMemcpyAsync(out, d, sizeof(int64_t) * shots, hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_sample_from_draws(bsim_state* h, int64_t shots, const int64_t* index, const double* u, int64_t* out) {
    if (int rc = check_state(h)) return rc;
    if (shots < 0 || (shots > 0 && (!index || !u || !out))) return fail(BSIM_EINVAL, "bad draws");
    if (!h->alias_ready) return fail(BSIM_EINVAL, "no alias table: call bsim_alias_build after the last change");
    int64_t n = (int64_t)1 << h->nq;
    for (int64_t s = 0; s < shots; ++s)
        if (index[s] < 0 || index[s] >= n) return fail(BSIM_EINVAL, "draw index out of range");
    if (shots == 0) return BSIM_OK;
    if (int rc = ensure_ws(h->scratch, (sizeof(int64_t) * 2 + sizeof(double)) * (size_t)shots)) return rc;
    int64_t* di = (int64_t*)h->scratch.ptr;
    double* du = (double*)(di + shots);
    int64_t* dout = (int64_t*)(du + shots);
    BSIM_HIP(hipMemcpyAsync(di, index, sizeof(int64_t) * shots, hipMemcpyHostToDevice, h->stream));
    BSIM_HIP(hipMemcpyAsync(du, u, sizeof(double) * shots, hipMemcpyHostToDevice, h->stream));
    int64_t g = grid_for(shots, 256, (int64_t)device_cu_count() * 16);
    hipLaunchKernelGGL(k_sample_draws, dim3((unsigned)g), dim3(256), 0, h->stream, (const double*)h->ap.ptr,
                       (const int32_t*)h->alias.ptr, shots, di, du, dout);
    BSIM_HIP(hipGetLastError());
    BSIM_HIP(hipMemcpyAsync(out, dout, sizeof(int64_t) * shots, hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

int bsim_shots_to_bits(bsim_state* h, int64_t shots, const int64_t* shot_index, int nm, const int* measured,
                       int8_t* out) {
    if (int rc = check_state(h)) return rc;
    if (shots < 0 || nm < 0 || (shots > 0 && nm > 0 && (!shot_index || !measured || !out)))
        return fail(BSIM_EINVAL, "bad arguments");
    for (int j = 0; j < nm; ++j)
        if (measured[j] < 0 || measured[j] >= h->nq) return fail(BSIM_EINVAL, "measured qubit out of range");
    if (shots == 0 || nm == 0) return BSIM_OK;
    size_t bytes = sizeof(int64_t) * shots + sizeof(int) * nm + (size_t)shots * nm + 256;
    if (int rc = ensure_ws(h->scratch, bytes)) return rc;
    int64_t* ds = (int64_t*)h->scratch.ptr;
    int* dm = (int*)(ds + shots);
    int8_t* dout = (int8_t*)(dm + nm);
    BSIM_HIP(hipMemcpyAsync(ds, shot_index, sizeof(int64_t) * shots, hipMemcpyHostToDevice, h->stream));
    BSIM_HIP(hipMemcpyAsync(dm, measured, sizeof(int) * nm, hipMemcpyHostToDevice, h->stream));
    int64_t g = grid_for(shots * nm, 256, (int64_t)device_cu_count() * 16);
    hipLaunchKernelGGL(k_shots_to_bits, dim3((unsigned)g), dim3(256), 0, h->stream, ds, shots, h->nq, nm, dm, dout);
    BSIM_HIP(hipGetLastError());
    BSIM_HIP(hipMemcpyAsync(out, dout, (size_t)shots * nm, hipMemcpyDeviceToHost, h->stream));
    BSIM_HIP(hipStreamSynchronize(h->stream));
    return BSIM_OK;
}

}  // extern "C"
