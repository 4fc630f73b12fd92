// Batched execution in one persistent launch (the reference's batch
// simulate, src/BraketSimulator.jl:347-395, runs one simulate per circuit on
// a pool of Julia tasks).
//
// Every circuit is planned on the host into fused tile passes (the same
// plans as bsim_evolve, tile = min(12, n) bits).  The work list is a ticket
// sequence: one ticket per (circuit, pass, tile) and one "finish" ticket per
// circuit, ordered pass-wave by pass-wave across circuits (longest circuits
// first), so a ticket's dependency -- every tile of the same circuit's
// previous pass -- always has smaller tickets.  A persistent grid pulls
// tickets from one atomic counter; a ticket waits (bounded spin) until its
// dependency's completion counter is full, runs, then releases its own
// counter.  Pass 0 builds |0...0> in LDS instead of reading the state.
//
// The finish ticket (one workgroup per circuit) reduces every <Z_q> in one
// read and draws the shots by inverse CDF over exact integer weights
// W_i = floor(p_i * 2^n * 2^32) (p_i = re^2 + im^2 rounded as written):
// prefix sums in place of the finished state, then per shot
// t = floor(r * total / 2^64) for r = (x1:x0) of Philox-4x32-10(shot, seed)
// and the shot is the first i whose inclusive prefix exceeds t.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "fused.h"
#include "tile_kernel.h"

namespace bsim {

struct BTask {
    double2* sv;
    int32_t n, pass0;
    int64_t shots;
    uint64_t seed;
    double* zout;
    int64_t* sout;
};

struct BEntry {
    int32_t task, pass;  // pass < 0: finish ticket
    int32_t dep, pad_;
    int64_t ntiles, prefix;
};

struct BCtl {
    unsigned long long counter;
    unsigned long long err;
    unsigned long long t0;
    unsigned long long pad_;
};

constexpr int ZROWS = 27;

// <Z_q> and shots of one finished circuit, by one workgroup.  `red` is LDS
// scratch of at least 27 * TB doubles.
__device__ __noinline__ void finish_task(const BTask& K, double* red) {
    const int n = K.n;
    const int64_t N = (int64_t)1 << n;
    const int t = threadIdx.x;
    // ---- <Z_q>: i = t + TB k; bits 0..7 of i are t, bits 8.. are k
    double tot = 0, hi[18];
    for (int b = 0; b < 18; ++b) hi[b] = 0;
    const int nhi = n > 8 ? n - 8 : 0;
    int64_t k = 0;
    for (int64_t i = t; i < N; i += TB, ++k) {
        const double2 a = K.sv[i];
        const double p = __dadd_rn(__dmul_rn(a.x, a.x), __dmul_rn(a.y, a.y));
        tot += p;
        for (int b = 0; b < 18; ++b)
            if (b < nhi && ((k >> b) & 1)) hi[b] += p;
    }
    // rows: 0 total, 1..8 low bits 0..7, 9..26 high bits 8..25
    double v[ZROWS];
    v[0] = tot;
    for (int b = 0; b < 8; ++b) v[1 + b] = ((t >> b) & 1) ? tot : 0.0;
    for (int b = 0; b < 18; ++b) v[9 + b] = hi[b];
    for (int r = 0; r < ZROWS; ++r) red[r * TB + t] = v[r];
    __syncthreads();
    for (int o = TB / 2; o > 0; o >>= 1) {
        if (t < o)
            for (int r = 0; r < ZROWS; ++r) red[r * TB + t] += red[r * TB + t + o];
        __syncthreads();
    }
    if (t < n) {
        const int b = n - 1 - t;  // qubit t <-> bit n-1-t
        const double s = b < 8 ? red[(1 + b) * TB] : red[(9 + b - 8) * TB];
        K.zout[t] = red[0] - 2.0 * s;
    }
    __syncthreads();
    if (K.shots <= 0) return;
    // ---- exact integer weights, inclusive prefix written over the state
    const double scale = (double)N * 4294967296.0;
    uint64_t* pre = (uint64_t*)K.sv;
    uint64_t* sh = (uint64_t*)red;  // TB scan slots + carry
    uint64_t carry = 0;
    for (int64_t c0 = 0; c0 < N; c0 += (int64_t)TB * 8) {
        uint64_t w[8];
        uint64_t mine = 0;
        for (int r = 0; r < 8; ++r) {
            const int64_t i = c0 + (int64_t)t * 8 + r;
            w[r] = 0;
            if (i < N) {
                const double2 a = K.sv[i];
                const double x = __dmul_rn(__dadd_rn(__dmul_rn(a.x, a.x), __dmul_rn(a.y, a.y)), scale);
                w[r] = x > 0 ? (uint64_t)x : 0;
            }
            mine += w[r];
        }
        sh[t] = mine;
        __syncthreads();  // every amplitude of this chunk has been read
        for (int o = 1; o < TB; o <<= 1) {
            uint64_t add = t >= o ? sh[t - o] : 0;
            __syncthreads();
            sh[t] += add;
            __syncthreads();
        }
        uint64_t run = carry + sh[t] - mine;
        for (int r = 0; r < 8; ++r) {
            const int64_t i = c0 + (int64_t)t * 8 + r;
            run += w[r];
            if (i < N) pre[i] = run;  // overwrites bytes of amplitudes < c0 + 8 t + r (already read)
        }
        carry += sh[TB - 1];
        __syncthreads();
    }
    const uint64_t total = carry;
    __threadfence_block();
    __syncthreads();
    for (int64_t s = t; s < K.shots; s += TB) {
        uint32_t c[4] = { (uint32_t)s, (uint32_t)((uint64_t)s >> 32), 0u, 0u };
        philox4x32_10(c, (uint32_t)K.seed, (uint32_t)(K.seed >> 32));
        const uint64_t r = ((uint64_t)c[1] << 32) | c[0];
        const uint64_t tt = __umul64hi(r, total);
        int64_t lo = 0, hi2 = N - 1;
        while (lo < hi2) {
            int64_t mid = (lo + hi2) >> 1;
            if (pre[mid] > tt) hi2 = mid;
            else lo = mid + 1;
        }
        K.sout[s] = lo;
    }
}

__global__ __launch_bounds__(TB, 2) void k_batch(const BTask* __restrict__ tasks, const BEntry* __restrict__ entries,
                                              int nentries, int64_t total, const Pass* __restrict__ passes,
                                              const Sweep* __restrict__ sweeps, const TOp* __restrict__ tops,
                                              BCtl* ctl, unsigned long long* done, unsigned long long* finished,
                                              unsigned long long wait_ticks) {
    extern __shared__ double2 tile[];
    __shared__ uint64_t scat_lo[256];
    __shared__ uint64_t scat_hi[64];
    __shared__ int64_t s_ticket;
    __shared__ int s_entry;
    __shared__ int s_ok;
    for (;;) {
        if (threadIdx.x == 0) {
            const int64_t tk = (int64_t)atomicAdd(&ctl->counter, 1ull);
            if (tk == 0) ctl->t0 = wall_clock64();
            s_ticket = tk;
            s_ok = 1;
            if (tk < total) {
                int lo = 0, hi = nentries - 1;  // last entry with prefix <= tk
                while (lo < hi) {
                    int mid = (lo + hi + 1) >> 1;
                    if (entries[mid].prefix <= tk) lo = mid;
                    else hi = mid - 1;
                }
                s_entry = lo;
                const int dep = entries[lo].dep;
                if (dep >= 0) {
                    // bounded wait: past `wait_ticks` of the device clock, or once
                    // another wait has failed, report and skip
                    const unsigned long long need = (unsigned long long)entries[dep].ntiles;
                    const unsigned long long t_begin = wall_clock64();
                    while (__hip_atomic_load(&done[dep], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
                        __builtin_amdgcn_s_sleep(4);
                        if (wall_clock64() - t_begin > wait_ticks ||
                            __hip_atomic_load(&ctl->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                            atomicOr(&ctl->err, 1ull);
                            s_ok = 0;
                            break;
                        }
                    }
                }
            }
        }
        __syncthreads();
        const int64_t tk = s_ticket;
        if (tk >= total) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        const BEntry& E = entries[s_entry];
        const BTask& K = tasks[E.task];
        if (s_ok) {
            if (E.pass >= 0)
                tile_body(K.sv, K.n, passes + K.pass0 + E.pass, sweeps, tops, (uint64_t)(tk - E.prefix), E.pass == 0,
                          tile, scat_lo, scat_hi);
            else
                finish_task(K, (double*)tile);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) {
            if (E.pass < 0) finished[E.task] = wall_clock64();
            __hip_atomic_fetch_add(&done[s_entry], 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_batch_run(const bsim_task* tasks, int n_tasks, double* latency_us) {
    if (n_tasks < 1 || !tasks) return fail(BSIM_EINVAL, "no tasks");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(BSIM_ENODEV, "no HIP device visible");
    // ---- plans
    std::vector<Plan> plans(n_tasks);
    std::vector<BTask> bt(n_tasks);
    std::vector<Pass> passes;
    std::vector<Sweep> sweeps;
    std::vector<TOp> tops;
    size_t arena = 0, zlen = 0, slen = 0;
    for (int i = 0; i < n_tasks; ++i) {
        const bsim_task& t = tasks[i];
        if (t.qubit_count < 4 || t.qubit_count > 26) return fail(BSIM_ENOTSUP, "batch circuits take 4..26 qubits");
        if (t.n_ops < 0 || (t.n_ops > 0 && !t.ops) || t.shots < 0 || !t.expect_z || (t.shots > 0 && !t.samples))
            return fail(BSIM_EINVAL, "bad batch task");
        std::vector<EOp> eops;
        if (int rc = lower_ops(t.qubit_count, BSIM_SV, t.ops, t.n_ops, eops, false)) return rc;
        TileConfig cfg = tile_config();
        if (int rc = make_plan(t.qubit_count, eops, cfg, plans[i])) return rc;
        for (auto& st : plans[i].steps)
            if (!st.tile) return fail(BSIM_ENOTSUP, "batch circuits use fusable gates only (dense <= 2 targets, diagonal <= 4)");
        if (plans[i].passes.empty()) {  // no gates: one empty pass still builds |0>
            Plan& p = plans[i];
            Pass P;
            memset(&P, 0, sizeof(P));
            P.T = std::min(tile_config().tile_bits, t.qubit_count);
            for (int b = 0; b < P.T; ++b) {
                P.lbits[b] = b;
                P.lmask |= 1ull << b;
            }
            P.sweep_begin = 0;
            P.nsweeps = 0;
            p.passes.push_back(P);
        }
        bt[i].n = t.qubit_count;
        bt[i].pass0 = (int)passes.size();
        bt[i].shots = t.shots;
        bt[i].seed = t.seed;
        const int sw0 = (int)sweeps.size(), op0 = (int)tops.size();
        for (Pass P : plans[i].passes) {
            P.sweep_begin += sw0;
            passes.push_back(P);
        }
        for (Sweep s : plans[i].sweeps) {
            s.op_begin += op0;
            s.op_end += op0;
            sweeps.push_back(s);
        }
        tops.insert(tops.end(), plans[i].tops.begin(), plans[i].tops.end());
        arena += (((size_t)1 << t.qubit_count) * sizeof(double2) + 255) & ~(size_t)255;
        zlen += (size_t)t.qubit_count;
        slen += (size_t)t.shots;
    }
    // ---- tickets, pass-wave by pass-wave, longest circuits first
    std::vector<int> order(n_tasks);
    for (int i = 0; i < n_tasks; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return plans[a].passes.size() > plans[b].passes.size();
    });
    std::vector<BEntry> entries;
    std::vector<int> last_entry(n_tasks, -1);
    size_t maxp = 0;
    for (auto& p : plans) maxp = std::max(maxp, p.passes.size());
    int64_t total = 0;
    for (size_t w = 0; w <= maxp; ++w)
        for (int i : order) {
            const size_t np = plans[i].passes.size();
            if (w > np) continue;
            BEntry e;
            memset(&e, 0, sizeof(e));
            e.task = i;
            e.pass = w < np ? (int)w : -1;
            e.dep = last_entry[i];
            e.ntiles = w < np ? ((int64_t)1 << (bt[i].n - plans[i].passes[w].T)) : 1;
            e.prefix = total;
            total += e.ntiles;
            last_entry[i] = (int)entries.size();
            entries.push_back(e);
        }
    // ---- device buffers
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off_tasks = 0;
    size_t off_entries = off_tasks + al(sizeof(BTask) * n_tasks);
    size_t off_passes = off_entries + al(sizeof(BEntry) * entries.size());
    size_t off_sweeps = off_passes + al(sizeof(Pass) * passes.size());
    size_t off_tops = off_sweeps + al(sizeof(Sweep) * std::max<size_t>(sweeps.size(), 1));
    size_t off_ctl = off_tops + al(sizeof(TOp) * std::max<size_t>(tops.size(), 1));
    size_t off_done = off_ctl + al(sizeof(BCtl));
    size_t off_fin = off_done + al(sizeof(unsigned long long) * entries.size());
    size_t off_z = off_fin + al(sizeof(unsigned long long) * n_tasks);
    size_t off_s = off_z + al(sizeof(double) * zlen);
    size_t off_arena = off_s + al(sizeof(int64_t) * std::max<size_t>(slen, 1));
    size_t bytes = off_arena + arena;
    char* dev = nullptr;
    hipError_t e = hipMalloc((void**)&dev, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(batch arena)");
    struct Guard {
        char* p;
        ~Guard() {
            if (p) (void)hipFree(p);
        }
    } guard{ dev };
    size_t a = 0, zo = 0, so = 0;
    for (int i = 0; i < n_tasks; ++i) {
        bt[i].sv = (double2*)(dev + off_arena + a);
        bt[i].zout = (double*)(dev + off_z) + zo;
        bt[i].sout = (int64_t*)(dev + off_s) + so;
        a += al(((size_t)1 << bt[i].n) * sizeof(double2));
        zo += (size_t)bt[i].n;
        so += (size_t)bt[i].shots;
    }
    hipStream_t s;
    BSIM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct SGuard {
        hipStream_t s;
        ~SGuard() { (void)hipStreamDestroy(s); }
    } sguard{ s };
    BSIM_HIP(hipMemcpyAsync(dev + off_tasks, bt.data(), sizeof(BTask) * n_tasks, hipMemcpyHostToDevice, s));
    BSIM_HIP(hipMemcpyAsync(dev + off_entries, entries.data(), sizeof(BEntry) * entries.size(), hipMemcpyHostToDevice, s));
    BSIM_HIP(hipMemcpyAsync(dev + off_passes, passes.data(), sizeof(Pass) * passes.size(), hipMemcpyHostToDevice, s));
    if (!sweeps.empty())
        BSIM_HIP(hipMemcpyAsync(dev + off_sweeps, sweeps.data(), sizeof(Sweep) * sweeps.size(), hipMemcpyHostToDevice, s));
    if (!tops.empty())
        BSIM_HIP(hipMemcpyAsync(dev + off_tops, tops.data(), sizeof(TOp) * tops.size(), hipMemcpyHostToDevice, s));
    std::vector<unsigned long long> zeros(std::max(entries.size(), (size_t)n_tasks) + 8, 0ull);
    BSIM_HIP(hipMemcpyAsync(dev + off_ctl, zeros.data(), sizeof(BCtl), hipMemcpyHostToDevice, s));
    BSIM_HIP(hipMemcpyAsync(dev + off_done, zeros.data(), sizeof(unsigned long long) * entries.size(),
                            hipMemcpyHostToDevice, s));
    if (int rc = prepare_tile_kernel()) return rc;
    int dev_id = 0, clock_khz = 0;
    BSIM_HIP(hipGetDevice(&dev_id));
    BSIM_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, dev_id));
    if (clock_khz <= 0) clock_khz = 100000;
    static bool attr = false;
    const int T = tile_config().tile_bits;
    const size_t shmem = std::max(sizeof(double2) * (size_t)((1 << T) + ((1 << T) >> 4)), sizeof(double) * ZROWS * TB);
    if (!attr) {
        BSIM_HIP(hipFuncSetAttribute((const void*)k_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(sizeof(double2) * ((1 << 13) + (1 << 9)))));
        attr = true;
    }
    int per_cu = 0;
    BSIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_batch, TB, shmem));
    if (per_cu < 1) per_cu = 1;
    const int grid = device_cu_count() * per_cu;
    BSIM_HIP(hipStreamSynchronize(s));
    hipLaunchKernelGGL(k_batch, dim3(grid), dim3(TB), shmem, s, (const BTask*)(dev + off_tasks),
                       (const BEntry*)(dev + off_entries), (int)entries.size(), total, (const Pass*)(dev + off_passes),
                       (const Sweep*)(dev + off_sweeps), (const TOp*)(dev + off_tops), (BCtl*)(dev + off_ctl),
                       (unsigned long long*)(dev + off_done), (unsigned long long*)(dev + off_fin),
                       (unsigned long long)clock_khz * 2000ull);  // 2 s
    BSIM_HIP(hipGetLastError());
    BSIM_HIP(hipStreamSynchronize(s));
    BCtl ctl;
    BSIM_HIP(hipMemcpy(&ctl, dev + off_ctl, sizeof(ctl), hipMemcpyDeviceToHost));
    if (ctl.err) return fail(BSIM_EHIP, "batch: a dependency wait exceeded its time limit");
    std::vector<double> z(zlen);
    BSIM_HIP(hipMemcpy(z.data(), dev + off_z, sizeof(double) * zlen, hipMemcpyDeviceToHost));
    std::vector<int64_t> shots(std::max<size_t>(slen, 1));
    if (slen) BSIM_HIP(hipMemcpy(shots.data(), dev + off_s, sizeof(int64_t) * slen, hipMemcpyDeviceToHost));
    std::vector<unsigned long long> fin(n_tasks);
    BSIM_HIP(hipMemcpy(fin.data(), dev + off_fin, sizeof(unsigned long long) * n_tasks, hipMemcpyDeviceToHost));
    zo = 0;
    so = 0;
    for (int i = 0; i < n_tasks; ++i) {
        memcpy(tasks[i].expect_z, z.data() + zo, sizeof(double) * bt[i].n);
        if (tasks[i].shots) memcpy(tasks[i].samples, shots.data() + so, sizeof(int64_t) * tasks[i].shots);
        zo += (size_t)bt[i].n;
        so += (size_t)tasks[i].shots;
        if (latency_us) latency_us[i] = fin[i] > ctl.t0 ? (double)(fin[i] - ctl.t0) * 1000.0 / (double)clock_khz : 0.0;
    }
    return BSIM_OK;
}

}  // extern "C"
