// Batched execution in one persistent launch (the reference's batch
// simulate, src/BraketSimulator.jl:347-395, runs one simulate per circuit on
// a pool of Julia tasks).
//
// Every circuit is planned on the host into fused tile passes (the same
// plans as bsim_evolve, tile = min(12, n) bits), then into stages:
//   pass p   one ticket per run of `chunk` tiles of pass p;
//   F1       one ticket per 4096-amplitude chunk of the final state: the
//            chunk's <Z> partial sums, its exact integer weight total and
//            the inclusive weight prefix of its 64-amplitude segments;
//   F2       one ticket: <Z_q> for every qubit, inclusive prefix of the
//            chunk totals;
//   F3       one ticket per 256 shots.
// Tickets are ordered stage-wave by stage-wave across circuits (circuits
// with the most passes first), so every ticket's dependency -- all tickets
// of the same circuit's previous stage -- has smaller ticket numbers.  A
// persistent grid pulls tickets from one atomic counter; a ticket waits
// (bounded in time) until its dependency's completion counter is full,
// runs, then releases its own counter.  Pass 0 builds |0...0> in LDS
// instead of reading the state.  Runs of several tiles per ticket are used
// when a wave holds many tiles: they cut the agent-scope release/acquire
// pairs, whose L2 write-back and invalidate dominate short tickets.
//
// Shots are drawn by inverse CDF over the exact integer weights
// W_i = floor(p_i * 2^n * 2^32) (p_i = re^2 + im^2 rounded as written):
// t = floor(r * sum(W) / 2^64) for r = (x1:x0) of Philox-4x32-10(counter
// (shot, 0, 0, 0), key seed), and the shot is the first i whose inclusive
// prefix W_0 + ... + W_i exceeds t -- found through the chunk prefix, the
// segment prefix and a scan of one segment.
#include <atomic>
#include <memory>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "fused.h"
#include "tile_kernel.h"

namespace bsim {

constexpr int TB = tile_threads<4>();  // batch workgroup size


constexpr int CHUNK_BITS = 12;  // F1 chunk: 4096 amplitudes
constexpr int SEG_BITS = 6;     // CDF segment: 64 amplitudes
constexpr int ZP = 1 + CHUNK_BITS;           // per-chunk Z partials: total + one per chunk bit
constexpr int ZR = ZP + (26 - CHUNK_BITS);   // F2 rows: + one per bit above the chunk
constexpr int MAX_CHUNK = 8;                 // tiles per ticket

enum : int32_t { K_F1 = -1, K_F2 = -2, K_F3 = -3 };

struct BTask {
    double2* sv;
    double* zpart;    // [nchunks][ZP]
    uint64_t* ctot;   // [nchunks]: chunk weight totals (F1), inclusive prefix (F2)
    uint64_t* seg;    // [nchunks][nseg]: inclusive segment prefix inside the chunk
    double* zout;
    int64_t* sout;
    int64_t shots;
    uint64_t seed;
    int32_t n, pass0;
    int32_t cbits;    // log2 amplitudes per chunk = min(12, n)
    int32_t sbits;    // log2 amplitudes per segment = min(6, cbits)
};

struct BEntry {
    int32_t task, kind;  // kind >= 0: tile pass index; K_F1/K_F2/K_F3
    int32_t dep;         // entry whose tickets must all be done first, -1: none
    int32_t chunk;       // tiles per ticket (tile passes)
    int64_t tickets, prefix;
    int32_t last, pad_;  // last entry of its circuit
};

struct BCtl {
    unsigned long long counter;
    unsigned long long err;
    unsigned long long t0;
    unsigned long long pad_;
};

__device__ __forceinline__ double prob(double2 a) { return __dadd_rn(__dmul_rn(a.x, a.x), __dmul_rn(a.y, a.y)); }

__device__ __forceinline__ uint64_t weight_of(double p, double scale) {
    const double x = __dmul_rn(p, scale);
    return x > 0 ? (uint64_t)x : 0;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// F1: chunk c.  `red`: LDS of ZP * TB doubles, `segs`: LDS of 64 uint64.
__device__ __noinline__ void f1_chunk(const BTask& K, int64_t c, double* red, uint64_t* segs) {
    const int L = 1 << K.cbits;
    const int nseg = 1 << (K.cbits - K.sbits);
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const double2* a = K.sv + (c << K.cbits);
    const double scale = ldexp(1.0, K.n + 32);
    double tot = 0, hb[4] = { 0, 0, 0, 0 };  // chunk bits 8..11 = row bits
    for (int j0 = 0, k = 0; j0 < L; j0 += TB, ++k) {
        const int j = j0 + t;
        double p = 0;
        uint64_t w = 0;
        if (j < L) {
            p = prob(a[j]);
            w = weight_of(p, scale);
        }
        tot += p;
        for (int r = 0; r < 4; ++r)
            if ((k >> r) & 1) hb[r] += p;
        const unsigned long long ws = wave_sum_u64(w);
        if (lane == 0 && j0 + 64 * wave < L) segs[(j0 + 64 * wave) >> K.sbits] = ws;
    }
    double v[ZP];
    v[0] = tot;
    for (int b = 0; b < 8; ++b) v[1 + b] = ((t >> b) & 1) ? tot : 0.0;
    for (int r = 0; r < 4; ++r) v[9 + r] = hb[r];
    for (int r = 0; r < ZP; ++r) red[r * TB + t] = v[r];
    __syncthreads();
    for (int o = TB / 2; o > 0; o >>= 1) {
        if (t < o)
            for (int r = 0; r < ZP; ++r) red[r * TB + t] += red[r * TB + t + o];
        __syncthreads();
    }
    if (t < ZP) K.zpart[c * ZP + t] = red[t * TB];
    if (wave == 0) {  // inclusive scan of the segment totals
        unsigned long long x = lane < nseg ? segs[lane] : 0ull;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned long long y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane < nseg) K.seg[c * nseg + lane] = x;
        if (lane == nseg - 1) K.ctot[c] = x;
    }
}

// F2: <Z_q> and the inclusive prefix of the chunk totals.  `red`: LDS of
// ZR * TB doubles, `sh`: LDS of TB uint64.
__device__ __noinline__ void f2_reduce(const BTask& K, double* red, uint64_t* sh) {
    const int n = K.n, hib = n - K.cbits;
    const int64_t nch = (int64_t)1 << hib;
    const int t = threadIdx.x;
    double v[ZR];
    for (int r = 0; r < ZR; ++r) v[r] = 0;
    for (int64_t c = t; c < nch; c += TB) {
        const double* zp = K.zpart + c * ZP;
        for (int r = 0; r < ZP; ++r) v[r] += zp[r];
        for (int b = 0; b < ZR - ZP; ++b)
            if (b < hib && ((c >> b) & 1)) v[ZP + b] += zp[0];
    }
    for (int r = 0; r < ZR; ++r) red[r * TB + t] = v[r];
    __syncthreads();
    for (int o = TB / 2; o > 0; o >>= 1) {
        if (t < o)
            for (int r = 0; r < ZR; ++r) red[r * TB + t] += red[r * TB + t + o];
        __syncthreads();
    }
    if (t < n) {
        const int bit = n - 1 - t;  // qubit t <-> amplitude bit n-1-t
        const double s = bit < K.cbits ? red[(1 + bit) * TB] : red[(ZP + bit - K.cbits) * TB];
        K.zout[t] = red[0] - 2.0 * s;
    }
    uint64_t carry = 0;
    for (int64_t c0 = 0; c0 < nch; c0 += TB) {
        const int64_t c = c0 + t;
        const uint64_t x = c < nch ? K.ctot[c] : 0;
        __syncthreads();
        sh[t] = x;
        __syncthreads();
        for (int o = 1; o < TB; o <<= 1) {
            const uint64_t y = t >= o ? sh[t - o] : 0;
            __syncthreads();
            sh[t] += y;
            __syncthreads();
        }
        if (c < nch) K.ctot[c] = carry + sh[t];
        carry += sh[TB - 1];
    }
}

// F3: shots s0 .. s0 + TB - 1.
__device__ __noinline__ void f3_shots(const BTask& K, int64_t s0) {
    const int64_t s = s0 + threadIdx.x;
    if (s >= K.shots) return;
    const int64_t nch = (int64_t)1 << (K.n - K.cbits);
    const int nseg = 1 << (K.cbits - K.sbits), S = 1 << K.sbits;
    const double scale = ldexp(1.0, K.n + 32);
    const uint64_t total = K.ctot[nch - 1];
    uint32_t ctr[4] = { (uint32_t)s, (uint32_t)((uint64_t)s >> 32), 0u, 0u };
    philox4x32_10(ctr, (uint32_t)K.seed, (uint32_t)(K.seed >> 32));
    const uint64_t r = ((uint64_t)ctr[1] << 32) | ctr[0];
    const uint64_t tt = __umul64hi(r, total);
    int64_t lo = 0, hi = nch - 1;  // first chunk whose inclusive prefix exceeds tt
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (K.ctot[mid] > tt) hi = mid;
        else lo = mid + 1;
    }
    const int64_t c = lo;
    const uint64_t t1 = tt - (c > 0 ? K.ctot[c - 1] : 0);
    const uint64_t* sg = K.seg + c * nseg;
    int g0 = 0, g1 = nseg - 1;
    while (g0 < g1) {
        const int mid = (g0 + g1) >> 1;
        if (sg[mid] > t1) g1 = mid;
        else g0 = mid + 1;
    }
    const uint64_t t2 = t1 - (g0 > 0 ? sg[g0 - 1] : 0);
    const int64_t base = (c << K.cbits) + ((int64_t)g0 << K.sbits);
    int64_t idx = base + S - 1;
    uint64_t run = 0;
    for (int j = 0; j < S; ++j) {
        run += weight_of(prob(K.sv[base + j]), scale);
        if (run > t2) {
            idx = base + j;
            break;
        }
    }
    K.sout[s] = idx;
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
                    const int mid = (lo + hi + 1) >> 1;
                    if (entries[mid].prefix <= tk) lo = mid;
                    else hi = mid - 1;
                }
                s_entry = lo;
                const int dep = entries[lo].dep;
                if (dep >= 0) {
                    // bounded wait: past `wait_ticks` of the device clock, or once
                    // another wait has failed, report and skip
                    const unsigned long long need = (unsigned long long)entries[dep].tickets;
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
        const int64_t local = tk - E.prefix;
        if (s_ok) {
            if (E.kind >= 0) {
                const uint64_t first = (uint64_t)local * (uint64_t)E.chunk;
                for (int c = 0; c < E.chunk; ++c) {
                    tile_body<4, false>(K.sv, K.n, passes + K.pass0 + E.kind, sweeps, tops, first + c, E.kind == 0, tile,
                              scat_lo, scat_hi);
                    __syncthreads();  // the next tile rewrites the scatter tables
                }
            } else if (E.kind == K_F1) {
                f1_chunk(K, local, (double*)tile, scat_lo);
            } else if (E.kind == K_F2) {
                f2_reduce(K, (double*)tile, scat_lo);
            } else {
                f3_shots(K, local * TB);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned long long old =
                __hip_atomic_fetch_add(&done[s_entry], 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (E.last && old + 1 == (unsigned long long)E.tickets) finished[E.task] = wall_clock64();
        }
        __syncthreads();
    }
}

// Host side -------------------------------------------------------------------

struct TaskPlan {
    Plan plan;
    int rc = BSIM_OK;
    std::string msg;
};

static void plan_task(const bsim_task& t, const TileConfig& base, TaskPlan& out) {
    std::vector<EOp> eops;
    if ((out.rc = lower_ops(t.qubit_count, BSIM_SV, t.ops, t.n_ops, eops, false)) != BSIM_OK) {
        out.msg = last_error_text();
        return;
    }
    TileConfig cfg = base;
    cfg.reg_bits = 4;  // the batch kernel runs 256-thread, 16-amplitude sweeps
    if ((out.rc = make_plan(t.qubit_count, eops, cfg, out.plan)) != BSIM_OK) {
        out.msg = last_error_text();
        return;
    }
    for (auto& st : out.plan.steps)
        if (!st.tile) {
            out.rc = BSIM_ENOTSUP;
            out.msg = "batch circuits use fusable gates only (dense <= 2 targets, diagonal <= 4)";
            return;
        }
    if (out.plan.passes.empty()) {  // no gates: one empty pass still builds |0>
        Pass P;
        memset(&P, 0, sizeof(P));
        P.T = std::min(base.tile_bits, t.qubit_count);
        P.R = 4;
        for (int b = 0; b < P.T; ++b) {
            P.lbits[b] = b;
            P.lmask |= 1ull << b;
        }
        out.plan.passes.push_back(P);
    }
}

template <class F>
static void parallel_for(int n, F f) {
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nth = std::max(1, std::min({ hw, 16, n / 8 }));
    if (nth == 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int w = 0; w < nth; ++w)
        th.emplace_back([&, w] {
            for (int i = w; i < n; i += nth) f(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace bsim

using namespace bsim;

extern "C" {

int bsim_batch_run(const bsim_task* tasks, int n_tasks, double* latency_us) {
    if (n_tasks < 1 || !tasks) return fail(BSIM_EINVAL, "no tasks");
    for (int i = 0; i < n_tasks; ++i) {
        const bsim_task& t = tasks[i];
        if (t.qubit_count < 4 || t.qubit_count > 26) return fail(BSIM_ENOTSUP, "batch circuits take 4..26 qubits");
        if (t.n_ops < 0 || (t.n_ops > 0 && !t.ops) || t.shots < 0 || !t.expect_z || (t.shots > 0 && !t.samples))
            return fail(BSIM_EINVAL, "bad batch task");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(BSIM_ENODEV, "no HIP device visible");
    // ---- plans, in parallel on the host
    std::vector<TaskPlan> tp(n_tasks);
    const TileConfig cfg = tile_config();  // one snapshot for planning and the LDS size
    parallel_for(n_tasks, [&](int i) { plan_task(tasks[i], cfg, tp[i]); });
    for (int i = 0; i < n_tasks; ++i)
        if (tp[i].rc != BSIM_OK) return fail(tp[i].rc, tp[i].msg);
    // ---- one table of passes / sweeps / ops for the whole batch
    std::vector<BTask> bt(n_tasks);
    std::vector<size_t> pass0(n_tasks + 1, 0), sw0(n_tasks + 1, 0), op0(n_tasks + 1, 0);
    for (int i = 0; i < n_tasks; ++i) {
        pass0[i + 1] = pass0[i] + tp[i].plan.passes.size();
        sw0[i + 1] = sw0[i] + tp[i].plan.sweeps.size();
        op0[i + 1] = op0[i] + tp[i].plan.tops.size();
    }
    std::vector<Pass> passes(pass0[n_tasks]);
    std::vector<Sweep> sweeps(std::max<size_t>(sw0[n_tasks], 1));
    // every entry is written below: left uninitialized (hundreds of MB for large batches)
    const size_t n_tops = std::max<size_t>(op0[n_tasks], 1);
    std::unique_ptr<TOp[]> tops_buf(new TOp[n_tops]);
    TOp* tops = tops_buf.get();
    parallel_for(n_tasks, [&](int i) {
        const Plan& p = tp[i].plan;
        for (size_t k = 0; k < p.passes.size(); ++k) {
            Pass P = p.passes[k];
            P.sweep_begin += (int32_t)sw0[i];
            passes[pass0[i] + k] = P;
        }
        for (size_t k = 0; k < p.sweeps.size(); ++k) {
            Sweep s = p.sweeps[k];
            s.op_begin += (int32_t)op0[i];
            s.op_end += (int32_t)op0[i];
            sweeps[sw0[i] + k] = s;
        }
        if (!p.tops.empty()) memcpy(&tops[op0[i]], p.tops.data(), sizeof(TOp) * p.tops.size());
    });
    // ---- device arena layout per circuit: state, Z partials, chunk totals, segment prefix
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    std::vector<size_t> off_sv(n_tasks);
    size_t arena = 0, zlen = 0, slen = 0;
    for (int i = 0; i < n_tasks; ++i) {
        const int n = tasks[i].qubit_count;
        const int cbits = std::min(CHUNK_BITS, n), sbits = std::min(SEG_BITS, cbits);
        const size_t nch = (size_t)1 << (n - cbits), nseg = (size_t)1 << (cbits - sbits);
        bt[i].n = n;
        bt[i].pass0 = (int32_t)pass0[i];
        bt[i].cbits = cbits;
        bt[i].sbits = sbits;
        bt[i].shots = tasks[i].shots;
        bt[i].seed = tasks[i].seed;
        off_sv[i] = arena;
        arena += al(((size_t)1 << n) * sizeof(double2)) + al(nch * ZP * sizeof(double)) +
                 al(nch * sizeof(uint64_t)) + al(nch * nseg * sizeof(uint64_t));
        zlen += (size_t)n;
        slen += (size_t)tasks[i].shots;
    }
    // ---- launch shape (needed for the ticket runs)
    if (int rc = prepare_tile_kernel()) return rc;
    int dev_id = 0, clock_khz = 0;
    BSIM_HIP(hipGetDevice(&dev_id));
    BSIM_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, dev_id));
    if (clock_khz <= 0) clock_khz = 100000;
    static std::atomic<bool> attr{ false };
    const int T = cfg.tile_bits;
    const size_t shmem = std::max(sizeof(double2) * (size_t)((1 << T) + ((1 << T) >> 4)), sizeof(double) * ZR * TB);
    if (!attr) {
        BSIM_HIP(hipFuncSetAttribute((const void*)k_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(sizeof(double2) * ((1 << 13) + (1 << 9)))));
        attr = true;
    }
    int per_cu = 0;
    BSIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_batch, TB, shmem));
    if (per_cu < 1) per_cu = 1;
    const int grid = device_cu_count() * per_cu;
    // ---- tickets: stage-wave by stage-wave, circuits with most passes first
    std::vector<int> order(n_tasks);
    for (int i = 0; i < n_tasks; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return tp[a].plan.passes.size() > tp[b].plan.passes.size(); });
    std::vector<int> stages(n_tasks), last_entry(n_tasks, -1);
    size_t maxs = 0;
    for (int i = 0; i < n_tasks; ++i) {
        stages[i] = (int)tp[i].plan.passes.size() + 2 + (tasks[i].shots > 0 ? 1 : 0);
        maxs = std::max(maxs, (size_t)stages[i]);
    }
    std::vector<BEntry> entries;
    int64_t total = 0;
    for (size_t w = 0; w < maxs; ++w) {
        int64_t wave_tiles = 0;
        for (int i : order)
            if (w < tp[i].plan.passes.size()) wave_tiles += (int64_t)1 << (bt[i].n - tp[i].plan.passes[w].T);
        int64_t run = 1;  // tiles per ticket: keep >= 4 tickets per resident workgroup in the wave
        while (run < MAX_CHUNK && wave_tiles / (run * 2) >= 4 * (int64_t)grid) run *= 2;
        for (int i : order) {
            if ((int)w >= stages[i]) continue;
            const int np = (int)tp[i].plan.passes.size();
            BEntry e;
            memset(&e, 0, sizeof(e));
            e.task = i;
            e.dep = last_entry[i];
            e.chunk = 1;
            const int64_t nch = (int64_t)1 << (bt[i].n - bt[i].cbits);
            if ((int)w < np) {
                const int64_t ntiles = (int64_t)1 << (bt[i].n - tp[i].plan.passes[w].T);
                e.kind = (int32_t)w;
                e.chunk = (int32_t)std::min(run, ntiles);
                e.tickets = ntiles / e.chunk;
            } else if ((int)w == np) {
                e.kind = K_F1;
                e.tickets = nch;
            } else if ((int)w == np + 1) {
                e.kind = K_F2;
                e.tickets = 1;
            } else {
                e.kind = K_F3;
                e.tickets = (tasks[i].shots + TB - 1) / TB;
            }
            e.last = (int)w == stages[i] - 1;
            e.prefix = total;
            total += e.tickets;
            last_entry[i] = (int)entries.size();
            entries.push_back(e);
        }
    }
    // ---- device buffers
    size_t off_tasks = 0;
    size_t off_entries = off_tasks + al(sizeof(BTask) * n_tasks);
    size_t off_passes = off_entries + al(sizeof(BEntry) * entries.size());
    size_t off_sweeps = off_passes + al(sizeof(Pass) * passes.size());
    size_t off_tops = off_sweeps + al(sizeof(Sweep) * sweeps.size());
    size_t off_ctl = off_tops + al(sizeof(TOp) * n_tops);
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
    size_t zo = 0, so = 0;
    for (int i = 0; i < n_tasks; ++i) {
        const size_t nch = (size_t)1 << (bt[i].n - bt[i].cbits);
        char* p = dev + off_arena + off_sv[i];
        bt[i].sv = (double2*)p;
        p += al(((size_t)1 << bt[i].n) * sizeof(double2));
        bt[i].zpart = (double*)p;
        p += al(nch * ZP * sizeof(double));
        bt[i].ctot = (uint64_t*)p;
        p += al(nch * sizeof(uint64_t));
        bt[i].seg = (uint64_t*)p;
        bt[i].zout = (double*)(dev + off_z) + zo;
        bt[i].sout = (int64_t*)(dev + off_s) + so;
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
    BSIM_HIP(hipMemcpyAsync(dev + off_sweeps, sweeps.data(), sizeof(Sweep) * sweeps.size(), hipMemcpyHostToDevice, s));
    BSIM_HIP(hipMemcpyAsync(dev + off_tops, tops, sizeof(TOp) * n_tops, hipMemcpyHostToDevice, s));
    std::vector<unsigned long long> zeros(std::max(entries.size(), (size_t)n_tasks) + 8, 0ull);
    BSIM_HIP(hipMemcpyAsync(dev + off_ctl, zeros.data(), sizeof(BCtl), hipMemcpyHostToDevice, s));
    BSIM_HIP(hipMemcpyAsync(dev + off_done, zeros.data(), sizeof(unsigned long long) * entries.size(),
                            hipMemcpyHostToDevice, s));
    BSIM_HIP(hipMemcpyAsync(dev + off_fin, zeros.data(), sizeof(unsigned long long) * n_tasks, hipMemcpyHostToDevice, s));
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
        if (latency_us)
            latency_us[i] = fin[i] > ctl.t0 ? (double)(fin[i] - ctl.t0) * 1000.0 / (double)clock_khz : 0.0;
    }
    return BSIM_OK;
}

}  // extern "C"
