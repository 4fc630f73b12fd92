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
// Gradient tasks (the parameter sweeps: a circuit with parametrized gates
// and a Pauli-sum observable H) share the launch.  Their stages after the
// forward passes are
//   LAM      one ticket per 4096-amplitude chunk: lambda = H psi on the
//            chunk and the chunk's part of <psi|H|psi>;
//   back b   one ticket per run of tiles of backward pass b: the fused
//            adjoint tile body (psi and lambda in LDS, U^dagger applied to
//            both, the <lambda|dU|psi> products of every parametrized op
//            summed per wave);
//   GFIN     one ticket: the fixed-order sums of the products and of <H>.
// The backward plans and the per-wave sums are those of the single-state
// fused adjoint (bsim_adjoint_gradient), so a batched gradient is the
// serial one up to the summation order of <H>.
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

#include "adj_tile_body.h"
#include "fused.h"
#include "reduce.h"
#include "tile_kernel.h"

namespace bsim {

constexpr int TB = tile_threads<4>();  // batch workgroup size


constexpr int CHUNK_BITS = 12;  // F1 chunk: 4096 amplitudes
constexpr int SEG_BITS = 6;     // CDF segment: 64 amplitudes
constexpr int ZP = 1 + CHUNK_BITS;           // per-chunk Z partials: total + one per chunk bit
constexpr int ZR = ZP + (26 - CHUNK_BITS);   // F2 rows: + one per bit above the chunk
constexpr int MAX_CHUNK = 8;                 // tiles per ticket

enum : int32_t { K_F1 = -1, K_F2 = -2, K_F3 = -3, K_LAM = -4, K_BWD = -5, K_GFIN = -6 };

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
    // gradient tasks (nwords > 0)
    double2* lam;
    const PWord* words;
    double* partials;  // [nslots * pm][ntiles][AT_W] per-wave products
    double* hdot;      // [nchunks]: chunk parts of <psi|H|psi>
    double* gres;      // [nslots * pm + 1]: summed products, then <H>
    int64_t ntiles;    // largest tile count of a backward pass
    int32_t nwords, bpass0, rows, pm;
};

struct BEntry {
    int32_t task, kind;  // kind >= 0: tile pass index; K_F1/K_F2/K_F3
    int32_t dep;         // entry whose tickets must all be done first, -1: none
    int32_t chunk;       // tiles per ticket (tile passes)
    int64_t tickets, prefix;
    int32_t last;        // last entry of its circuit
    int32_t sub;         // K_BWD: backward pass index
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

// LAM: lambda = H psi on chunk c, and the chunk's Re <psi|lambda> into
// hdot[c] (a fixed tree over the workgroup).  `red`: LDS of TB doubles.
__device__ __noinline__ void lam_chunk(const BTask& K, int64_t c, double* red) {
    const int L = 1 << K.cbits;
    const uint64_t base = (uint64_t)c << K.cbits;
    double acc = 0.0;
    for (int j = threadIdx.x; j < L; j += TB) {
        const uint64_t i = base + (uint64_t)j;
        double2 a = make_double2(0.0, 0.0);
        for (int w = 0; w < K.nwords; ++w) a = cfma(pphase(i, K.words[w]), K.sv[i ^ K.words[w].f], a);
        K.lam[i] = a;
        const double2 p = K.sv[i];
        acc += p.x * a.x + p.y * a.y;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = TB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) K.hdot[c] = red[0];
}

// GFIN: every product row summed in the fixed order of the single-state
// finish, then <H> over the chunks.  `sh`: LDS of TB doubles.
__device__ __noinline__ void grad_finish(const BTask& K, double* sh) {
    const int64_t per = K.ntiles * AT_W;
    for (int r = 0; r < K.rows; ++r) {
        const double v = adj_row_sum(K.partials + (size_t)r * per, per, sh);
        if (threadIdx.x == 0) K.gres[r] = v;
    }
    const double e = adj_row_sum(K.hdot, (int64_t)1 << (K.n - K.cbits), sh);
    if (threadIdx.x == 0) K.gres[K.rows] = e;
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
            } else if (E.kind == K_BWD) {
                const uint64_t first = (uint64_t)local * (uint64_t)E.chunk;
                for (int c = 0; c < E.chunk; ++c) {
                    adj_tile_body(K.sv, K.lam, K.n, passes + K.bpass0 + E.sub, sweeps, tops, K.partials, K.pm,
                                  K.ntiles, first + c, tile, scat_lo, scat_hi);
                    __syncthreads();  // the next tile rewrites the scatter tables
                }
            } else if (E.kind == K_F1) {
                f1_chunk(K, local, (double*)tile, scat_lo);
            } else if (E.kind == K_F2) {
                f2_reduce(K, (double*)tile, scat_lo);
            } else if (E.kind == K_F3) {
                f3_shots(K, local * TB);
            } else if (E.kind == K_LAM) {
                lam_chunk(K, local, (double*)tile);
            } else {
                grad_finish(K, (double*)tile);
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
    Plan plan;        // forward passes
    bool grad = false;
    Plan bplan;       // gradient tasks: backward passes
    int nslots = 0, pm = 0;
    std::vector<std::vector<int>> slot_params;  // parameter index of each (slot, p), slots in forward order
    std::vector<PWord> words;
    int rc = BSIM_OK;
    std::string msg;
};

static void plan_task(const bsim_task& t, const TileConfig& base, TaskPlan& out) {
    out.grad = t.n_words > 0;
    std::vector<EOp> eops;
    if ((out.rc = lower_ops(t.qubit_count, BSIM_SV, t.ops, t.n_ops, eops, out.grad)) != BSIM_OK) {
        out.msg = last_error_text();
        return;
    }
    TileConfig cfg = base;
    cfg.reg_bits = 4;  // the batch kernel runs 256-thread, 16-amplitude sweeps
    cfg.chains = false;  // and the full tile op set, which has no register chains
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
    if (!out.grad) return;
    for (const EOp& e : eops) {
        if (e.nparams == 0) continue;
        std::vector<int> ps;
        for (int p = 0; p < e.nparams; ++p) {
            if (e.pidx[p] < 0 || e.pidx[p] >= t.n_params) {
                out.rc = BSIM_EINVAL;
                out.msg = "batch gradient: param_index out of range";
                return;
            }
            ps.push_back(e.pidx[p]);
        }
        out.slot_params.push_back(ps);
    }
    std::vector<EOp> back;
    if ((out.rc = adjoint_back_plan(t.qubit_count, eops, back, out.bplan, out.nslots, out.pm, &base)) != BSIM_OK) {
        out.msg = last_error_text();
        return;
    }
    for (auto& st : out.bplan.steps)
        if (!st.tile) {
            out.rc = BSIM_ENOTSUP;
            out.msg = "batch gradients use fusable gates only (dense <= 2 targets, diagonal <= 4)";
            return;
        }
    for (const Pass& P : out.bplan.passes)
        if (P.R != 3 || P.T > 11) {
            out.rc = BSIM_ENOTSUP;
            out.msg = "batch gradient: backward pass shape";
            return;
        }
    out.pm = std::max(out.pm, 1);
    out.words.resize(t.n_words);
    for (int w = 0; w < t.n_words; ++w) {
        PauliMask pmk;
        if ((out.rc = parse_pauli(t.qubit_count, t.paulis + (size_t)w * t.qubit_count, pmk)) != BSIM_OK) {
            out.msg = last_error_text();
            return;
        }
        out.words[w] = { pmk.x | pmk.y, pmk.y, pmk.z, pmk.ny, 0, t.coeffs[w] };
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
        if (t.n_ops < 0 || (t.n_ops > 0 && !t.ops) || t.shots < 0) return fail(BSIM_EINVAL, "bad batch task");
        if (t.n_words > 0) {
            if (t.qubit_count < 5 || t.qubit_count > 26) return fail(BSIM_ENOTSUP, "batch gradients take 5..26 qubits");
            if (t.shots != 0 || !t.paulis || !t.coeffs || t.n_params < 0 || (t.n_params > 0 && !t.grad))
                return fail(BSIM_EINVAL, "bad batch gradient task");
        } else {
            if (t.qubit_count < 4 || t.qubit_count > 26) return fail(BSIM_ENOTSUP, "batch circuits take 4..26 qubits");
            if (t.n_words < 0 || !t.expect_z || (t.shots > 0 && !t.samples)) return fail(BSIM_EINVAL, "bad batch task");
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(BSIM_ENODEV, "no HIP device visible");
    // ---- plans, in parallel on the host
    std::vector<TaskPlan> tp(n_tasks);
    const TileConfig cfg = tile_config();  // one snapshot for planning and the LDS size
    parallel_for(n_tasks, [&](int i) { plan_task(tasks[i], cfg, tp[i]); });
    for (int i = 0; i < n_tasks; ++i)
        if (tp[i].rc != BSIM_OK) return fail(tp[i].rc, tp[i].msg);
    // ---- one table of passes / sweeps / ops for the whole batch: per task
    // its forward plan, then (gradient tasks) its backward plan
    std::vector<BTask> bt(n_tasks);
    std::vector<size_t> pass0(n_tasks + 1, 0), sw0(n_tasks + 1, 0), op0(n_tasks + 1, 0);
    for (int i = 0; i < n_tasks; ++i) {
        const Plan& f = tp[i].plan;
        const Plan& b = tp[i].bplan;
        pass0[i + 1] = pass0[i] + f.passes.size() + b.passes.size();
        sw0[i + 1] = sw0[i] + f.sweeps.size() + b.sweeps.size();
        op0[i + 1] = op0[i] + f.tops.size() + b.tops.size();
    }
    std::vector<Pass> passes(std::max<size_t>(pass0[n_tasks], 1));
    std::vector<Sweep> sweeps(std::max<size_t>(sw0[n_tasks], 1));
    // every entry is written below: left uninitialized (hundreds of MB for large batches)
    const size_t n_tops = std::max<size_t>(op0[n_tasks], 1);
    std::unique_ptr<TOp[]> tops_buf(new TOp[n_tops]);
    TOp* tops = tops_buf.get();
    parallel_for(n_tasks, [&](int i) {
        size_t pbase = pass0[i], sbase = sw0[i], obase = op0[i];
        for (const Plan* p : { &tp[i].plan, &tp[i].bplan }) {
            for (size_t k = 0; k < p->passes.size(); ++k) {
                Pass P = p->passes[k];
                P.sweep_begin += (int32_t)sbase;
                passes[pbase + k] = P;
            }
            for (size_t k = 0; k < p->sweeps.size(); ++k) {
                Sweep sw = p->sweeps[k];
                sw.op_begin += (int32_t)obase;
                sw.op_end += (int32_t)obase;
                sweeps[sbase + k] = sw;
            }
            if (!p->tops.empty()) memcpy(&tops[obase], p->tops.data(), sizeof(TOp) * p->tops.size());
            pbase += p->passes.size();
            sbase += p->sweeps.size();
            obase += p->tops.size();
        }
    });
    // ---- device arena layout per task.  Circuits: state, Z partials, chunk
    // totals, segment prefix.  Gradients: psi, lambda, chunk <H> parts,
    // results, Pauli words (their per-wave products live in one region of
    // their own, cleared before the launch).
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    std::vector<size_t> off_sv(n_tasks), off_part(n_tasks, 0);
    size_t arena = 0, zlen = 0, slen = 0, plen = 0;
    size_t bwd_lds = 0;
    for (int i = 0; i < n_tasks; ++i) {
        const int n = tasks[i].qubit_count;
        const int cbits = std::min(CHUNK_BITS, n), sbits = std::min(SEG_BITS, cbits);
        const size_t nch = (size_t)1 << (n - cbits), nseg = (size_t)1 << (cbits - sbits);
        const size_t amps = al(((size_t)1 << n) * sizeof(double2));
        bt[i].n = n;
        bt[i].pass0 = (int32_t)pass0[i];
        bt[i].cbits = cbits;
        bt[i].sbits = sbits;
        bt[i].shots = tasks[i].shots;
        bt[i].seed = tasks[i].seed;
        off_sv[i] = arena;
        if (tp[i].grad) {
            const TaskPlan& g = tp[i];
            int64_t nt = 1;
            for (const Pass& P : g.bplan.passes) {
                nt = std::max<int64_t>(nt, (int64_t)1 << (n - P.T));
                bwd_lds = std::max(bwd_lds, 2 * sizeof(double2) * (size_t)((1 << P.T) + ((1 << P.T) >> 4)));
            }
            bt[i].nwords = (int32_t)g.words.size();
            bt[i].bpass0 = (int32_t)(pass0[i] + g.plan.passes.size());
            bt[i].rows = g.nslots * g.pm;
            bt[i].pm = g.pm;
            bt[i].ntiles = nt;
            arena += 2 * amps + al(nch * sizeof(double)) + al(((size_t)bt[i].rows + 1) * sizeof(double)) +
                     al(g.words.size() * sizeof(PWord));
            off_part[i] = plen;
            plen += (size_t)bt[i].rows * (size_t)nt * AT_W;
        } else {
            arena += amps + al(nch * ZP * sizeof(double)) + al(nch * sizeof(uint64_t)) + al(nch * nseg * sizeof(uint64_t));
            zlen += (size_t)n;
            slen += (size_t)tasks[i].shots;
        }
    }
    // ---- launch shape (needed for the ticket runs)
    if (int rc = prepare_tile_kernel()) return rc;
    int dev_id = 0, clock_khz = 0;
    BSIM_HIP(hipGetDevice(&dev_id));
    BSIM_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, dev_id));
    if (clock_khz <= 0) clock_khz = 100000;
    static std::atomic<bool> attr{ false };
    const int T = cfg.tile_bits;
    const size_t shmem = std::max({ sizeof(double2) * (size_t)((1 << T) + ((1 << T) >> 4)), sizeof(double) * ZR * TB,
                                    bwd_lds });
    if (!attr) {
        BSIM_HIP(hipFuncSetAttribute((const void*)k_batch, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)(sizeof(double2) * ((1 << 13) + (1 << 9)))));
        attr = true;
    }
    int per_cu = 0;
    BSIM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_batch, TB, shmem));
    if (per_cu < 1) per_cu = 1;
    const int grid = device_cu_count() * per_cu;
    // ---- tickets: stage-wave by stage-wave, tasks with most stages first.
    // Circuit stages: forward passes, F1, F2, F3 (with shots).  Gradient
    // stages: forward passes, LAM, backward passes, GFIN.
    std::vector<int> stages(n_tasks), last_entry(n_tasks, -1);
    size_t maxs = 0;
    for (int i = 0; i < n_tasks; ++i) {
        const int np = (int)tp[i].plan.passes.size();
        stages[i] = tp[i].grad ? np + 2 + (int)tp[i].bplan.passes.size() : np + 2 + (tasks[i].shots > 0 ? 1 : 0);
        maxs = std::max(maxs, (size_t)stages[i]);
    }
    std::vector<int> order(n_tasks);
    for (int i = 0; i < n_tasks; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return stages[a] > stages[b]; });
    // the tile pass of stage w of task i (nullptr: not a tile stage) and, for
    // backward passes, its index
    auto stage_pass = [&](int i, int w, int& bwd) -> const Pass* {
        const int np = (int)tp[i].plan.passes.size();
        bwd = -1;
        if (w < np) return &tp[i].plan.passes[w];
        if (tp[i].grad && w > np && w <= np + (int)tp[i].bplan.passes.size()) {
            bwd = w - np - 1;
            return &tp[i].bplan.passes[bwd];
        }
        return nullptr;
    };
    std::vector<BEntry> entries;
    int64_t total = 0;
    for (size_t w = 0; w < maxs; ++w) {
        int64_t wave_tiles = 0;
        for (int i : order) {
            int bw;
            if (const Pass* P = stage_pass(i, (int)w, bw)) wave_tiles += (int64_t)1 << (bt[i].n - P->T);
        }
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
            int bw;
            if (const Pass* P = stage_pass(i, (int)w, bw)) {
                const int64_t ntiles = (int64_t)1 << (bt[i].n - P->T);
                e.kind = bw < 0 ? (int32_t)w : K_BWD;
                e.sub = bw < 0 ? 0 : bw;
                e.chunk = (int32_t)std::min(run, ntiles);
                e.tickets = ntiles / e.chunk;
            } else if (tp[i].grad) {
                e.kind = (int)w == np ? K_LAM : K_GFIN;
                e.tickets = (int)w == np ? nch : 1;
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
    size_t off_s = off_z + al(sizeof(double) * std::max<size_t>(zlen, 1));
    size_t off_p = off_s + al(sizeof(int64_t) * std::max<size_t>(slen, 1));
    size_t off_arena = off_p + al(sizeof(double) * std::max<size_t>(plen, 1));
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
    std::vector<std::vector<PWord>*> word_src(n_tasks, nullptr);
    std::vector<size_t> off_words(n_tasks, 0), off_gres(n_tasks, 0);
    size_t zo = 0, so = 0;
    for (int i = 0; i < n_tasks; ++i) {
        const size_t nch = (size_t)1 << (bt[i].n - bt[i].cbits);
        const size_t amps = al(((size_t)1 << bt[i].n) * sizeof(double2));
        char* p = dev + off_arena + off_sv[i];
        bt[i].sv = (double2*)p;
        p += amps;
        if (tp[i].grad) {
            bt[i].lam = (double2*)p;
            p += amps;
            bt[i].hdot = (double*)p;
            p += al(nch * sizeof(double));
            bt[i].gres = (double*)p;
            off_gres[i] = (size_t)(p - dev);
            p += al(((size_t)bt[i].rows + 1) * sizeof(double));
            bt[i].words = (const PWord*)p;
            off_words[i] = (size_t)(p - dev);
            word_src[i] = &tp[i].words;
            bt[i].partials = (double*)(dev + off_p) + off_part[i];
            continue;
        }
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
    for (int i = 0; i < n_tasks; ++i)
        if (word_src[i] && !word_src[i]->empty())
            BSIM_HIP(hipMemcpyAsync(dev + off_words[i], word_src[i]->data(), sizeof(PWord) * word_src[i]->size(),
                                    hipMemcpyHostToDevice, s));
    if (plen) BSIM_HIP(hipMemsetAsync(dev + off_p, 0, sizeof(double) * plen, s));
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
    std::vector<double> z(std::max<size_t>(zlen, 1));
    if (zlen) BSIM_HIP(hipMemcpy(z.data(), dev + off_z, sizeof(double) * zlen, hipMemcpyDeviceToHost));
    std::vector<int64_t> shots(std::max<size_t>(slen, 1));
    if (slen) BSIM_HIP(hipMemcpy(shots.data(), dev + off_s, sizeof(int64_t) * slen, hipMemcpyDeviceToHost));
    std::vector<unsigned long long> fin(n_tasks);
    BSIM_HIP(hipMemcpy(fin.data(), dev + off_fin, sizeof(unsigned long long) * n_tasks, hipMemcpyDeviceToHost));
    zo = 0;
    so = 0;
    std::vector<double> gres;
    for (int i = 0; i < n_tasks; ++i) {
        const bsim_task& t = tasks[i];
        if (tp[i].grad) {
            // grad[p] = 2 Re <lambda|dU|psi> summed over the ops carrying p
            gres.resize((size_t)bt[i].rows + 1);
            BSIM_HIP(hipMemcpy(gres.data(), dev + off_gres[i], sizeof(double) * gres.size(), hipMemcpyDeviceToHost));
            for (int p = 0; p < t.n_params; ++p) t.grad[p] = 0.0;
            for (size_t slot = 0; slot < tp[i].slot_params.size(); ++slot)
                for (size_t p = 0; p < tp[i].slot_params[slot].size(); ++p)
                    t.grad[tp[i].slot_params[slot][p]] += 2.0 * gres[slot * (size_t)bt[i].pm + p];
            if (t.expval) *t.expval = gres[(size_t)bt[i].rows];
        } else {
            memcpy(t.expect_z, z.data() + zo, sizeof(double) * bt[i].n);
            if (t.shots) memcpy(t.samples, shots.data() + so, sizeof(int64_t) * t.shots);
            zo += (size_t)bt[i].n;
            so += (size_t)t.shots;
        }
        if (latency_us)
            latency_us[i] = fin[i] > ctl.t0 ? (double)(fin[i] - ctl.t0) * 1000.0 / (double)clock_khz : 0.0;
    }
    return BSIM_OK;
}

}  // extern "C"
