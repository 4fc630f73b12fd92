"""Benchmark of the evolution hot path on one MI355X, ComplexF64 throughout,
every state resident in HBM.  Prints one JSON line.

Headline (``value``, gates/s): the 30-qubit brick-wall random circuit of
depth 20 (600 U + 290 CZ, 16 GiB state).  One step = reset to |0...0> +
fused evolve + one-pass <Z_i> for all i + alias table + 1,000,000 shots;
``value`` is the evolve's gates/s, the other phases are reported beside it.
The tile configuration is chosen by the autotuner on the circuit's first
five layers before the timed steps.

Other legs (each a workload named by its sizes): the 20-qubit RY + QFT
circuit with probabilities, <Z_i> and 10,000 shots; the 14-qubit noisy
density matrix in both Kraus semantics; the 24- and 28-qubit
hardware-efficient ansatz adjoint gradient against parameter shift; the
4096-circuit uneven batch run together with 256 16-qubit gradient sets in
one persistent launch; the per-gate and dense k-target kernels; HBM stream
probes (the measured bandwidth ceiling); and the oracle (the reference's
loops restated in C, OpenMP) on bounded samples of every workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--qubits 30] [--depth 20]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "braketsimulator.jl_amd", ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

METRIC = "30q SV HBM GB/s + gates/s; 24q adjoint grads/s vs param-shift at 1e-10"
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--shots", type=int, default=1000000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--autotune", type=int, default=1, help="1: autotune the tile configuration first")
    ap.add_argument("--qft-qubits", type=int, default=20, help="RY + QFT leg (0 skips it)")
    ap.add_argument("--adjoint-qubits", type=int, default=24, help="0 skips the adjoint leg")
    ap.add_argument("--adjoint-large", type=int, default=28,
                    help="qubits of the second adjoint workload (the same ansatz; 0 skips it)")
    ap.add_argument("--dense-k", type=int, default=28, help="qubits of the dense 3..6-target gate leg (0 skips it)")
    ap.add_argument("--ps-sample", type=int, default=24, help="parameters timed by parameter shift")
    ap.add_argument("--batch", type=int, default=4096, help="circuits in the batch leg (0 skips it)")
    ap.add_argument("--dm-qubits", type=int, default=14, help="noisy density-matrix leg (0 skips it)")
    ap.add_argument("--graph-reps", type=int, default=500, help="captured-evolution replays (0 skips the leg)")
    ap.add_argument("--per-gate", type=int, default=1, help="1: time the per-gate launch path on one layer")
    ap.add_argument("--probe", type=int, default=1, help="1: HBM stream probes (the bandwidth ceiling)")
    ap.add_argument("--sweep-sets", type=int, default=256, help="16-qubit ansatz parameter sets with gradients")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="profiling aid: one fused evolve of the workload and one 4 GiB stream copy, no JSON")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: libbsim has no multi-GPU path (a state "
              "would shard by its top index bits with qubit swaps over xGMI); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np

    from bsim import _lib as L
    from bsim import workloads as W
    from bsim.simulators import StateVectorSimulator, lower_instructions

    n, depth = a.qubits, a.depth
    if a.pmc_probe:
        pmc_probe(n, depth)
        return
    ins = W.brickwall(n, depth, seed=0)
    n_gates = len(ins)
    ol = lower_instructions(ins, L.BSIM_SV)
    sim = StateVectorSimulator(n, a.shots)
    lib = L.load()

    tuned = None
    if a.autotune:
        tuned = autotune_leg(sim, ins)
    stats = (ctypes.c_int64 * 4)()
    L.call("bsim_plan_stats", n, 0, ol.ptr, ol.n, stats)
    passes = int(stats[0])
    cfg = current_config()
    log(f"workload: {n}q brick-wall depth {depth}: {n_gates} gates -> {passes} fused passes ({cfg})")

    phases = {"evolve": 0.0, "expect_z": 0.0, "alias_build": 0.0, "shots": 0.0}

    def step(timed):
        t = time.perf_counter()
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
        L.call("bsim_synchronize", sim.handle)
        t1 = time.perf_counter()
        z = sim.expect_z_all()
        t2 = time.perf_counter()
        L.call("bsim_alias_build", sim.handle, L.ALIAS_PARALLEL, 0.0)
        L.call("bsim_synchronize", sim.handle)
        t3 = time.perf_counter()
        out = np.empty(a.shots, dtype=np.int64)
        L.call("bsim_sample", sim.handle, a.shots, ctypes.c_uint64(7), L.i64ptr(out))
        t4 = time.perf_counter()
        if timed:
            phases["evolve"] += t1 - t
            phases["expect_z"] += t2 - t1
            phases["alias_build"] += t3 - t2
            phases["shots"] += t4 - t3
        return z

    for i in range(a.warmup):
        log(f"warmup step {i}")
        step(False)
    L.call("bsim_synchronize", sim.handle)
    lib.bsim_set_profiling(1)
    lib.bsim_kernel_stats(1, None, None)
    log(f"timing {a.steps} steps")
    t0 = time.perf_counter()
    for i in range(a.steps):
        z = step(True)
    L.call("bsim_synchronize", sim.handle)
    dt = time.perf_counter() - t0
    launches, kms = ctypes.c_int64(), ctypes.c_double()
    lib.bsim_kernel_stats(1, ctypes.byref(launches), ctypes.byref(kms))
    lib.bsim_set_profiling(0)

    pass_bytes = 32.0 * 2 ** n
    flop_per_launch = algorithmic_flops(ins, n) / max(passes, 1)
    evolve_s = phases["evolve"] / a.steps
    value = n_gates / evolve_s
    avg_kernel_ms = kms.value / max(launches.value, 1)
    achieved = pass_bytes / (avg_kernel_ms * 1e-3) / 1e9
    ms_per_step = dt / a.steps * 1e3
    rec = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex128 (f64)",
        "data": "synthetic (SplitMix64 angles, seed 0)",
        "config": {
            "workload": f"brick-wall random circuit, {n} qubits, depth {depth}, ComplexF64 state "
                        f"({2 ** n * 16 / 2 ** 30:.0f} GiB), <Z_i> all i + {a.shots} alias-method shots; "
                        f"value = gates/s of the fused evolve",
            "qubits": n, "depth": depth, "gates": n_gates, "shots": a.shots,
            "fused_passes": passes, **cfg,
        },
        "evolve_ms": round(evolve_s * 1e3, 3),
        "hbm_GBps_algorithmic": round(passes * pass_bytes / evolve_s / 1e9, 1),
        "phase_ms": {k: round(v / a.steps * 1e3, 3) for k, v in phases.items()},
        "expect_z_GBps": round(16.0 * 2 ** n / (phases["expect_z"] / a.steps) / 1e9, 1),
        "shots_per_s": round(a.shots / (phases["shots"] / a.steps), 1),
        "autotune": tuned,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_tile_stream_lean (fused LDS-tiled gate pass, persistent, next tile prefetched)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(n, depth),
            "traffic_source": "profiles/k_tile_brickwall_30q.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
            "algorithmic_bytes_per_launch": int(pass_bytes),
            "avg_kernel_ms": round(avg_kernel_ms, 4),
            "launches_timed": int(launches.value),
            # the pass does many gates per HBM round trip and is bound by FP64
            # issue, not by HBM: its arithmetic against the FP64 vector peak
            "fp64": {
                "flop_per_launch": flop_per_launch,
                "achieved_tflops": round(flop_per_launch / (avg_kernel_ms * 1e-3) / 1e12, 2),
                "peak_tflops": FP64_PEAK_TFLOPS,
                "frac": round(flop_per_launch / (avg_kernel_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                "flop_model": "complex FP64 per amplitude: dense 1q 14 (2 cmul + 1 cadd), CZ 0 (sign flip)",
            },
        },
        "z_checksum": float(np.sum(z)),
    }
    del sim
    if a.probe:
        rec["stream_ceiling"] = probe_leg()
        ceil = rec["stream_ceiling"]["copy_GBps"]
        rec["roofline"]["frac_of_stream_copy"] = round(achieved / ceil, 4)
        rec["expect_z_frac_of_stream_read"] = round(rec["expect_z_GBps"] / rec["stream_ceiling"]["read_GBps"], 4)
    if a.per_gate:
        rec["per_gate"] = per_gate_leg(n, ins, rec.get("stream_ceiling"))
    if a.qft_qubits > 0:
        rec["qft"] = qft_leg(a.qft_qubits)
    if a.adjoint_qubits > 0:
        rec["adjoint"] = adjoint_leg(a.adjoint_qubits, a.ps_sample)
    if a.adjoint_large > 0:
        rec["adjoint_large"] = adjoint_large_leg(a.adjoint_large)
    if a.dense_k > 0:
        rec["dense_k"] = dense_k_leg(a.dense_k)
    if a.batch > 0:
        rec["batch"] = batch_leg(a.batch, a.sweep_sets)
    if a.dm_qubits > 0:
        rec["density_matrix"] = dm_leg(a.dm_qubits, a.autotune)
    if a.graph_reps > 0:
        rec["captured"] = graph_leg(a.graph_reps)
    if not a.no_cpu:
        log("cpu baseline (oracle) ...")
        rec["cpu_baseline"] = cpu_baselines(n, depth, ins, rec)
    print(json.dumps(rec), flush=True)


def cpu_baselines(n: int, depth: int, ins, rec: dict) -> dict:
    """The oracle -- the reference's per-gate, per-channel and result loops
    restated in C (oracle/refcpu.c), OpenMP over the host cores -- on a
    bounded sample of every workload, each phase timed and extrapolated to
    the full workload where the sample is a part of it.  The top-level value
    is the 30-qubit evolve in gates/s, the headline's unit."""
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(cores)
    configs = {}
    for name, fn in (("C1_qft_20q", lambda: _cpu_qft(20)),
                     ("C2_brickwall_30q", lambda: _cpu_brickwall(n, ins)),
                     ("C3_noisy_dm_14q", lambda: _cpu_dm(14)),
                     ("C4_hea_adjoint_24q", lambda: _cpu_hea(24)),
                     ("C5_batch", lambda: _cpu_batch())):
        t = time.perf_counter()
        configs[name] = fn()
        configs[name]["cpu_wall_s"] = round(time.perf_counter() - t, 2)
        log(f"cpu baseline {name}: {configs[name]}")
    c2 = configs["C2_brickwall_30q"]
    return {
        "value": c2["evolve_gates_per_s"],
        "unit": "gates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"30-qubit brick-wall: the first {c2['evolve_sample_gates']} of {len(ins)} gates through the "
                  f"reference per-gate loops restated in C (oracle/refcpu.c, {cores} OpenMP threads); reference "
                  f"Julia unavailable on the host.  Per-config records below (their samples named in each)",
        "configs": configs,
    }


def _timed(f, *args):
    t = time.perf_counter()
    r = f(*args)
    return r, time.perf_counter() - t


def _cpu_qft(nq: int) -> dict:
    """C1 in full, median of 3: evolve, probabilities, <Z_i> for every
    qubit, the reference-order alias table and 10,000 shots."""
    import numpy as np

    import oracle as O
    from bsim import workloads as W

    ins, _ = W.ry_qft(nq, seed=0)
    runs = []
    for rep in range(3):
        psi, t_ev = _timed(O.evolve_sv, nq, ins)
        p, t_p = _timed(O.sv_probabilities, psi, nq)
        t = time.perf_counter()
        for q in range(nq):
            O.sv_expect_pauli(psi, nq, "I" * q + "Z" + "I" * (nq - 1 - q))
        t_z = time.perf_counter() - t
        t = time.perf_counter()
        ap, al = O.alias_reference(p, float(p.sum()))
        idx, u = O.shot_draws(rep, nq, 10000)
        O.lookup(ap, al, idx, u)
        t_s = time.perf_counter() - t
        runs.append((t_ev, t_p, t_z, t_s))
    m = np.median(np.array(runs), axis=0)
    return {"sample": f"full workload ({len(ins)} gates), median of 3", "evolve_ms": round(m[0] * 1e3, 2),
            "probabilities_ms": round(m[1] * 1e3, 2), "expect_z_ms": round(m[2] * 1e3, 2),
            "alias_and_shots_ms": round(m[3] * 1e3, 2), "total_ms": round(float(m.sum()) * 1e3, 2),
            "evolve_gates_per_s": round(len(ins) / m[0], 1)}


def _cpu_brickwall(n: int, ins, budget_s: float = 8.0) -> dict:
    """C2: the evolve on the gates that fit in ``budget_s`` (extrapolated to
    the circuit), then on that state: one <Z_q> (x n), the probabilities,
    the reference-order alias table on 2^26 of them (x 2^(n-26)) and the
    reference's n >= 30 direct sampling (StatsBase.sample, a linear scan)
    on 3 shots, extrapolated to 1,000,000."""
    import numpy as np

    import oracle as O
    from bsim.gates import kernel_call

    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi.fill(0)  # touch every page outside the timed region
    psi[0] = 1
    t0 = time.perf_counter()
    done = 0
    for i in ins:
        O.apply_call(psi, n, kernel_call(i.operator, i.target))
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    t_ev = time.perf_counter() - t0
    _, t_z = _timed(O.sv_expect_pauli, psi, n, "Z" + "I" * (n - 1))
    p, t_p = _timed(O.sv_probabilities, psi, n)
    del psi
    m = min(26, n)
    sub = np.ascontiguousarray(p[: 2 ** m])
    _, t_a = _timed(O.alias_reference, sub, float(sub.sum()))
    del sub
    u = np.array([0.25, 0.5, 0.75])
    _, t_d = _timed(O.direct_samples, p, float(p.sum()), u)
    del p
    rate = done / t_ev
    per_shot = t_d / len(u)
    return {"sample": f"evolve: first {done} of {len(ins)} gates; <Z>: 1 of {n} qubits; alias: 2^{m} of 2^{n} "
                      f"weights; direct sampling: 3 shots (u = 0.25, 0.5, 0.75) of 10^6",
            "evolve_sample_gates": done, "evolve_gates_per_s": round(rate, 3),
            "evolve_s_extrapolated": round(len(ins) / rate, 2),
            "expect_z_all_s_extrapolated": round(t_z * n, 3),
            "probabilities_s": round(t_p, 3),
            "alias_table_s_extrapolated": round(t_a * 2 ** (n - m), 3),
            "direct_sampling_s_per_shot": round(per_shot, 4),
            "direct_sampling_1e6_shots_s_extrapolated": round(per_shot * 1e6, 1)}


def _cpu_dm(nq: int) -> dict:
    """C3, reference Kraus semantics: the first op of each kind of the noisy
    brick-wall (U, Depolarizing, CZ, TwoQubitDepolarizing,
    AmplitudeDamping) applied to the 4 GiB density matrix (best of two), the evolve
    estimated as the sum over the circuit of its ops' kind times; one
    <Z_i Z_i+1> (x (n-1)) and the diagonal probabilities."""
    import numpy as np

    import oracle as O
    from bsim import workloads as W

    ins = W.brickwall(nq, 10, seed=0, noise=True)
    kinds = {}
    for i in ins:
        kinds.setdefault(type(i.operator).__name__, []).append(i)
    rho = np.zeros(4 ** nq, dtype=np.complex128)
    rho.fill(0)  # touch every page outside the timed ops
    rho[0] = 1
    per_kind = {}
    for k, lst in kinds.items():  # best of two applications
        per_kind[k] = min(_timed(O.evolve_dm, nq, lst[:1], "reference", rho)[1] for _ in range(2))
    est = sum(per_kind[k] * len(lst) for k, lst in kinds.items())
    _, t_zz = _timed(O.dm_expect_pauli, rho, nq, "ZZ" + "I" * (nq - 2))
    _, t_p = _timed(O.dm_probabilities, rho, nq)
    del rho
    return {"sample": "one op of each kind (" + ", ".join(f"{k} x{len(v)}" for k, v in kinds.items())
                      + f" in the circuit); one of {nq - 1} <Z_i Z_i+1>",
            "op_ms": {k: round(v * 1e3, 1) for k, v in per_kind.items()},
            "evolve_s_estimated": round(est, 2), "ops_per_s": round(len(ins) / est, 2),
            "zz_all_s_extrapolated": round(t_zz * (nq - 1), 3), "probabilities_s": round(t_p, 3)}


def _cpu_hea(nq: int) -> dict:
    """C4: one energy evaluation of the 24-qubit ansatz -- the first layer's
    gates (x 8 layers) and one ZZ and one X term of <H> (x their counts) --
    then the parameter-shift gradient (2 energies per parameter) built from
    it; the reference has no adjoint."""
    import numpy as np

    import oracle as O
    from bsim import workloads as W
    from bsim.gates import kernel_call

    ins, pmap, th = W.hea(nq, 8, seed=0)
    words, _ = W.tfim(nq)
    layer = len(ins) // 8
    psi = np.zeros(2 ** nq, dtype=np.complex128)
    psi.fill(0)
    psi[0] = 1
    t = time.perf_counter()
    for i in ins[:layer]:
        O.apply_call(psi, nq, kernel_call(i.operator, i.target))
    t_layer = time.perf_counter() - t
    zz = [w for w in words if "Z" in w]
    xx = [w for w in words if "X" in w]
    _, t_zz = _timed(O.sv_expect_pauli, psi, nq, zz[0])
    _, t_x = _timed(O.sv_expect_pauli, psi, nq, xx[0])
    del psi
    t_energy = 8 * t_layer + len(zz) * t_zz + len(xx) * t_x
    t_ps = 2 * len(th) * t_energy
    return {"sample": f"first layer of 8 ({layer} of {len(ins)} gates); 1 of {len(zz)} ZZ and 1 of {len(xx)} X terms",
            "energy_s_estimated": round(t_energy, 3),
            "param_shift_gradient_s_estimated": round(t_ps, 1),
            "param_shift_grads_per_s": round(1.0 / t_ps, 6), "params": len(th)}


def _cpu_batch() -> dict:
    """C5: 16 circuits of the uneven batch (evolve + <Z_q> for every qubit +
    reference-order alias + 1000 shots each), extrapolated to the batch by
    gate count; one energy of the 16-qubit ansatz, extrapolated to the
    parameter-shift gradients of 256 sets (2 energies per parameter)."""
    import numpy as np

    import oracle as O
    from bsim import workloads as W

    count = 4096
    spec = W.batch_tasks(count, seed=0)
    pick = spec[:: count // 16][:16]
    t = time.perf_counter()
    g_sample = 0
    for n, d, s in pick:
        ins = W.brickwall(n, d, s)
        g_sample += len(ins)
        psi = O.evolve_sv(n, ins)
        for q in range(n):
            O.sv_expect_pauli(psi, n, "I" * q + "Z" + "I" * (n - 1 - q))
        p = O.sv_probabilities(psi, n)
        ap, al = O.alias_reference(p, float(p.sum()))
        idx, u = O.shot_draws(s, n, 1000)
        O.lookup(ap, al, idx, u)
    t_sample = time.perf_counter() - t
    # brick-wall gates: n U per layer, then n // 2 CZ (even layers) or (n - 1) // 2 (odd)
    g_total = sum(n * d + (n // 2) * ((d + 1) // 2) + ((n - 1) // 2) * (d // 2) for n, d, _ in spec)
    t_circ = t_sample * g_total / g_sample
    words, _ = W.tfim(16)
    ins, pmap, th = W.hea(16, 8, seed=0)
    t = time.perf_counter()
    psi = O.evolve_sv(16, ins)
    for w in words:
        O.sv_expect_pauli(psi, 16, w)
    t_energy = time.perf_counter() - t
    t_grad = 256 * 2 * len(th) * t_energy
    total = t_circ + t_grad
    return {"sample": f"16 of {count} circuits ({g_sample} gates, extrapolated by gate count); one 16-qubit energy "
                      f"(x 2 x {len(th)} parameters x 256 sets)",
            "circuits_s_extrapolated": round(t_circ, 2), "gradients_s_extrapolated": round(t_grad, 2),
            "tasks_per_s": round((count + 256) / total, 2)}


def current_config() -> dict:
    """The library's tile configuration: tile bits, contiguous low bits,
    register bits per sweep, fusion limit."""
    from bsim import _lib as L

    t, lo, rb, fl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    L.call("bsim_get_tile_config", ctypes.byref(t), ctypes.byref(lo))
    L.call("bsim_set_sweep_bits", 0, ctypes.byref(rb))  # 0: read only
    L.call("bsim_set_fusion_limit", 0, ctypes.byref(fl))
    L.call("bsim_set_fusion_limit", fl.value, None)
    mw = ctypes.c_int()
    L.call("bsim_set_merge_width", -1, ctypes.byref(mw))
    return {"tile_bits": t.value, "low_bits": lo.value, "sweep_bits": rb.value, "fusion_limit": fl.value,
            "merge_width": mw.value}


def set_config(c: dict):
    from bsim import _lib as L

    L.call("bsim_set_tile_config", c["tile_bits"], c["low_bits"])
    L.call("bsim_set_sweep_bits", c["sweep_bits"], None)
    L.call("bsim_set_fusion_limit", c["fusion_limit"], None)
    L.call("bsim_set_merge_width", c["merge_width"], None)


def autotune_leg(sim, ins) -> dict:
    """bsim_autotune on the workload (every tile width, contiguous low bits
    and sweep width, then the fusion width at the best shape); the winner
    becomes the library configuration.  Reports it with the default
    configuration's evolve time."""
    from bsim import _lib as L
    from bsim.simulators import lower_instructions

    default = current_config()
    ol = lower_instructions(ins, L.BSIM_SV)

    def timed():
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_synchronize", sim.handle)
        t = time.perf_counter()
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
        L.call("bsim_synchronize", sim.handle)
        return time.perf_counter() - t

    timed()
    t_default = min(timed() for _ in range(2))
    t0 = time.perf_counter()
    best = sim.autotune(ins)
    t_search = time.perf_counter() - t0
    log(f"autotune: {best} in {t_search:.1f} s; default {default} {t_default * 1e3:.1f} ms")
    return {"chosen": {k: best[k] for k in ("tile_bits", "low_bits", "sweep_bits", "fusion_limit", "merge_width")},
            "chosen_ms": round(best["ms"], 3), "default": default, "default_ms": round(t_default * 1e3, 3),
            "search_s": round(t_search, 2)}


def probe_leg() -> dict:
    """The HBM ceiling: bsim_stream_probe's plain copy, read and write
    kernels over 4 GiB (copy: 4 GiB read + 4 GiB written), best of 5
    launches at 4, 8 and 16 workgroups per CU."""
    from bsim import _lib as L

    cus = ctypes.c_int()
    L.call("bsim_device_info", None, 0, ctypes.byref(cus), None)
    nbytes = 2 ** 32
    out = {"bytes": nbytes}
    for mode, name, moved in ((0, "copy", 2 * nbytes), (1, "read", nbytes), (2, "write", nbytes)):
        best = float("inf")
        for per_cu in (4, 8, 16):
            ms = ctypes.c_double()
            L.call("bsim_stream_probe", mode, nbytes, per_cu * cus.value, 5, ctypes.byref(ms))
            best = min(best, ms.value)
        out[f"{name}_GBps"] = round(moved / (best * 1e-3) / 1e9, 1)
        out[f"{name}_frac_of_peak"] = round(moved / (best * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    log(f"stream probes: copy {out['copy_GBps']} GB/s, read {out['read_GBps']}, write {out['write_GBps']}")
    return out


def qft_leg(nq: int) -> dict:
    """The 20-qubit RY + textbook QFT circuit (240 gates): fused evolve,
    probabilities, <Z_i> for every qubit and 10,000 alias-method shots
    (exact-integer parallel table), median of 5 after a warm-up; then once
    with the reference-order table.  The 16 MiB state stays in cache:
    launch-bound."""
    import numpy as np

    from bsim import _lib as L
    from bsim import workloads as W
    from bsim.simulators import StateVectorSimulator, lower_instructions

    ins, _ = W.ry_qft(nq, seed=0)
    ol = lower_instructions(ins, L.BSIM_SV)
    sim = StateVectorSimulator(nq, 10000)
    runs = []
    for rep in range(6):
        t0 = time.perf_counter()
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
        L.call("bsim_synchronize", sim.handle)
        t1 = time.perf_counter()
        p = sim.probabilities()
        t2 = time.perf_counter()
        z = sim.expect_z_all()
        t3 = time.perf_counter()
        shots = sim.samples(seed=rep)
        t4 = time.perf_counter()
        if rep:
            runs.append((t1 - t0, t2 - t1, t3 - t2, t4 - t3, t4 - t0))
    med = np.median(np.array(runs), axis=0)
    # the reference-order table (bit-compatible with make_alias_table!): a
    # sequential LIFO pairing, one lane
    t = time.perf_counter()
    sim.samples(seed=0, alias_mode=L.ALIAS_REFERENCE)
    t_ref = time.perf_counter() - t
    del sim
    log(f"qft {nq}q: evolve {med[0] * 1e3:.2f} ms, total {med[4] * 1e3:.2f} ms")
    return {"workload": f"RY + QFT, {nq} qubits ({len(ins)} gates), probabilities + <Z_i> + 10000 shots",
            "evolve_ms": round(med[0] * 1e3, 3), "probabilities_ms": round(med[1] * 1e3, 3),
            "expect_z_ms": round(med[2] * 1e3, 3), "alias_and_shots_ms": round(med[3] * 1e3, 3),
            "total_ms": round(med[4] * 1e3, 3), "gates_per_s": round(len(ins) / med[0], 1),
            "alias_reference_order_and_shots_ms": round(t_ref * 1e3, 3),
            "z_checksum": float(np.sum(z)), "prob_sum": float(np.sum(p)), "shots_drawn": int(shots.size)}


def adjoint_leg(nq: int, ps_sample: int):
    """24-qubit hardware-efficient ansatz (8 layers, 384 parameters) with the
    open-chain transverse-field Ising H: full adjoint gradient vs parameter
    shift through the plain GPU path on a sample of parameters.

    The timed figure (adjoint_s) is an optimizer's step: the circuit is
    lowered once, and each step binds a new parameter vector into the op
    list (OpList.bind: every rotation's matrix and derivative recomputed)
    and takes the gradient.  adjoint_from_instructions_s is the same
    gradient from the instruction list, lowered again inside every call."""
    import numpy as np

    from bsim import _lib as L
    from bsim import adjoint as A
    from bsim import workloads as W
    from bsim.simulators import StateVectorSimulator, lower_instructions

    ins, pmap, th = W.hea(nq, 8, seed=0)
    words, coeffs = W.tfim(nq)
    sim = StateVectorSimulator(nq, 0)
    log(f"adjoint: {nq} qubits, {len(th)} parameters")
    A.adjoint_gradient(sim, ins, pmap, words, coeffs)  # warm-up
    reps = 5
    t = time.perf_counter()
    ol = lower_instructions(ins, L.BSIM_SV, param_map=pmap)
    ol.ptr
    t_lower = time.perf_counter() - t
    A.adjoint_gradient(sim, ol.bind(th + 0.5), pmap, words, coeffs)  # warm-up of the bound path
    steps = [th + 0.01 * (k + 1) for k in range(reps)]
    t = time.perf_counter()
    for tk in steps:
        A.adjoint_gradient(sim, ol.bind(tk), pmap, words, coeffs)
    t_adj = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    for _ in range(reps):
        ev, grad = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    t_ins = (time.perf_counter() - t) / reps
    ev_b, grad_b = A.adjoint_gradient(sim, ol.bind(th), pmap, words, coeffs)
    # the same gradient with one backward kernel per op
    A.adjoint_gradient(sim, ins, pmap, words, coeffs, fused=False)
    t = time.perf_counter()
    for _ in range(reps):
        _, grad_op = A.adjoint_gradient(sim, ins, pmap, words, coeffs, fused=False)
    t_op = (time.perf_counter() - t) / reps
    log(f"adjoint: bind + gradient {t_adj * 1e3:.1f} ms, from instructions {t_ins * 1e3:.1f} ms, "
        f"per-op backward {t_op * 1e3:.1f} ms")
    # parameter shift on the first `ps_sample` parameters (2 evolutions each)
    sample = list(range(min(ps_sample, len(th))))
    t = time.perf_counter()
    ps = np.zeros(len(sample))
    for k, p in enumerate(sample):
        tp = th.copy(); tp[p] += np.pi / 2
        tm = th.copy(); tm[p] -= np.pi / 2
        ps[k] = 0.5 * (A.energy(sim, A.with_params(ins, pmap, tp), words, coeffs)
                       - A.energy(sim, A.with_params(ins, pmap, tm), words, coeffs))
    t_ps_per_param = (time.perf_counter() - t) / len(sample)
    t_ps = t_ps_per_param * len(th)
    return {
        "workload": f"HEA {nq} qubits x 8 layers (RY RZ + CNOT ladder), {len(th)} params, TFIM H (47 Pauli words)",
        "expval": ev,
        "adjoint_s": round(t_adj, 4),
        "grads_per_s": round(1.0 / t_adj, 3),
        "timing": f"mean of {reps} optimizer steps: bind a new parameter vector + gradient",
        "adjoint_from_instructions_s": round(t_ins, 4),
        "lowering_once_s": round(t_lower, 4),
        "max_abs_diff_bound_vs_instructions": float(max(abs(ev_b - ev), np.max(np.abs(grad_b - grad)))),
        "per_op_backward_s": round(t_op, 4),
        "max_abs_diff_fused_vs_per_op": float(np.max(np.abs(grad - grad_op))),
        "param_shift_s_extrapolated": round(t_ps, 3),
        "param_shift_grads_per_s": round(1.0 / t_ps, 4),
        "speedup_vs_param_shift": round(t_ps / t_adj, 1),
        "max_abs_diff_vs_param_shift": float(np.max(np.abs(grad[sample] - ps))),
        "param_shift_sample": f"first {len(sample)} of {len(th)} parameters, 2 fused evolutions + <H> each",
    }


def algorithmic_flops(ins, n: int) -> float:
    """FP64 flops an evolve of ``ins`` needs: a dense k-qubit gate costs
    2^k complex multiplies and 2^k - 1 complex adds per amplitude (6 and 2
    flops), a diagonal one multiply unless its entries are +-1, controls
    scale by the fraction of amplitudes they select."""
    import numpy as np

    from bsim.gates import kernel_call

    total = 0.0
    for i in ins:
        call = kernel_call(i.operator, i.target)
        m = np.asarray(call.matrix)
        frac = 0.5 ** len(getattr(call, "controls", []) or [])
        if m.ndim == 2 and np.count_nonzero(m - np.diag(np.diagonal(m))) == 0:
            m = np.diagonal(m)
        if getattr(call, "diagonal", False) or m.ndim == 1:
            d = m.reshape(-1)
            per = 0.0 if np.all(np.isin(d, (1, -1))) else 6.0
        else:
            k2 = m.shape[0]
            per = 6.0 * k2 + 2.0 * (k2 - 1)
        total += per * frac * 2 ** n
    return total


def adjoint_large_leg(nq: int):
    """The 8-layer hardware-efficient ansatz at ``nq`` qubits (28: 4 GiB per
    vector): fused adjoint gradients (best of 2 after a warm-up) and one
    gradient with one backward kernel per op, for the difference."""
    import numpy as np

    from bsim import adjoint as A
    from bsim import workloads as W
    from bsim.simulators import StateVectorSimulator

    ins, pmap, th = W.hea(nq, 8, seed=0)
    words, coeffs = W.tfim(nq)
    sim = StateVectorSimulator(nq, 0)
    log(f"adjoint: {nq} qubits, {len(th)} parameters")
    _, g = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    best = float("inf")
    for _ in range(2):
        t = time.perf_counter()
        ev, g = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
        best = min(best, time.perf_counter() - t)
    t = time.perf_counter()
    _, g_op = A.adjoint_gradient(sim, ins, pmap, words, coeffs, fused=False)
    t_op = time.perf_counter() - t
    log(f"adjoint {nq}q: fused {best * 1e3:.1f} ms, per-op {t_op * 1e3:.1f} ms")
    del sim
    return {
        "workload": f"HEA {nq} qubits x 8 layers, {len(th)} params, TFIM H",
        "expval": ev, "adjoint_s": round(best, 4), "grads_per_s": round(1.0 / best, 3),
        "per_op_backward_s": round(t_op, 4), "max_abs_diff_fused_vs_per_op": float(np.max(np.abs(g - g_op))),
    }


def dense_k_leg(nq: int):
    """Single dense gates of k = 3..6 targets (random unitaries) at ``nq``
    qubits through the per-gate path: k = 4..6 run on the FP64 matrix cores,
    k = 3 through LDS.  Low, high and scattered targets; 12 timed gates each
    after one warm-up evolve of 4.  Bytes per gate: 32 B x 2^n (read + write)."""
    import numpy as np

    from bsim import _lib as L
    from bsim import gates as G
    from bsim.simulators import Instruction, StateVectorSimulator

    rng = np.random.default_rng(0)
    sim = StateVectorSimulator(nq, 0)
    out = {"workload": f"{nq} qubits, one dense k-target gate per launch", "ms_per_gate": {}}
    for k in (3, 4, 5, 6):
        for name, tg in (("low", list(range(k))), ("high", [nq - 1 - i for i in range(k)]),
                         ("scattered", [0, 9, 15, 22, 27, 5][:k])):
            ins = []
            for _ in range(4):
                z = rng.normal(size=(2 ** k, 2 ** k)) + 1j * rng.normal(size=(2 ** k, 2 ** k))
                ins.append(Instruction(G.Unitary(np.linalg.qr(z)[0]), tg))
            sim.evolve(ins, mode=1)
            L.call("bsim_synchronize", sim._h)
            t = time.perf_counter()
            for _ in range(3):
                sim.evolve(ins, mode=1)
            L.call("bsim_synchronize", sim._h)
            dt = (time.perf_counter() - t) / 12
            out["ms_per_gate"][f"k{k}_{name}"] = round(dt * 1e3, 3)
            log(f"dense k={k} {name} {tg}: {dt * 1e3:.2f} ms per gate, {32 * 2 ** nq / dt / 1e9:.0f} GB/s")
    del sim
    return out


def measured_traffic(n: int, depth: int):
    """HBM bytes per k_tile launch measured by hardware counters for the
    30-qubit depth-20 workload (committed profile); None for other shapes."""
    f = ROOT / "profiles" / "k_tile_brickwall_30q.md"
    if (n, depth) != (30, 20) or not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def dm_leg(nq: int, autotune: int):
    """Noisy density matrix: the brick-wall of the state-vector leg at nq
    qubits, depth 10, with Depolarizing(0.01) after every U,
    TwoQubitDepolarizing(0.02) after every CZ and AmplitudeDamping(0.005) on
    every qubit per layer; <Z_i Z_i+1> for all pairs and the diagonal
    probabilities, in both noise semantics, each with the configuration the
    autotuner picks for it (restored afterwards).  Then
    the 12-qubit noisy hardware-efficient ansatz (depolarizing 0.005) with
    its full adjoint gradient through the checkpointed density-matrix
    adjoint."""
    import numpy as np

    from bsim import _lib as L
    from bsim import adjoint as A
    from bsim import workloads as W
    from bsim.simulators import DensityMatrixSimulator, lower_instructions

    ins = W.brickwall(nq, 10, seed=0, noise=True)
    words = ["I" * i + "ZZ" + "I" * (nq - 2 - i) for i in range(nq - 1)]
    out = {"workload": f"noisy brick-wall {nq} qubits depth 10 ({len(ins)} ops), "
                       f"{2 ** (2 * nq) * 16 / 2 ** 30:.0f} GiB density matrix, <Z_i Z_i+1> + probabilities"}
    keep = current_config()
    for mode in ("reference", "physical"):
        sim = DensityMatrixSimulator(nq, 0, noise_mode=mode)
        tuned = None
        if autotune:
            t0 = time.perf_counter()
            best = sim.autotune(ins)
            tuned = {k: best[k] for k in ("tile_bits", "low_bits", "sweep_bits", "fusion_limit", "merge_width")}
            tuned["search_s"] = round(time.perf_counter() - t0, 2)
        ol = lower_instructions(ins, L.BSIM_DM, mode)
        stats = (ctypes.c_int64 * 4)()
        L.call("bsim_plan_stats", 2 * nq, 1, ol.ptr, ol.n, stats)
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)  # warm-up
        L.call("bsim_synchronize", sim.handle)
        best_ev = float("inf")
        for _ in range(2):
            L.call("bsim_state_reset", sim.handle)
            L.call("bsim_synchronize", sim.handle)
            t = time.perf_counter()
            L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
            L.call("bsim_synchronize", sim.handle)
            best_ev = min(best_ev, time.perf_counter() - t)
        t = time.perf_counter()
        zz = sim.expect_pauli_words(words)
        p = sim.probabilities()
        t_res = time.perf_counter() - t
        out[mode] = {"evolve_ms": round(best_ev * 1e3, 3), "results_ms": round(t_res * 1e3, 3),
                     "ops_per_s": round(len(ins) / best_ev, 1),
                     "passes": int(stats[0]), "single_launches": int(stats[1]), "autotune": tuned,
                     "zz_checksum": float(np.sum(zz)), "trace": float(np.sum(p))}
        log(f"density matrix {nq}q {mode}: evolve {best_ev * 1e3:.1f} ms ({tuned})")
        del sim
        set_config(keep)
    n12 = 12
    hins, pmap, th = W.hea(n12, 8, seed=0, noise_p=0.005)
    hw, hc = W.tfim(n12)
    sim = DensityMatrixSimulator(n12, 0)
    A.adjoint_gradient(sim, hins, pmap, hw, hc)  # warm-up
    t = time.perf_counter()
    ev, g = A.adjoint_gradient(sim, hins, pmap, hw, hc)
    t_adj = time.perf_counter() - t
    out["adjoint"] = {"workload": f"noisy HEA {n12} qubits x 8 layers (depolarizing 0.005), {len(th)} params, "
                                  f"{len(hins)} ops, TFIM H, checkpoint every ceil(sqrt(ops))",
                      "s": round(t_adj, 4), "grads_per_s": round(1.0 / t_adj, 3), "expval": ev}
    log(f"density matrix adjoint {n12}q: {t_adj:.3f} s")
    return out


def per_gate_leg(n: int, ins, ceiling: dict | None):
    """The per-gate launch path (bsim_evolve mode 1, the reference's loop
    shape) on the workload's first layer: its 30 U gates (each one HBM pass,
    32 B x 2^n) and its CZ gates (a diagonal over the control-set half,
    16 B x 2^n), timed apart."""
    from bsim import _lib as L
    from bsim import gates as G
    from bsim.simulators import StateVectorSimulator, lower_instructions

    layer_u = [i for i in ins[: 2 * n] if isinstance(i.operator, G.U)][:n]
    layer_cz = [i for i in ins[n: 2 * n] if isinstance(i.operator, G.CZ)]
    sim = StateVectorSimulator(n, 0)
    out = {"workload": f"first brick-wall layer at {n} qubits, one launch per gate"}
    for name, part, unit_bytes in (("u", layer_u, 32.0 * 2 ** n), ("cz", layer_cz, 16.0 * 2 ** n)):
        ol = lower_instructions(part, L.BSIM_SV)
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 1)  # warm-up
        L.call("bsim_synchronize", sim.handle)
        t = time.perf_counter()
        for _ in range(2):
            L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 1)
        L.call("bsim_synchronize", sim.handle)
        per = (time.perf_counter() - t) / (2 * len(part))
        gbs = unit_bytes / per / 1e9
        out[name] = {"gates": len(part), "ms_per_gate": round(per * 1e3, 3), "pass_GBps": round(gbs, 1),
                     "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
        if ceiling:
            out[name]["frac_of_stream_copy"] = round(gbs / ceiling["copy_GBps"], 4)
        log(f"per-gate {name}: {per * 1e3:.2f} ms per gate, {gbs:.0f} GB/s")
    out["note"] = "pass_GBps: algorithmic bytes per gate (U: every amplitude read and written once; CZ: the control-set half)"
    del sim
    out["device_copy_GBps"] = copy_bandwidth()
    log(f"device-to-device copy: {out['device_copy_GBps']:.0f} GB/s")
    return out


class _DeviceCopy:
    """Two device buffers of ``nbytes`` and timed device-to-device copies
    between them, through the HIP runtime libbsim itself links (ctypes: no
    second GPU runtime in the process)."""

    def __init__(self, nbytes: int):
        import ctypes as C

        self.C = C
        self.hip = C.CDLL("libamdhip64.so")
        self.n = nbytes
        self.src, self.dst = C.c_void_p(), C.c_void_p()
        self._ok(self.hip.hipMalloc(C.byref(self.src), C.c_size_t(nbytes)), "hipMalloc")
        self._ok(self.hip.hipMalloc(C.byref(self.dst), C.c_size_t(nbytes)), "hipMalloc")
        self._ok(self.hip.hipMemset(self.src, 0, C.c_size_t(nbytes)), "hipMemset")
        self._ok(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def copy_seconds(self) -> float:
        C = self.C
        e0, e1 = C.c_void_p(), C.c_void_p()
        self._ok(self.hip.hipEventCreate(C.byref(e0)), "hipEventCreate")
        self._ok(self.hip.hipEventCreate(C.byref(e1)), "hipEventCreate")
        self._ok(self.hip.hipEventRecord(e0, None), "hipEventRecord")
        self._ok(self.hip.hipMemcpyAsync(self.dst, self.src, C.c_size_t(self.n), 3, None), "hipMemcpyAsync")
        self._ok(self.hip.hipEventRecord(e1, None), "hipEventRecord")
        self._ok(self.hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = C.c_float()
        self._ok(self.hip.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
        self.hip.hipEventDestroy(e0)
        self.hip.hipEventDestroy(e1)
        return ms.value * 1e-3

    def close(self):
        self.hip.hipFree(self.src)
        self.hip.hipFree(self.dst)


def copy_bandwidth() -> float:
    """Reference point for the per-gate figures: a plain 4 GiB
    device-to-device copy (read + write bytes / time), best of 3."""
    cp = _DeviceCopy(2 ** 32)
    cp.copy_seconds()
    best = min(cp.copy_seconds() for _ in range(3))
    cp.close()
    return round(2 * 2 ** 32 / best / 1e9, 1)


def graph_leg(reps: int):
    """Launch-bound small circuits: a 12-qubit brick-wall of depth 20 (460
    gates) evolved from |0...0> ``reps`` times, eagerly (bsim_evolve: lower,
    plan, upload, launch per call) and as replays of one captured graph
    (reset + passes).  Reports the wall time per evolution."""
    import numpy as np

    from bsim import _lib as L
    from bsim import workloads as W
    from bsim.simulators import StateVectorSimulator, lower_instructions

    nq, depth = 12, 20
    ins = W.brickwall(nq, depth, seed=5)
    ol = lower_instructions(ins, L.BSIM_SV)
    sim = StateVectorSimulator(nq, 0)
    for _ in range(10):
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
    L.call("bsim_synchronize", sim.handle)
    t = time.perf_counter()
    for _ in range(reps):
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
    L.call("bsim_synchronize", sim.handle)
    t_eager = (time.perf_counter() - t) / reps
    eager = sim.state_vector()
    g = sim.capture(ins, reset=True)
    for _ in range(10):
        g.launch()
    L.call("bsim_synchronize", sim.handle)
    t = time.perf_counter()
    for _ in range(reps):
        g.launch()
    L.call("bsim_synchronize", sim.handle)
    t_graph = (time.perf_counter() - t) / reps
    same = bool(np.array_equal(sim.state_vector(), eager))
    g.close()
    log(f"captured {nq}q depth {depth}: eager {t_eager * 1e6:.1f} us, graph {t_graph * 1e6:.1f} us per evolve")
    return {"workload": f"brick-wall {nq} qubits depth {depth} ({len(ins)} gates) from |0...0>, {reps} evolutions",
            "eager_us": round(t_eager * 1e6, 2), "graph_us": round(t_graph * 1e6, 2),
            "speedup": round(t_eager / t_graph, 2), "bit_identical": same}


def pmc_probe(n: int, depth: int):
    """Kernels for hardware-counter runs: the fused passes of one evolve and
    the stream probe's kernels over 4 GiB (k_probe_write fills the source:
    4 GiB written; k_probe_copy twice: 4 GiB read + 4 GiB written each),
    whose FETCH_SIZE / WRITE_SIZE calibrate the counters."""
    from bsim import _lib as L
    from bsim import workloads as W
    from bsim.simulators import StateVectorSimulator, lower_instructions

    ol = lower_instructions(W.brickwall(n, depth, seed=0), L.BSIM_SV)
    sim = StateVectorSimulator(n, 0)
    L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
    L.call("bsim_synchronize", sim.handle)
    log(f"pmc probe: evolve {n}q done")
    del sim
    ms = ctypes.c_double()
    L.call("bsim_stream_probe", 0, 2 ** 32, 0, 1, ctypes.byref(ms))
    log(f"pmc probe: stream copy of 4 GiB done ({ms.value:.3f} ms)")


def batch_leg(count: int, sweep_sets: int, shots: int = 1000):
    """The uneven batch of one persistent launch: ``count`` brick-wall
    circuits (qubits uniform in 8..22, depth uniform in 10..40, <Z_q> + 1000
    shots each) together with ``sweep_sets`` parameter sets of the 16-qubit
    8-layer ansatz, each with its adjoint gradient of the TFIM <H>.  Against
    the serial path on the same tasks: one reset + fused evolve + <Z> +
    alias table + shots launch sequence per circuit, then one
    adjoint_gradient call per parameter set.  Latency = completion time of
    each task from the start (persistent: the device clock from the launch
    start); throughput = tasks / wall time.  The circuits and the gradient
    sets are lowered to op lists once, outside both timed regions (timed
    apart: host_lowering_s); the planning and packing of the persistent
    launch are inside its time."""
    import numpy as np

    from bsim import _lib as L
    from bsim import adjoint as A
    from bsim import workloads as W
    from bsim.batch import GradientTask, run_batch
    from bsim.simulators import StateVectorSimulator, lower_instructions

    spec = W.batch_tasks(count, seed=0)
    circuits = [(n, W.brickwall_ops(n, d, s)) for n, d, s in spec]
    seeds = [s for _, _, s in spec]
    gates = sum(c[1].n for c in circuits)
    words, coeffs = W.tfim(16)
    sets = [W.hea(16, 8, seed=k) for k in range(sweep_sets)]
    # both paths take the same lowered op lists (as the circuits' packed ops)
    t = time.perf_counter()
    lowered = [lower_instructions(list(ins), L.BSIM_SV, param_map=pmap) for ins, pmap, _ in sets]
    for ol in lowered:
        ol.ptr  # an OpList writes its bsim_op records on first use: part of the lowering
    t_lower = time.perf_counter() - t
    grads = [GradientTask(16, ol, pmap, words, coeffs) for ol, (_, pmap, _) in zip(lowered, sets)]
    log(f"batch: {count} circuits ({gates} gates) + {sweep_sets} gradient sets in one launch")
    nw = min(4, sweep_sets)
    run_batch(circuits[:64] + grads[:nw], [shots] * 64 + [0] * nw, seeds[:64] + [0] * nw)  # warm-up
    t = time.perf_counter()
    res = run_batch(circuits + grads, [shots] * count + [0] * sweep_sets, seeds + [0] * sweep_sets)
    t_batch = time.perf_counter() - t
    lat_c = np.array([r.latency_us for r in res[:count]]) * 1e-3
    lat_g = np.array([r.latency_us for r in res[count:]]) * 1e-3
    log("batch: serial path")
    sims = {n: StateVectorSimulator(n, shots) for n in sorted({c[0] for c in circuits})}
    out = np.empty(shots, dtype=np.int64)
    done_c = np.empty(count)
    z_diff = 0.0
    t = time.perf_counter()
    for i, ((n, ops), seed) in enumerate(zip(circuits, seeds)):
        sim = sims[n]
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_evolve", sim.handle, ops.ptr, ops.n, 0)
        z = sim.expect_z_all()
        L.call("bsim_alias_build", sim.handle, L.ALIAS_PARALLEL, 0.0)
        L.call("bsim_sample", sim.handle, shots, ctypes.c_uint64(seed), L.i64ptr(out))
        done_c[i] = time.perf_counter() - t
        z_diff = max(z_diff, float(np.max(np.abs(z - res[i].expect_z))))
    done_g = np.empty(sweep_sets)
    g_diff = 0.0
    gsim = sims.get(16) or StateVectorSimulator(16, 0)
    for k, (ol, (_, pmap, _)) in enumerate(zip(lowered, sets)):
        ev, g = A.adjoint_gradient(gsim, ol, pmap, words, coeffs)
        done_g[k] = time.perf_counter() - t
        g_diff = max(g_diff, float(np.max(np.abs(g - res[count + k].grad))), abs(ev - res[count + k].expval))
    t_serial = time.perf_counter() - t
    del sims, gsim

    def pct(x):
        return {"p50": round(float(np.percentile(x, 50)), 3), "p99": round(float(np.percentile(x, 99)), 3)} \
            if len(x) else None

    ntask = count + sweep_sets
    log(f"batch: persistent {t_batch:.3f} s, serial {t_serial:.3f} s")
    return {
        "workload": f"{count} brick-wall SV circuits (qubits U{{8..22}}, depth U{{10..40}}, <Z_i> + {shots} shots "
                    f"each) + {sweep_sets} parameter sets of the 16-qubit 8-layer ansatz with adjoint gradients "
                    f"(TFIM H), one persistent launch",
        "gates": int(gates),
        "persistent_s": round(t_batch, 4),
        "persistent_tasks_per_s": round(ntask / t_batch, 1),
        "persistent_kernel_ms": round(float(max(lat_c.max(initial=0), lat_g.max(initial=0))), 3),
        "persistent_latency_ms": {"circuit": pct(lat_c), "gradient": pct(lat_g)},
        "serial_s": round(t_serial, 4),
        "serial_tasks_per_s": round(ntask / t_serial, 1),
        "serial_latency_ms": {"circuit": pct(done_c * 1e3), "gradient": pct(done_g * 1e3)},
        "speedup": round(t_serial / t_batch, 2),
        "host_lowering_s": round(t_lower, 4),
        "max_abs_z_diff_vs_serial": z_diff,
        "max_abs_grad_diff_vs_serial": g_diff,
        "latency_note": "completion time from the start; persistent = device clock from the launch start",
    }


if __name__ == "__main__":
    main()
