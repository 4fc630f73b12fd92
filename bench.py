"""Benchmark: 30-qubit ComplexF64 brick-wall random circuit (depth 20, U on
every qubit + nearest-neighbour CZ layers), <Z_i> for all i and 1,000,000
alias-method shots, on one MI355X.

One step = reinit to |0...0> + fused evolve + one-pass <Z_i> + alias table
build + 1e6 shots, with the state resident in HBM.  Prints one JSON line.

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


def cpu_baseline(n: int, ins, budget_s: float = 15.0):
    """The oracle (C restatement of the reference's per-gate loops, OpenMP
    over every host core) on the same 30-qubit circuit, on a bounded prefix
    of its gates."""
    import numpy as np

    import oracle as O
    from bsim.gates import kernel_call

    cores = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1
    t0 = time.perf_counter()
    done = 0
    bytes_ = 0.0
    for i in ins:
        call = kernel_call(i.operator, i.target)
        O.apply_call(psi, n, call)
        done += 1
        # reference traffic: dense gates stream the whole state, controlled
        # gates (CZ) the control-set half (src/gate_kernels.jl:579-606)
        bytes_ += (16.0 if call.controls else 32.0) * 2 ** n
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    del psi
    return {
        "value": bytes_ / dt / 1e9,
        "unit": "GB/s",
        "gates_per_s": done / dt,
        "cores": int(os.environ.get("OMP_NUM_THREADS", cores)),
        "kind": "port",
        "sample": f"first {done} of {len(ins)} gates of the {n}-qubit brick-wall, reference per-gate loops "
                  f"restated in C (oracle/refcpu.c), {dt:.1f} s; reference Julia unavailable",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=20)
    ap.add_argument("--shots", type=int, default=1000000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--adjoint-qubits", type=int, default=24, help="0 skips the adjoint leg")
    ap.add_argument("--adjoint-large", type=int, default=28,
                    help="qubits of the second adjoint workload (the same ansatz; 0 skips it)")
    ap.add_argument("--dense-k", type=int, default=28, help="qubits of the dense 3..6-target gate leg (0 skips it)")
    ap.add_argument("--ps-sample", type=int, default=24, help="parameters timed by parameter shift")
    ap.add_argument("--batch", type=int, default=4096, help="circuits in the batch leg (0 skips it)")
    ap.add_argument("--dm-qubits", type=int, default=14, help="noisy density-matrix leg (0 skips it)")
    ap.add_argument("--graph-reps", type=int, default=500, help="captured-evolution replays (0 skips the leg)")
    ap.add_argument("--per-gate", type=int, default=1, help="1: time the per-gate launch path on one layer")
    ap.add_argument("--sweep-sets", type=int, default=256, help="16-qubit ansatz parameter sets with gradients")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="profiling aid: one fused evolve of the workload and one 8 GiB device copy, no JSON")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(sharding is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import bsim
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
    stats = (ctypes.c_int64 * 4)()
    L.call("bsim_plan_stats", n, 0, ol.ptr, ol.n, stats)
    passes = int(stats[0])
    tile_bits, low_bits = ctypes.c_int(), ctypes.c_int()
    L.call("bsim_get_tile_config", ctypes.byref(tile_bits), ctypes.byref(low_bits))
    log(f"workload: {n}q brick-wall depth {depth}: {n_gates} gates -> {passes} fused passes "
        f"(tile {tile_bits.value} bits, {low_bits.value} contiguous)")

    sim = StateVectorSimulator(n, a.shots)
    lib = L.load()
    phases = {"evolve": 0.0, "expect_z": 0.0, "alias_build": 0.0, "shots": 0.0}

    def step(timed):
        t = time.perf_counter()
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
        t1 = time.perf_counter()
        z = sim.expect_z_all()
        t2 = time.perf_counter()
        L.call("bsim_alias_build", sim.handle, L.ALIAS_PARALLEL, 0.0)
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
    value = passes * pass_bytes / evolve_s / 1e9
    avg_kernel_ms = kms.value / max(launches.value, 1)
    achieved = pass_bytes / (avg_kernel_ms * 1e-3) / 1e9
    ms_per_step = dt / a.steps * 1e3
    rec = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
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
                        f"({2 ** n * 16 / 2 ** 30:.0f} GiB), <Z_i> all i + {a.shots} alias-method shots",
            "qubits": n, "depth": depth, "gates": n_gates, "shots": a.shots,
            "fused_passes": passes, "tile_bits": tile_bits.value, "low_bits": low_bits.value,
        },
        "gates_per_s": round(n_gates / evolve_s, 1),
        "phase_ms": {k: round(v / a.steps * 1e3, 3) for k, v in phases.items()},
        "shots_per_s": round(a.shots / (phases["shots"] / a.steps), 1),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_tile (fused LDS-tiled gate pass)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(n, depth),
            "traffic_source": "profiles/k_tile_brickwall_30q.md (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
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
    if a.per_gate:
        rec["per_gate"] = per_gate_leg(n, ins)
    if a.adjoint_qubits > 0:
        rec["adjoint"] = adjoint_leg(a.adjoint_qubits, a.ps_sample)
    if a.adjoint_large > 0:
        rec["adjoint_large"] = adjoint_large_leg(a.adjoint_large)
    if a.dense_k > 0:
        rec["dense_k"] = dense_k_leg(a.dense_k)
    if a.batch > 0:
        rec["batch"] = batch_leg(a.batch, a.sweep_sets)
    if a.dm_qubits > 0:
        rec["density_matrix"] = dm_leg(a.dm_qubits)
    if a.graph_reps > 0:
        rec["captured"] = graph_leg(a.graph_reps)
    if not a.no_cpu:
        log("cpu baseline (oracle) ...")
        rec["cpu_baseline"] = cpu_baseline(n, ins)
    print(json.dumps(rec), flush=True)


def adjoint_leg(nq: int, ps_sample: int):
    """24-qubit hardware-efficient ansatz (8 layers, 384 parameters) with the
    open-chain transverse-field Ising H: full adjoint gradient vs parameter
    shift through the plain GPU path on a sample of parameters."""
    import numpy as np

    from bsim import adjoint as A
    from bsim import workloads as W
    from bsim.simulators import StateVectorSimulator

    ins, pmap, th = W.hea(nq, 8, seed=0)
    words, coeffs = W.tfim(nq)
    sim = StateVectorSimulator(nq, 0)
    log(f"adjoint: {nq} qubits, {len(th)} parameters")
    A.adjoint_gradient(sim, ins, pmap, words, coeffs)  # warm-up
    reps = 3
    t = time.perf_counter()
    for _ in range(reps):
        ev, grad = A.adjoint_gradient(sim, ins, pmap, words, coeffs)
    t_adj = (time.perf_counter() - t) / reps
    # the same gradient with one backward kernel per op
    A.adjoint_gradient(sim, ins, pmap, words, coeffs, fused=False)
    t = time.perf_counter()
    for _ in range(reps):
        _, grad_op = A.adjoint_gradient(sim, ins, pmap, words, coeffs, fused=False)
    t_op = (time.perf_counter() - t) / reps
    log(f"adjoint: fused backward {t_adj * 1e3:.1f} ms, per-op backward {t_op * 1e3:.1f} ms")
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


def dm_leg(nq: int):
    """Noisy density matrix: the brick-wall of the state-vector leg at nq
    qubits, depth 10, with Depolarizing(0.01) after every U,
    TwoQubitDepolarizing(0.02) after every CZ and AmplitudeDamping(0.005) on
    every qubit per layer; <Z_i Z_i+1> for all pairs and the diagonal
    probabilities, in both noise semantics.  Then the 12-qubit noisy
    hardware-efficient ansatz (depolarizing 0.005) with its full adjoint
    gradient through the checkpointed density-matrix adjoint."""
    import numpy as np

    from bsim import _lib as L
    from bsim import adjoint as A
    from bsim import workloads as W
    from bsim.simulators import DensityMatrixSimulator, lower_instructions

    ins = W.brickwall(nq, 10, seed=0, noise=True)
    words = ["I" * i + "ZZ" + "I" * (nq - 2 - i) for i in range(nq - 1)]
    out = {"workload": f"noisy brick-wall {nq} qubits depth 10 ({len(ins)} ops), "
                       f"{2 ** (2 * nq) * 16 / 2 ** 30:.0f} GiB density matrix, <Z_i Z_i+1> + probabilities"}
    for mode in ("reference", "physical"):
        sim = DensityMatrixSimulator(nq, 0, noise_mode=mode)
        ol = lower_instructions(ins, L.BSIM_DM, mode)
        stats = (ctypes.c_int64 * 4)()
        L.call("bsim_plan_stats", 2 * nq, 1, ol.ptr, ol.n, stats)
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)  # warm-up
        L.call("bsim_synchronize", sim.handle)
        t = time.perf_counter()
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
        L.call("bsim_synchronize", sim.handle)
        t_ev = time.perf_counter() - t
        t = time.perf_counter()
        zz = sim.expect_pauli_words(words)
        p = sim.probabilities()
        t_res = time.perf_counter() - t
        out[mode] = {"evolve_ms": round(t_ev * 1e3, 3), "results_ms": round(t_res * 1e3, 3),
                     "passes": int(stats[0]), "single_launches": int(stats[1]),
                     "zz_checksum": float(np.sum(zz)), "trace": float(np.sum(p))}
        log(f"density matrix {nq}q {mode}: evolve {t_ev * 1e3:.1f} ms")
        del sim
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


def per_gate_leg(n: int, ins):
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
    """Kernels for hardware-counter runs: the fused passes of one evolve
    (k_tile) and a plain device copy of 2^29 doubles (4 GiB read + 4 GiB
    written), whose FETCH_SIZE / WRITE_SIZE calibrate the counters."""
    from bsim import _lib as L
    from bsim import workloads as W
    from bsim.simulators import StateVectorSimulator, lower_instructions

    ol = lower_instructions(W.brickwall(n, depth, seed=0), L.BSIM_SV)
    sim = StateVectorSimulator(n, 0)
    L.call("bsim_evolve", sim.handle, ol.ptr, ol.n, 0)
    L.call("bsim_synchronize", sim.handle)
    log(f"pmc probe: evolve {n}q done")
    del sim
    cp = _DeviceCopy(2 ** 32)
    cp.copy_seconds()
    cp.close()
    log("pmc probe: copy of 4 GiB done")


def batch_leg(count: int, sweep_sets: int, shots: int = 1000):
    """Uneven batch: ``count`` brick-wall circuits, qubits uniform in 8..22,
    depth uniform in 10..40, <Z_q> + 1000 shots each, in one persistent
    launch, against the per-circuit launch path on the same circuits; then
    ``sweep_sets`` parameter sets of the 16-qubit 8-layer ansatz, each with
    its adjoint gradient."""
    import numpy as np

    from bsim import _lib as L
    from bsim import adjoint as A
    from bsim import workloads as W
    from bsim.batch import run_batch
    from bsim.simulators import StateVectorSimulator

    spec = W.batch_tasks(count, seed=0)
    circuits = [(n, W.brickwall_ops(n, d, s)) for n, d, s in spec]
    seeds = [s for _, _, s in spec]
    gates = sum(c[1].n for c in circuits)
    log(f"batch: {count} circuits, {gates} gates")
    run_batch(circuits[:64], shots, seeds[:64])  # warm-up
    t = time.perf_counter()
    res = run_batch(circuits, shots, seeds)
    t_batch = time.perf_counter() - t
    lat = np.array([r.latency_us for r in res]) * 1e-3
    # per-circuit launches: reset, fused evolve, <Z>, parallel alias, shots
    log("batch: per-circuit launch path")
    sims = {n: StateVectorSimulator(n, shots) for n in sorted({c[0] for c in circuits})}
    out = np.empty(shots, dtype=np.int64)
    done = np.empty(count)
    z_diff = 0.0
    t = time.perf_counter()
    for i, ((n, ops), seed) in enumerate(zip(circuits, seeds)):
        sim = sims[n]
        L.call("bsim_state_reset", sim.handle)
        L.call("bsim_evolve", sim.handle, ops.ptr, ops.n, 0)
        z = sim.expect_z_all()
        L.call("bsim_alias_build", sim.handle, L.ALIAS_PARALLEL, 0.0)
        L.call("bsim_sample", sim.handle, shots, ctypes.c_uint64(seed), L.i64ptr(out))
        done[i] = time.perf_counter() - t
        z_diff = max(z_diff, float(np.max(np.abs(z - res[i].expect_z))))
    t_single = done[-1]
    del sims
    rec = {
        "workload": f"{count} brick-wall SV circuits, qubits U{{8..22}}, depth U{{10..40}}, <Z_i> + {shots} shots each",
        "gates": int(gates),
        "persistent_s": round(t_batch, 4),
        "persistent_circuits_per_s": round(count / t_batch, 1),
        "persistent_kernel_ms": round(float(lat.max()), 3),
        "persistent_latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
        "persistent_latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
        "per_circuit_s": round(t_single, 4),
        "per_circuit_circuits_per_s": round(count / t_single, 1),
        "per_circuit_latency_ms_p50": round(float(np.percentile(done, 50)) * 1e3, 3),
        "per_circuit_latency_ms_p99": round(float(np.percentile(done, 99)) * 1e3, 3),
        "speedup": round(t_single / t_batch, 2),
        "max_abs_z_diff_vs_per_circuit": z_diff,
        "latency_note": "completion time from the start of the batch; persistent = device clock from the launch",
    }
    if sweep_sets > 0:
        words, coeffs = W.tfim(16)
        sim = StateVectorSimulator(16, 0)
        sets = [W.hea(16, 8, seed=k) for k in range(sweep_sets)]
        A.adjoint_gradient(sim, sets[0][0], sets[0][1], words, coeffs)
        log(f"batch: {sweep_sets} parameter sets x 16-qubit ansatz gradients")
        t = time.perf_counter()
        for ins, pmap, _ in sets:
            A.adjoint_gradient(sim, ins, pmap, words, coeffs)
        t_sw = time.perf_counter() - t
        rec["sweep"] = {
            "workload": f"{sweep_sets} parameter sets, HEA 16 qubits x 8 layers ({len(sets[0][2])} params), TFIM H",
            "s": round(t_sw, 4),
            "grads_per_s": round(sweep_sets / t_sw, 1),
        }
    return rec


if __name__ == "__main__":
    main()
