"""Benchmark: 30-qubit complex64 brickwork gate application on one MI355X.

Workload (one step): depth-24 brickwork of Haar-random SU(4) gates on 30
qubits (348 gates, seed 0), applied to a resident 8 GiB complex64 state.
The circuit is compiled once in the launch shape the program autotuner
measures fastest on this GPU (fusion width 2-5, 3-5-qubit blocks on MFMA or
VALU, tile of 2^11-2^13 amplitudes; cache-blocked stages with the tile in LDS,
in registers between LDS round trips, or one whole tile per wave in
registers); a step is one replay of the compiled program.

Printed JSON line: gates/s as `value`; the whole-job HBM rate; the roofline
of the dominant kernel (algorithmic FLOPs or bytes per launch / its average
launch time, from HIP events recorded between stages on the launch stream);
the CPU baseline (the oracle's C + OpenMP restatement applying the same gates
at 30 qubits on every host core, a bounded sample of gates); and extra workloads
(per-gate streaming passes, full-size c64-vs-c128 parity, 32 and 34 qubits, the
20-qubit plain circuit, VQE adjoint vs parameter shift, the 1024-instance
batch, the 28-qubit XXZ precision study, Lanczos ground states of the XXZ
chain).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dense-evolution_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_info():
    """CPU model, cores this process may use, and RAM of the host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    try:
        ram = os.sysconf("SC_PAGE_SIZE") * os.sysconf("SC_PHYS_PAGES") / 2 ** 30
    except (ValueError, OSError):
        ram = float("nan")
    return {"cpu_model": model, "ram_gib": round(ram, 1), "cores_available": len(os.sched_getaffinity(0))}


def median_of_runs(step, units, budget_s, runs=5):
    """The host-baseline protocol: one warm-up run, then the median wall time
    of `runs` timed runs.  A run is a bounded sample of k consecutive units
    (gates, evaluations) of the workload, k fixed by the warm-up run, which
    takes as many units as fit in budget_s / (runs + 1); step(i) processes
    unit i (taken modulo `units`).  Returns units per second of the median
    run, k and the run times."""
    t0 = time.perf_counter()
    k = 0
    while k < units and (k == 0 or time.perf_counter() - t0 < budget_s / (runs + 1)):
        step(k)
        k += 1
    times, at = [], k
    for _ in range(runs):
        t1 = time.perf_counter()
        for u in range(k):
            step((at + u) % units)
        times.append(time.perf_counter() - t1)
        at += k
    med = sorted(times)[len(times) // 2]
    return k / med, k, times


def cpu_baseline(n, gates, budget_s=20.0):
    """The oracle's C + OpenMP restatement (oracle/c) applying the same SU(4)
    gates to a complex64 2^n state on the host, on every core this process may
    use: one warm-up run and the median of 5 runs, each run a bounded sample
    of consecutive gates of the circuit."""
    import numpy as np

    from oracle import sv_oracle_c as OC
    threads = OC.default_threads()
    sv = np.zeros(1 << n, dtype=np.complex64)
    sv[0] = 1
    OC.apply_matrix(sv, np.eye(4), [0, 1], threads=threads)   # page the state in

    def step(i):
        u, a, b = gates[i]
        OC.apply_matrix(sv, u, [a, b], threads=threads)
    rate, k, times = median_of_runs(step, len(gates), budget_s)
    return {"value": rate, "unit": "gates/s", "cores": threads, "kind": "port",
            "sample": f"one warm-up run, then the median of 5 runs of {k} consecutive SU(4) gates "
                      f"(of the {len(gates)}) at {n} qubits, complex64, oracle/c (C + OpenMP, one pass "
                      f"over the state per gate); run times {[round(t, 2) for t in times]} s",
            **host_info()}


def cpu_baselines_small(budget_s=8.0):
    """Host baselines of the other workloads on the oracle, each by the same
    protocol (one warm-up run, median of 5 runs of a bounded sample): the C +
    OpenMP restatement (all host cores) and the NumPy per-gate gather
    formulation of the reference's kernel (one thread of NumPy), complex64 for
    gates, complex128 for the VQE energy."""
    import numpy as np
    from desv import workloads
    from oracle import sv_oracle as O
    from oracle import sv_oracle_c as OC
    threads = OC.default_threads()

    out = {"cores": threads}
    # plain 20-qubit circuit (1580 gates)
    n = 20
    ops = O.encode_ops(workloads.random_layers_circuit(n, 20, seed=0), n)
    sv = np.zeros(1 << n, dtype=np.complex64)
    sv[0] = 1
    out["plain20_c_gates_per_s"] = median_of_runs(
        lambda i: OC.run_ops(sv, ops[i][None, :], threads), len(ops), budget_s)[0]
    box = [O.zero_state(n, np.complex64)]
    out["plain20_numpy_gather_gates_per_s"] = median_of_runs(
        lambda i: box.__setitem__(0, O.apply_op_gather(box[0], ops[i])), len(ops), budget_s / 2)[0]
    # 20-qubit XXZ Trotter (200 steps)
    circ = workloads.xxz_trotter_circuit(n, 0.05, 200, seed=0)
    xops = O.encode_ops(circ, n)
    sv = np.zeros(1 << n, dtype=np.complex64)
    sv[0] = 1
    out["xxz20_c_gates_per_s"] = median_of_runs(
        lambda i: OC.run_ops(sv, xops[i][None, :], threads), len(xops), budget_s)[0]
    out["xxz20_gates"] = len(xops)
    # 20-qubit VQE energy (forward on the C oracle in complex128, energy with
    # the NumPy Pauli sum), one evaluation a unit; a parameter-shift gradient
    # costs 2 x 320 of them
    vc, theta, terms = workloads.vqe_instance(n, 8, 200, seed=0)
    tmpl = O.build_template(vc.ops, n)
    table = O.patch_template(tmpl, theta)
    energy = [0.0]

    def vqe_eval(_):
        st = np.zeros(1 << n, dtype=np.complex128)
        st[0] = 1
        OC.run_ops(st, table, threads)
        energy[0] = O.pauli_sum_expectation(st, terms, n)
    rate = median_of_runs(vqe_eval, 1, budget_s)[0]
    out["vqe20_energy_evals_per_s"] = rate
    out["vqe20_param_shift_grads_per_s"] = rate / (2 * len(theta))
    out["vqe20_energy"] = float(energy[0])
    out["sample"] = (f"each figure: one warm-up run, then the median of 5 runs of consecutive gates "
                     f"(about {budget_s:.0f} s in all); the VQE figure times whole energy evaluations, "
                     "the gradient figure is that rate / 640")
    return out


def _timed(fn, reps=1):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, out


def suite_plain20():
    """20-qubit random rx/ry/rz + cx-ladder circuit (1580 gates), complex64 vs
    the complex128 run of the same engine: gates/s, fidelity, max |d<Z_i>|."""
    import numpy as np
    import desv
    from desv import workloads
    n = 20
    circ = workloads.random_layers_circuit(n, 20, seed=0)
    s64 = desv.DenseSVSimulator(n, use_float32=True)
    ops = s64.compile_ops(circ)
    prog = desv.Program(n).add_ops(ops)
    tuned = desv.tune.autotune_program(prog, "complex64", reps=5, cache=False)["best"]
    prog.run(s64.handle)
    dt, _ = _timed(lambda: (s64.handle.reset(), prog.run(s64.handle)), reps=20)
    s128 = desv.DenseSVSimulator(n)
    s128.run_circuit(circ)
    a, b = s64.get_statevector(), s128.get_statevector()
    a = a.astype(np.complex128)
    fid = float(abs(np.vdot(a, b)) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real))
    dz = max(abs(desv.pauli_expectation(s64.sv, {q: 'Z'}) - desv.pauli_expectation(s128.sv, {q: 'Z'}))
             for q in range(n))
    return {"workload": "20q random rx/ry/rz + cx ladder, 20 layers, 1580 gates, c64",
            "gates_per_s": len(ops) / dt, "ms_per_circuit": dt * 1e3, "hbm_passes": prog.stats()["passes"],
            "autotuned": tuned, "fidelity_vs_c128": fid, "max_abs_dZ_vs_c128": dz}


def suite_batchjit20(n=20, rows=256):
    """run_batch_jit on the 20-qubit plain circuit (20 layers of rx, ry, rz on
    every qubit + a CX ladder, 1200 angle columns), 256 parameter rows in one
    device batch, complex64 (compensated, the simulator default): rows/s
    including the host encode, the batch build and the (256, 2^20) download,
    and the device-only forward time of a second run."""
    import numpy as np
    import desv
    from desv import workloads
    circ = workloads.random_layers_circuit(n, 20, seed=0)
    sim = desv.DenseSVSimulator(n, use_float32=True)
    cols = sum(1 for c in circ if c[0] in ('rx', 'ry', 'rz'))
    P = np.random.default_rng(1).uniform(0, 2 * np.pi, (rows, cols))
    sim.run_batch_jit(circ, P[:2])                       # warm up (library load, first launch)
    t, out = _timed(lambda: sim.run_batch_jit(circ, P), reps=1)
    # device-only: the same batch object run again
    from desv.batch import Batch
    from desv.simulator import encode_op_table
    from desv.transpiler import QuantumTranspiler
    tmpl = encode_op_table(QuantumTranspiler.transpile(circ), n, param_slots=True)
    slots = np.nonzero(tmpl[:, 3] == -1.0)[0]
    b = Batch("complex64")
    b.set_forward_only(True)
    b.configure(use_graph=False)
    pm = np.full(len(tmpl), -1, dtype=np.int64)
    for r in P:
        ops = tmpl.copy()
        ops[slots, 3] = r
        b.add_ops(n, ops, pm, 0, np.zeros(0, np.uint64), np.zeros(0, np.uint64), np.zeros(0))
    b.run_forward()
    td, _ = _timed(lambda: b.run_forward(), reps=3)
    b.close()
    return {"workload": f"run_batch_jit, {n}q plain circuit ({len(circ)} gates, {cols} columns), {rows} rows, c64",
            "rows_per_s": rows / t, "s_per_call": t, "device_rows_per_s": rows / td,
            "device_ms_per_run": td * 1e3, "gates_per_s_device": rows * len(circ) / td}


def suite_vqe20():
    """20-qubit HEA VQE (8 layers, 320 params, 200 Pauli terms): adjoint energy+gradient
    evals/s (c64 with FP64 reductions, and c128) vs parameter shift on the plain forward
    path, and the c64 error against the c128 result."""
    import numpy as np
    import desv
    from desv import workloads
    n = 20
    circ, theta, terms = workloads.vqe_instance(n, 8, 200, seed=0)
    H = desv.PauliSumOperator(terms, n)
    f128, _ = desv.circuit_to_energy_fn(circ, n, dtype="complex128")
    f64, _ = desv.circuit_to_energy_fn(circ, n, dtype="complex64")
    tuned = {name: f.autotune(theta, H, cache=False, reps=10)["best"] for name, f in (("c64", f64), ("c128", f128))}
    log(f"vqe20 autotune: {tuned}")
    f128.value_and_grad(theta, H)
    f64.value_and_grad(theta, H)
    t128, (e128, g128) = _timed(lambda: f128.value_and_grad(theta, H), reps=5)
    t64, (e64, g64) = _timed(lambda: f64.value_and_grad(theta, H), reps=5)
    tps, gps = _timed(lambda: desv.parameter_shift_grad(f64.program, theta, H), reps=1)
    return {"workload": "20q HEA 8 layers (472 gates, 320 params), 200 random Pauli terms",
            "adjoint_evals_per_s_c64": 1.0 / t64, "adjoint_evals_per_s_c128": 1.0 / t128,
            "autotuned": tuned,
            "param_shift_evals_per_s_c64": 1.0 / tps,
            "adjoint_over_param_shift_c64": tps / t64,
            "energy_c128": e128, "abs_dE_c64_vs_c128": abs(e64 - e128),
            "max_abs_dgrad_c64_vs_c128": float(np.max(np.abs(g64 - g128))),
            "max_abs_dgrad_param_shift_vs_adjoint_c64": float(np.max(np.abs(gps - g64)))}


def suite_batch(count=1024, solo_sample=128):
    """1024 heterogeneous VQE instances (n ~ U{10..22}, 4-8 layers, 50 terms):
    instances/s of energy + adjoint gradient in one graph-captured batch
    (complex64 with compensated matrices, the default, and plain), the
    per-instance latency inside the batch (device wall-clock completion stamp
    of each instance's last result, p50 / p99), the error against
    single-instance complex128 runs on a sample, and, run alone (its own
    compiled program, energy + gradient), the latency distribution over a
    sample of the instances."""
    import numpy as np
    import desv
    from desv import workloads
    from desv.batch import Batch
    insts = workloads.batch_instances(count=count)
    res = {}
    for label, comp in (("compensated", True), ("plain", False)):
        b = Batch("complex64", compensated=comp)
        for n, circ, theta, terms in insts:
            b.add(circ, n, theta, terms)
        tuned = b.autotune(reps=1, cache=False)   # launch shape measured on this device
        log(f"batch ({label}) autotune: best {tuned['best']} "
            f"{[(c['tile_bits'], c['fuse_bits'], c['use_mfma'], c['ms'] and round(c['ms'], 1)) for c in tuned['candidates']]}")
        b.run()
        dt, (es, gs) = _timed(lambda: b.run(), reps=2)
        done_ms = b.completion_us() / 1e3
        b.close()
        res[label] = (dt, tuned, es, gs, done_ms)
    dt, tuned, es, gs, done_ms = res["compensated"]
    # error of the batch against single-instance complex128 runs (a sample)
    errs = []
    gs_plain = res["plain"][3]
    for i in range(0, count, max(1, count // 64)):
        n, circ, theta, terms = insts[i]
        fn, _ = desv.circuit_to_energy_fn(circ, n, dtype="complex128")
        e1, g1 = fn.value_and_grad(theta, desv.PauliSumOperator(terms, n))
        fn.program.close()
        sc = max(1.0, float(np.max(np.abs(g1))))
        errs.append((abs(es[i] - e1) / max(1.0, abs(e1)), float(np.max(np.abs(gs[i] - g1))) / sc,
                     float(np.max(np.abs(gs_plain[i] - g1))) / sc))
    errs = np.array(errs)
    lat = []
    for i in range(0, count, max(1, count // solo_sample)):
        n, circ, theta, terms = insts[i]
        fn, _ = desv.circuit_to_energy_fn(circ, n, dtype="complex64")
        H = desv.PauliSumOperator(terms, n)
        fn.value_and_grad(theta, H)
        t, _ = _timed(lambda: fn.value_and_grad(theta, H), reps=3)
        lat.append(t * 1e3)
    lat = np.array(lat)
    return {"workload": f"{count} VQE instances, n ~ U{{10..22}}, 4-8 HEA layers, 50 terms, c64",
            "instances_per_s": count / dt, "ms_per_batch": dt * 1e3,
            "autotuned": tuned["best"],
            "autotune_candidates_ms": {f"t{c['tile_bits']}f{c['fuse_bits']}{'m' if c['use_mfma'] else ''}": c["ms"]
                                       for c in tuned["candidates"]},
            "instances_per_s_plain": count / res["plain"][0], "autotuned_plain": res["plain"][1]["best"],
            "max_rel_dE_vs_c128_sample": float(errs[:, 0].max()),
            "max_rel_dgrad_vs_c128_sample": float(errs[:, 1].max()),
            "max_rel_dgrad_vs_c128_sample_plain": float(errs[:, 2].max()),
            "error_sample": f"{len(errs)} instances (every {max(1, count // 64)}th) vs their single-instance "
                            "complex128 runs; relative to max(1, max |ref|)",
            "completion_ms_p50": float(np.percentile(done_ms, 50)),
            "completion_ms_p99": float(np.percentile(done_ms, 99)),
            "completion_ms_min": float(done_ms.min()), "completion_ms_max": float(done_ms.max()),
            "completion_note": "device wall-clock time from the batch's first kernel to each instance's last "
                               "reduced result (compensated run)",
            "amortized_ms_per_instance": dt * 1e3 / count,
            "solo_latency_ms_p50": float(np.percentile(lat, 50)),
            "solo_latency_ms_p99": float(np.percentile(lat, 99)),
            "solo_sample": len(lat),
            "solo_instances_per_s": len(lat) / (lat.sum() / 1e3),
            "batch_over_solo": (count / dt) / (len(lat) / (lat.sum() / 1e3))}


def suite_precision(n=28, steps=200):
    """XXZ chain (J=1, Delta=0.5) first-order Trotter, dt=0.05, 200 steps, from a
    random product state: complex64 with compensated matrices (the simulator's
    default) and with plain FP32-rounded matrices, each against complex128 on
    the GPU (FP64 norm, <Z_i Z_i+1> and fidelity |<c128|c64>|^2)."""
    import numpy as np
    import desv
    from desv import workloads
    circ = workloads.xxz_trotter_circuit(n, 0.05, steps, seed=0)
    # the launch shape is tuned on the first 20 steps (same preparation, same
    # per-step structure), then the whole evolution runs with it
    head = workloads.xxz_trotter_circuit(n, 0.05, min(steps, 20), seed=0)
    from desv.simulator import encode_op_table
    from desv.transpiler import QuantumTranspiler
    ops = encode_op_table(QuantumTranspiler.transpile(circ), n)
    hops = encode_op_table(QuantumTranspiler.transpile(head), n)
    bonds = [(1.0, {i: 'Z', i + 1: 'Z'}) for i in range(n - 1)]
    res, tuned = {}, {}
    ref = None
    for name, dtype, comp in (("c128", "complex128", False), ("c64_compensated", "complex64", True),
                              ("c64_plain", "complex64", False)):
        probe = desv.Program(n).add_ops(hops)
        tuned[name] = desv.tune.autotune_program(probe, dtype, reps=1, cache=False, compensated=comp)["best"]
        probe.close()
        prog = desv.Program(n).add_ops(ops).compile(dtype, compensated=comp, **tuned[name])
        h = desv.StateHandle(n, complex128=dtype == "complex128")
        t, _ = _timed(lambda: prog.run(h), reps=1)
        zzs = np.array([desv.pauli_expectation(desv.DeviceState(h), t_[1]) for t_ in bonds])
        r = {"gates_per_s": len(ops) / t, "hbm_passes": prog.stats()["passes"], "norm_minus_1": h.norm2() - 1.0}
        prog.close()
        if ref is None:
            ref, zref = h, zzs
        else:
            up = desv.StateHandle(n, complex128=True)
            up.copy_from(h)
            h.close()
            r["fidelity_vs_c128"] = abs(up.inner(ref)) ** 2
            r["infidelity_vs_c128"] = 1.0 - r["fidelity_vs_c128"]
            r["max_abs_dZZ_vs_c128"] = float(np.max(np.abs(zzs - zref)))
            up.close()
        res[name] = r
    ref.close()
    return {"workload": f"{n}q XXZ Trotter, {steps} steps, {len(ops)} gates", "autotuned": tuned,
            **{f"{k}": v for k, v in res.items()},
            "compensation_cost_time_ratio": res["c64_plain"]["gates_per_s"] / res["c64_compensated"]["gates_per_s"]}


MFMA_FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix peak
VALU_FP32_TFLOPS = 78.6          # FP32 VALU without packed-FP32 instructions


def _load_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 PMC summary (FETCH_SIZE and WRITE_SIZE, copy-kernel scaled)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    for e in table.get("entries", []):
        if e.get("kernel") == kernel and e.get("workload") == workload:
            return e.get("hbm_bytes_per_launch")
    return None


def suite_streaming30(n=30, depth=24, steps=2):
    """The same brickwork with one streaming pass per gate (no fusion, no cache
    blocking): the HBM-bound gate-application path, in GB/s and fraction of
    the 8 TB/s peak."""
    import desv
    from desv import workloads
    gates = workloads.brickwork_su4(n, depth, seed=0)
    prog = desv.Program(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    prog.compile("complex64", fuse_bits=1, per_gate=True)
    h = desv.StateHandle(n, complex128=False)
    prog.run(h)
    dt, _ = _timed(lambda: prog.run(h), reps=steps)
    h.close()
    passes = prog.stats()["passes"]
    gbps = passes * 2 * (1 << n) * 8 / dt / 1e9
    return {"workload": f"brickwork {n}q depth {depth}, one streaming pass per gate (k_direct), c64",
            "gates_per_s": len(gates) / dt, "passes_per_s": passes / dt, "hbm_gbps": gbps,
            "hbm_frac": gbps / HBM_PEAK_GBPS}


def suite_parity30(n=30, depth=24, shape=None):
    """Full-size parity: the benchmark circuit in complex64, compiled in the
    launch shape the headline was timed with (`shape`), against complex128;
    fidelity and norms computed on the GPU."""
    import desv
    from desv import workloads
    gates = workloads.brickwork_su4(n, depth, seed=0)
    prog = desv.Program(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    shape = dict(shape or dict(fuse_bits=4, use_mfma=True))
    prog.compile("complex64", **shape)
    plan_stats = prog.stats()
    h64 = desv.StateHandle(n, complex128=False)
    prog.run(h64)
    n64 = h64.norm2()
    prog.compile("complex128")
    h128 = desv.StateHandle(n, complex128=True)
    prog.run(h128)
    n128 = h128.norm2()
    up = desv.StateHandle(n, complex128=True)
    up.copy_from(h64)
    h64.close()
    raw = abs(up.inner(h128)) ** 2
    fid = raw / (n64 * n128)
    up.close()
    h128.close()
    return {"workload": f"brickwork {n}q depth {depth}: c64 in the timed launch shape vs c128",
            "shape": shape, "plan": plan_stats,
            "fidelity_c64_vs_c128": raw, "infidelity": 1 - raw, "fidelity_normalised": fid,
            "norm_c64_minus_1": n64 - 1, "norm_c128_minus_1": n128 - 1}


def suite_brickwork32(n=32, depth=24, shape=None):
    """The same circuit structure at 32 qubits (32 GiB complex64 state), in the
    headline's launch shape."""
    import desv
    from desv import workloads
    gates = workloads.brickwork_su4(n, depth, seed=0)
    prog = desv.Program(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    prog.compile("complex64", **dict(shape or dict(fuse_bits=4, use_mfma=True)))
    h = desv.StateHandle(n, complex128=False)
    prog.run(h)
    h.reset()
    dt, _ = _timed(lambda: prog.run(h), reps=1)
    nrm = h.norm2()
    h.close()
    st = prog.stats()
    return {"workload": f"brickwork {n}q depth {depth}, {len(gates)} gates, c64 (32 GiB state)",
            "gates_per_s": len(gates) / dt, "ms_per_circuit": dt * 1e3, "hbm_passes": st["passes"],
            "hbm_gbps": st["passes"] * 2 * (1 << n) * 8 / dt / 1e9, "norm_minus_1": nrm - 1}


def cpu_baseline_lanczos(terms, n, ncv):
    """The reference's ground-state pattern on the host: scipy eigsh (ARPACK)
    over a LinearOperator whose matvec is the oracle's NumPy Pauli-sum matvec.
    Returns (seconds, E0)."""
    import numpy as np
    import scipy.sparse.linalg as sla
    from oracle import sv_oracle as O
    dim = 1 << n
    op = sla.LinearOperator((dim, dim), matvec=lambda v: O.pauli_sum_matvec(np.ravel(v), terms),
                            dtype=np.complex128)
    t0 = time.perf_counter()
    e = sla.eigsh(op, k=1, which="SA", ncv=ncv, return_eigenvectors=False)[0]
    return time.perf_counter() - t0, float(e)


def suite_lanczos(n_big=26, n_cmp=16, ncv=20):
    """Lanczos ground state of the open XXZ chain (J = 1, delta = 0.5),
    complex128.  At `n_cmp` qubits: the device-resident solver against the
    reference's pattern, scipy eigsh over a LinearOperator (ARPACK, one host
    round trip per product), with the GPU matvec and with the oracle's NumPy
    matvec on the host.  At `n_big` qubits: time to solution and the rates of
    the three kernels of one Lanczos step (Pauli-sum matvec, 20-vector inner
    products, 21-vector combination), as bytes streamed per second."""
    import numpy as np
    import scipy.sparse.linalg as sla
    import desv
    from desv import lanczos, workloads
    out = {"workload": f"open XXZ chain (J=1, delta=0.5), ground state, complex128, ncv={ncv}"}
    terms = workloads.xxz_terms(n_cmp)
    dim = 1 << n_cmp
    lanczos.ground_state_energy(terms, n_cmp, ncv=ncv)
    info = {}
    t_dev, e_dev = _timed(lambda: lanczos.ground_state_energy(terms, n_cmp, ncv=ncv, info=info))
    op = desv.PauliSumOperator(terms, n_cmp).as_linear_operator()
    t_op, e_op = _timed(lambda: sla.eigsh(op, k=1, which="SA", ncv=ncv, return_eigenvectors=False)[0])
    t_cpu, e_cpu = cpu_baseline_lanczos(terms, n_cmp, ncv)
    out[f"{n_cmp}q"] = {"device_lanczos_s": t_dev, "device_matvecs": info["matvecs"],
                        "scipy_eigsh_gpu_matvec_s": t_op, "scipy_eigsh_numpy_matvec_s": t_cpu,
                        "speedup_vs_scipy_numpy": t_cpu / t_dev, "speedup_vs_scipy_gpu_matvec": t_op / t_dev,
                        "E0": e_dev, "abs_dE_vs_scipy_numpy": abs(e_dev - e_cpu),
                        "abs_dE_vs_scipy_gpu": abs(e_dev - e_op)}
    n = n_big
    terms = workloads.xxz_terms(n)
    info = {}
    t_big, e_big = _timed(lambda: lanczos.ground_state_energy(terms, n, ncv=ncv, info=info))
    res = {"time_to_solution_s": t_big, "E0": e_big, "E0_per_bond": e_big / (n - 1), **info}
    # kernel rates at n_big
    hs = [desv.StateHandle(n, True) for _ in range(ncv + 2)]
    for i, h in enumerate(hs):
        desv._capi.check(desv._capi.lib.desv_fill_random(h.ptr, i + 1))
    xm, zm, cs = desv.PauliSumOperator(terms, n).encoded()
    groups = len(set(int(x) for x in xm))
    nb = (1 << n) * 16
    from desv.observables import matvec_into
    matvec_into(hs[1], hs[0], xm, zm, cs)
    t_mv, _ = _timed(lambda: matvec_into(hs[1], hs[0], xm, zm, cs), reps=5)
    t_in, _ = _timed(lambda: lanczos.inner_many(hs[-1], hs[:ncv]), reps=5)
    coef = np.concatenate([[1.0], -np.full(ncv, 1e-3)])
    t_cb, _ = _timed(lambda: lanczos.combine([hs[-1]], [hs[-1]] + hs[:ncv], coef), reps=5)
    for h in hs:
        h.close()
    rows = (ncv + 7) // 8
    res["kernels"] = {
        "pauli_matvec": {"ms": t_mv * 1e3, "flip_mask_groups": groups,
                         "stream_gbps": (groups + 1) * nb / t_mv / 1e9,
                         "min_traffic_gbps": 2 * nb / t_mv / 1e9},
        "inner_many": {"ms": t_in * 1e3, "vectors": ncv, "stream_gbps": (ncv + rows) * nb / t_in / 1e9},
        "combine": {"ms": t_cb * 1e3, "inputs": ncv + 1, "stream_gbps": (ncv + 2) * nb / t_cb / 1e9},
    }
    out[f"{n}q"] = res
    return out


def suite_readout(n=30, shots=1 << 20):
    """Device readout and noise on a 30-qubit complex64 state: shot sampling
    (fixed-point chunk totals of |a|^2 in one pass, then the chunks that hold
    shots, each shot a binary search in its chunk's prefix), a
    projective single-qubit measurement (probabilities, collapse,
    renormalisation) and one noise trajectory per model (every qubit draws;
    Pauli channels cost one streaming pass per fired qubit, amplitude damping
    a probability pass and an update pass per qubit)."""
    import numpy as np
    import desv
    from desv import _capi, workloads
    sim = desv.DenseSVSimulator(n, use_float32=True)
    sim.run_circuit(workloads.random_layers_circuit(n, 2, seed=3))
    h = sim.handle
    rng = np.random.default_rng(0)
    u = np.ascontiguousarray(rng.random(shots), dtype=np.float64)
    out = np.empty(shots, dtype=np.int64)
    sample = lambda: _capi.check(_capi.lib.desv_sample(h.ptr, _capi.dptr(u), shots, _capi.i64ptr(out)))
    sample()
    t_sample, _ = _timed(sample, reps=3)
    # one shot: the chunk-total pass over the state and one chunk
    one = np.ascontiguousarray(u[:1])
    o1 = np.empty(1, dtype=np.int64)
    t_one, _ = _timed(lambda: _capi.check(_capi.lib.desv_sample(h.ptr, _capi.dptr(one), 1, _capi.i64ptr(o1))),
                      reps=3)
    t_measure, _ = _timed(lambda: sim.measure(n // 2, rng=np.random.default_rng(1)), reps=3)
    noise = {}
    for model, p in (("depolarizing", 0.05), ("amplitude_damping", 0.05)):
        t, _ = _timed(lambda: desv.NoiseModel.apply_to_sv(sim.sv, n, model, p, rng=np.random.default_rng(7)),
                      reps=2)
        noise[model] = {"p": p, "ms_per_trajectory": t * 1e3}
    sim.close()
    state_bytes = (1 << n) * 8  # the chunk-total pass reads the state once
    return {"workload": f"{n}q complex64 state: sampling, measurement, noise trajectories",
            "shots": shots, "sample_ms": t_sample * 1e3, "shots_per_s": shots / t_sample,
            "one_shot_ms": t_one * 1e3, "one_shot_state_gbps": state_bytes / t_one / 1e9,
            "measure_ms": t_measure * 1e3, "noise": noise}


def suite_brickwork34(n=34, depth=24, shape=None):
    """The brickwork at 34 qubits: a 128 GiB complex64 state, the largest
    power of two that fits one MI355X's 288 GB (the state is updated in
    place; the fused program has no out-of-place streaming stage).  Skipped
    with a note if the device cannot hold it."""
    import desv
    from desv import workloads
    gates = workloads.brickwork_su4(n, depth, seed=0)
    prog = desv.Program(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    prog.compile("complex64", **dict(shape or dict(fuse_bits=3, use_mfma=True, tile_bits=12)))
    try:
        h = desv.StateHandle(n, complex128=False)
    except MemoryError as e:
        return {"workload": f"brickwork {n}q depth {depth}, c64 (128 GiB state)", "skipped": str(e)}
    try:
        h.set_double_buffer(False)
        prog.run(h)
        h.reset()
        dt, _ = _timed(lambda: prog.run(h), reps=1)
        nrm = h.norm2()
    finally:
        h.close()
    st = prog.stats()
    return {"workload": f"brickwork {n}q depth {depth}, {len(gates)} gates, c64 (128 GiB state, in place)",
            "gates_per_s": len(gates) / dt, "ms_per_circuit": dt * 1e3, "hbm_passes": st["passes_in_place"],
            "hbm_gbps": st["passes_in_place"] * 2 * (1 << n) * 8 / dt / 1e9, "norm_minus_1": nrm - 1}


def copy_probe(n, steps):
    """Copy one 2^n complex64 state into another with the library's streaming
    copy (desv_copy): read + write bytes per second.  Under rocprofv3 --pmc,
    FETCH_SIZE and WRITE_SIZE of this kernel against its known 2^n * 8 bytes
    give the scale factors of the byte counters."""
    import torch
    import desv
    stream = torch.cuda.Stream()
    a = desv.StateHandle(n, complex128=False)
    b = desv.StateHandle(n, complex128=False)
    for h in (a, b):
        desv._capi.check(desv._capi.lib.desv_set_stream(h.ptr, desv._capi.C.c_void_p(stream.cuda_stream)))
    b.copy_from(a)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for _ in range(steps):
        b.copy_from(a)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    nbytes = (1 << n) * 8
    print(json.dumps({"copy_probe": f"desv_copy of a 2^{n} complex64 state", "bytes_each_way": nbytes,
                      "ms": ms, "gbps_read_plus_write": 2 * nbytes / ms / 1e6,
                      "frac_of_peak": 2 * nbytes / ms / 1e6 / HBM_PEAK_GBPS}), flush=True)
    a.close()
    b.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=24)
    ap.add_argument("--fuse", type=int, default=None,
                    help="fusion width; default: the program autotuner picks fusion, tile and MFMA use")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the program autotuner (then fusion 4, MFMA blocks, default tile)")
    ap.add_argument("--tile-bits", type=int, default=0)
    ap.add_argument("--low-bits", type=int, default=0)
    ap.add_argument("--per-gate", action="store_true")
    ap.add_argument("--no-mfma", action="store_true", help="fused blocks on VALU instead of MFMA")
    ap.add_argument("--reg-groups", type=int, default=0,
                    help="with --fuse: 1 = register-group stages, 2 = one tile per wave in registers")
    ap.add_argument("--in-place", action="store_true",
                    help="streaming passes update the state in place (no second buffer)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time the library's device copy of a 2^qubits complex64 state "
                         "(calibration of HBM rate and of the rocprofv3 byte counters)")
    ap.add_argument("--suite", default="streaming30,parity30,brickwork32,brickwork34,plain20,batchjit20,vqe20,batch,"
                    "precision,lanczos,readout",
                    help="extra workloads to measure after the headline ('' for none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: every run here has a single MI355X; "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import desv
    from desv import workloads

    n = args.qubits
    if args.copy_probe:
        copy_probe(n, args.steps)
        return
    mfma = not args.no_mfma
    gates = workloads.brickwork_su4(n, args.depth, seed=0)
    prog = desv.Program(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    tuned = None
    if args.fuse is None and not args.no_autotune and not args.per_gate and not args.tile_bits:
        # launch shape chosen by measurement (desv.tune), not cached: every run
        # times the candidates on this GPU
        t0 = time.perf_counter()
        tuned = desv.tune.autotune_program(prog, "complex64", cache=False)
        if args.no_mfma and tuned["best"].get("use_mfma"):
            tuned = desv.tune.autotune_program(
                prog, "complex64", cache=False,
                candidates=[c for c in desv.tune.program_candidates("complex64", n) if not c["use_mfma"]])
        log(f"autotune ({time.perf_counter() - t0:.1f} s): {tuned['best']} "
            f"{[(c['fuse_bits'], c['tile_bits'], c['use_mfma'], c['ms']) for c in tuned['candidates']]}")
        kw = dict(tuned["best"])
    else:
        kw = dict(fuse_bits=args.fuse if args.fuse is not None else 4, per_gate=args.per_gate, use_mfma=mfma)
        if args.tile_bits:
            kw["tile_bits"] = args.tile_bits
        if args.reg_groups:
            kw["reg_groups"] = args.reg_groups
    if args.low_bits:
        kw["low_bits"] = args.low_bits
    prog.compile("complex64", **kw)
    fuse, mfma = kw["fuse_bits"], bool(kw.get("use_mfma"))
    st = prog.stats()
    kernels = prog.stage_kernels()
    log(f"compiled: {st}")

    torch.cuda.init()
    dev = torch.device("cuda:0")
    # a dedicated stream (default priority): every kernel of the program is
    # launched on it, so events recorded on it bracket exactly that work
    stream = torch.cuda.Stream(device=dev)
    h = desv.StateHandle(n, complex128=False, device=0)
    desv._capi.check(desv._capi.lib.desv_set_stream(h.ptr, desv._capi.C.c_void_p(stream.cuda_stream)))
    h.set_double_buffer(not args.in_place)

    for i in range(args.warmup):
        h.reset()
        prog.run(h)
    torch.cuda.synchronize()
    log("warmup done")
    h.reset()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    stage_ms = np.zeros(len(kernels))
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        # the same launches as prog.run, with an event between stages
        stage_ms += prog.run_timed(h)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    log(f"timed {args.steps} steps: {gpu_ms:.2f} ms (wall {wall*1e3:.1f} ms, stages {stage_ms.sum():.2f} ms)")

    # sanity: the state stays normalised
    nrm = h.norm2()
    log(f"norm^2 after {args.steps} steps: {nrm:.9f}")

    N = 1 << n
    bytes_per_pass = 2 * N * 8
    # a relabelled tail runs in its in-place form when the handle is in place
    passes_run = st["passes_in_place"] if args.in_place else st["passes"]
    ms_per_step = gpu_ms / args.steps
    gates_per_s = len(gates) * args.steps / (gpu_ms / 1e3)

    # per-kernel attribution from the per-stage events
    per = {}
    for (name, flops, mflops), ms in zip(kernels, stage_ms):
        d = per.setdefault(name, {"launches": 0, "ms": 0.0, "flops": 0.0, "mfma_flops": 0.0})
        d["launches"] += args.steps
        d["ms"] += float(ms)
        d["flops"] += flops * args.steps
        d["mfma_flops"] += mflops * args.steps
    dom = max(per, key=lambda k: per[k]["ms"])
    d = per[dom]
    avg_ms = d["ms"] / d["launches"]
    workload_key = (f"brickwork{n}_depth{args.depth}_fuse{fuse}{'_mfma' if mfma and fuse >= 3 else ''}"
                    f"{'_tile' + str(kw['tile_bits']) if kw.get('tile_bits') else ''}"
                    f"{['', '_groups', '_waves'][kw.get('reg_groups', 0)]}"
                    f"{'_pergate' if args.per_gate else ''}")
    traffic = _load_traffic(dom, workload_key)
    if dom == "k_direct":
        achieved = bytes_per_pass / (avg_ms / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBPS}
    else:
        achieved = d["flops"] / d["launches"] / (avg_ms / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / MFMA_FP32_PEAK_TFLOPS}
        roof["mfma_only_tflops"] = d["mfma_flops"] / (d["ms"] / 1e3) / 1e12
        roof["valu_fp32_peak_tflops"] = VALU_FP32_TFLOPS
        roof["hbm_gbps_per_launch"] = bytes_per_pass / (avg_ms / 1e3) / 1e9
    roof.update({"traffic": traffic, "kernel": dom, "launches": d["launches"],
                 "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": bytes_per_pass,
                 "algorithmic_flops_per_launch": d["flops"] / d["launches"]})
    kernel_table = {k: {"launches": v["launches"], "avg_ms": v["ms"] / v["launches"],
                        "share": v["ms"] / max(stage_ms.sum(), 1e-9)} for k, v in per.items()}

    out = {
        "metric": "gates/s + HBM GB/s at 30q c64",
        "value": gates_per_s,
        "unit": "gates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"brickwork Haar-SU(4), {n} qubits, depth {args.depth}, "
                               f"{len(gates)} gates, complex64 state resident in HBM",
                   "qubits": n, "depth": args.depth, "gates": len(gates),
                   "fuse_bits": fuse, "tile_bits": kw.get("tile_bits"), "per_gate": args.per_gate,
                   "mfma": mfma, "reg_groups": kw.get("reg_groups", 0), "shape_from_autotuner": tuned is not None,
                   "hbm_passes_per_step": passes_run, "device_ops_per_step": st["blocks"],
                   "mfma_blocks_per_step": st["mfma_blocks"]},
        # whole-job HBM rate: algorithmic bytes of the passes executed / time
        "hbm_gbps": passes_run * bytes_per_pass * args.steps / (gpu_ms / 1e3) / 1e9,
        # the unfused per-gate equivalent: gates x 2 N b / time
        "gate_equivalent_gbps": len(gates) * bytes_per_pass * args.steps / (gpu_ms / 1e3) / 1e9,
        "roofline": roof,
        "kernels": kernel_table,
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(n, gates)
        log("cpu baselines of the other workloads ...")
        out["cpu_baselines_other"] = cpu_baselines_small()
        log(f"{out['cpu_baselines_other']}")
    del prog, h
    torch.cuda.synchronize()
    extras = {}
    fns = {"streaming30": suite_streaming30, "parity30": suite_parity30, "brickwork32": suite_brickwork32,
           "batchjit20": suite_batchjit20,
           "plain20": suite_plain20, "vqe20": suite_vqe20, "batch": suite_batch,
           "precision": suite_precision, "lanczos": suite_lanczos, "brickwork34": suite_brickwork34,
           "readout": suite_readout}
    for name in [x for x in args.suite.split(",") if x]:
        log(f"suite {name} ...")
        t0 = time.perf_counter()
        extras[name] = fns[name](shape=kw) if name in ("parity30", "brickwork32", "brickwork34") else fns[name]()
        log(f"suite {name}: {time.perf_counter() - t0:.1f} s {extras[name]}")
    if extras:
        out["extras"] = extras
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
