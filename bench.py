"""Benchmark: 30-qubit complex64 brickwork gate application on one MI355X.

Workload (one step): depth-24 brickwork of Haar-random SU(4) gates on 30
qubits (348 gates, seed 0), applied to a resident 8.6 GB complex64 state.
The circuit is compiled once (fusion + cache-blocked LDS stages); a step is
one replay of the compiled program.

Printed JSON line: gates/s as `value`, plus the whole-job HBM rate, the
roofline of the stage kernel (algorithmic bytes per launch / measured average
launch time from HIP events on the launch stream) and the CPU baseline (the
NumPy oracle applying the same gates at 30 qubits on the host, single BLAS
thread, a bounded sample of gates).
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


def cpu_baseline(n, gates, max_gates=2, budget_s=30.0):
    """NumPy oracle (eager k-qubit contraction) on the host, 1 BLAS thread."""
    import numpy as np
    from threadpoolctl import threadpool_limits

    from oracle import sv_oracle as O
    sv = np.zeros(1 << n, dtype=np.complex64)
    sv[0] = 1
    done = 0
    t0 = time.perf_counter()
    with threadpool_limits(limits=1):
        for u, a, b in gates[:max_gates]:
            sv = O.apply_matrix(sv, u, [a, b])
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "gates/s", "cores": 1, "kind": "port",
            "sample": f"{done} of the {len(gates)} SU(4) gates at {n} qubits, complex64, "
                      f"oracle.sv_oracle.apply_matrix (NumPy), {dt:.1f} s"}


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
    prog = desv.Program(n).add_ops(ops).compile("complex64")
    prog.run(s64.handle)
    dt, _ = _timed(lambda: (s64.handle.reset(), prog.run(s64.handle)), reps=20)
    s128 = desv.DenseSVSimulator(n)
    s128.run_circuit(circ)
    a, b = s64.get_statevector(), s128.get_statevector()
    fid = float(abs(np.vdot(a.astype(np.complex128), b)) ** 2)
    dz = max(abs(desv.pauli_expectation(s64.sv, {q: 'Z'}) - desv.pauli_expectation(s128.sv, {q: 'Z'}))
             for q in range(n))
    return {"workload": "20q random rx/ry/rz + cx ladder, 20 layers, 1580 gates, c64",
            "gates_per_s": len(ops) / dt, "ms_per_circuit": dt * 1e3, "hbm_passes": prog.stats()["passes"],
            "fidelity_vs_c128": fid, "max_abs_dZ_vs_c128": dz}


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
    f128.value_and_grad(theta, H)
    f64.value_and_grad(theta, H)
    t128, (e128, g128) = _timed(lambda: f128.value_and_grad(theta, H), reps=5)
    t64, (e64, g64) = _timed(lambda: f64.value_and_grad(theta, H), reps=5)
    tps, gps = _timed(lambda: desv.parameter_shift_grad(f64.program, theta, H), reps=1)
    return {"workload": "20q HEA 8 layers (472 gates, 320 params), 200 random Pauli terms",
            "adjoint_evals_per_s_c64": 1.0 / t64, "adjoint_evals_per_s_c128": 1.0 / t128,
            "param_shift_evals_per_s_c64": 1.0 / tps,
            "adjoint_over_param_shift_c64": tps / t64,
            "energy_c128": e128, "abs_dE_c64_vs_c128": abs(e64 - e128),
            "max_abs_dgrad_c64_vs_c128": float(np.max(np.abs(g64 - g128))),
            "max_abs_dgrad_param_shift_vs_adjoint_c64": float(np.max(np.abs(gps - g64)))}


def suite_batch(count=1024):
    """1024 heterogeneous VQE instances (n ~ U{10..22}, 4-8 layers, 50 terms):
    instances/s of energy + adjoint gradient."""
    import numpy as np
    from desv import workloads
    from desv.batch import Batch
    insts = workloads.batch_instances(count=count)
    b = Batch("complex64")
    for n, circ, theta, terms in insts:
        b.add(circ, n, theta, terms)
    b.configure(0, True)
    b.run()
    dt, (es, gs) = _timed(lambda: b.run(), reps=2)
    return {"workload": f"{count} VQE instances, n ~ U{{10..22}}, 4-8 HEA layers, 50 terms, c64",
            "instances_per_s": count / dt, "ms_per_batch": dt * 1e3,
            "mean_latency_ms": dt * 1e3 / count}


def suite_precision(n=28, steps=200):
    """XXZ chain (J=1, Delta=0.5) first-order Trotter, dt=0.05, 200 steps, from a
    random product state: complex64 (FP64 norm and <Z_i Z_i+1>) vs complex128."""
    import numpy as np
    import desv
    from desv import workloads
    circ = workloads.xxz_trotter_circuit(n, 0.05, steps, seed=0)
    res = {}
    for dt_name, f32 in (("c64", True), ("c128", False)):
        sim = desv.DenseSVSimulator(n, use_float32=f32)
        ops = sim.compile_ops(circ)
        prog = desv.Program(n).add_ops(ops).compile("complex64" if f32 else "complex128")
        t, _ = _timed(lambda: prog.run(sim.handle), reps=1)
        zz = desv.pauli_sum_expectation(sim.sv, [(1.0, {i: 'Z', i + 1: 'Z'}) for i in range(n - 1)])
        zzs = [desv.pauli_expectation(sim.sv, {i: 'Z', i + 1: 'Z'}) for i in range(n - 1)]
        res[dt_name] = (t, sim.handle.norm2(), np.array(zzs), prog.stats(), len(ops))
        sim.close()
    (t64, n64, z64, st64, g), (t128, n128, z128, st128, _) = res["c64"], res["c128"]
    return {"workload": f"{n}q XXZ Trotter, {steps} steps, {g} gates",
            "gates_per_s_c64": g / t64, "gates_per_s_c128": g / t128,
            "hbm_passes_c64": st64["passes"],
            "norm_c64_minus_1": n64 - 1.0, "norm_c128_minus_1": n128 - 1.0,
            "max_abs_dZZ_c64_vs_c128": float(np.max(np.abs(z64 - z128)))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=24)
    ap.add_argument("--fuse", type=int, default=2)
    ap.add_argument("--tile-bits", type=int, default=0)
    ap.add_argument("--low-bits", type=int, default=0)
    ap.add_argument("--per-gate", action="store_true")
    ap.add_argument("--mfma", action="store_true")
    ap.add_argument("--in-place", action="store_true",
                    help="streaming passes update the state in place (no second buffer)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--suite", default="plain20,vqe20,batch,precision",
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
    gates = workloads.brickwork_su4(n, args.depth, seed=0)
    prog = desv.Program(n)
    for u, a, b in gates:
        prog.add_matrix(u, [a, b])
    kw = dict(fuse_bits=args.fuse, per_gate=args.per_gate, use_mfma=args.mfma)
    if args.tile_bits:
        kw["tile_bits"] = args.tile_bits
    if args.low_bits:
        kw["low_bits"] = args.low_bits
    prog.compile("complex64", **kw)
    st = prog.stats()
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
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        prog.run(h)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    log(f"timed {args.steps} steps: {gpu_ms:.2f} ms (wall {wall*1e3:.1f} ms)")

    # sanity: the state stays normalised
    nrm = h.norm2()
    log(f"norm^2 after {args.steps + 0} steps: {nrm:.9f}")

    amp_bytes = 8
    N = 1 << n
    bytes_per_pass = 2 * N * amp_bytes
    launches = st["passes"] * args.steps
    avg_launch_s = gpu_ms / 1e3 / launches
    achieved = bytes_per_pass / avg_launch_s / 1e9
    ms_per_step = gpu_ms / args.steps
    gates_per_s = len(gates) * args.steps / (gpu_ms / 1e3)

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
                   "fuse_bits": args.fuse, "per_gate": args.per_gate, "mfma": args.mfma,
                   "hbm_passes_per_step": st["passes"], "device_ops_per_step": st["blocks"]},
        "hbm_gbps": st["passes"] * bytes_per_pass * args.steps / (gpu_ms / 1e3) / 1e9,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "k_stage<float,1>",
                     "bytes_per_launch": bytes_per_pass, "avg_launch_ms": avg_launch_s * 1e3},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(n, gates)
    del prog, h
    extras = {}
    fns = {"plain20": suite_plain20, "vqe20": suite_vqe20, "batch": suite_batch,
           "precision": suite_precision}
    for name in [x for x in args.suite.split(",") if x]:
        log(f"suite {name} ...")
        t0 = time.perf_counter()
        extras[name] = fns[name]()
        log(f"suite {name}: {time.perf_counter() - t0:.1f} s {extras[name]}")
    if extras:
        out["extras"] = extras
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
