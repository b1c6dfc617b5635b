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
    ap.add_argument("--no-cpu-baseline", action="store_true")
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
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
