"""Throughput of the Rosenbrock path on one MI355X, in cell-solves per second.

Workload (default): seeded random mechanism of 210 species and 560 reactions
(75 % Arrhenius, 10 % Troe, 5 % ternary chemical activation, 10 % photolysis),
100,000 grid cells with T uniform in 180-320 K, P log-uniform in 1-101 kPa and
photolysis zero in a random 50 % of cells, FP64, 4-stage Rosenbrock, one
Solve of 900 s per step.  Inputs are resident in HBM before timing starts; a
step restores the initial concentrations (device-to-device copy) and runs
Solver::Solve with the persistent form.

Prints one JSON line (rank 0).  `roofline` relates the algorithmic bytes of the
path -- per attempt F x forcing + Jacobian + LU + S x substitution + error, with
per-cell byte counts of each kernel as in the design notes -- summed over the
attempts the cells actually took, to the measured time of the solve kernels
(HIP events on the solver's stream).  `cpu_baseline` times the reference's own
CPU solver (or the CPU restatement, where the reference build is absent) on a
bounded sample of the same cells with one 1-cell State per cell and one worker
thread per host core.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(info, n_species, n_reactions, stats):
    """Bytes of the path's unique reads and writes per kernel invocation, summed over
    the invocations the cells performed (stats columns: function calls, Jacobian
    updates, steps, accepted, rejected, decompositions, solves)."""
    N, R = n_species, n_reactions
    nnzJ, nnzL, nnzU = info.nnz_jacobian, info.nnz_lower, info.nnz_upper
    forcing = 8 * (2 * N + R)
    jacobian = 8 * (N + R + nnzJ)
    lu = 8 * (nnzJ + nnzL + nnzU)
    solve = 8 * ((nnzL + nnzU - 2 * N) + 2 * N)
    error = 8 * 3 * N
    s = stats.sum(axis=0)
    return int(s[0] * forcing + s[1] * jacobian + s[5] * lu + s[6] * solve + s[2] * error)


def cpu_baseline(desc, cells, order, dt, sample, threads):
    """Cell-solves/s of the reference CPU solver (per-cell States) on `sample` cells."""
    from oracle import Oracle, Reference, reference_available
    kind = "reference" if reference_available() else "port"
    idx = np.arange(sample)
    sub = cells.subset(idx)
    y = sub.variables[:, order]
    if kind == "port":
        os.environ["OMP_NUM_THREADS"] = str(threads)
        solver = Oracle(desc, 4)
        t0 = time.perf_counter()
        solver.solve(sub.conditions, sub.custom, y, dt, 1)
        el = time.perf_counter() - t0
        return kind, sample / el, el
    solvers = [Reference(desc, 4) for _ in range(threads)]
    chunks = np.array_split(idx, threads)

    def work(k):
        c = chunks[k]
        if len(c):
            solvers[k].solve(sub.conditions[c], sub.custom[c], y[c], dt, 1)

    t0 = time.perf_counter()
    ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    el = time.perf_counter() - t0
    return kind, sample / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=100000)
    ap.add_argument("--form", default="persistent", choices=["persistent", "plain", "cooperative"])
    ap.add_argument("--cpu-sample", type=int, default=512, help="cells in the CPU-baseline sample (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU-baseline threads (0: all host cores)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a device-to-device copy of 1 GiB (counter calibration for profiles)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 8, "copy_bytes_written": b.numel() * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the per-cell forms need no exchange between GPUs and this build "
              f"runs on a single MI355X; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import torch

    from micm_amd import _lib, synthetic

    torch.cuda.set_device(0)
    desc = synthetic.mechanism_210_560()
    cells = synthetic.inputs_210_560(desc, args.cells)
    dt = 900.0
    mech = _lib.Mechanism(desc)
    order = mech.species_order()
    state = _lib.DeviceState(mech, args.cells)
    stream = torch.cuda.current_stream()
    state.set_stream(stream.cuda_stream)
    state.set_variables(cells.variables[:, order])
    state.set_conditions(cells.conditions)
    state.set_custom_parameters(cells.custom)
    t0 = time.perf_counter()
    state.update_state_parameters()
    torch.cuda.synchronize()
    update_ms = (time.perf_counter() - t0) * 1e3
    params = _lib.rosenbrock_parameters(4)
    form = {"persistent": _lib.SOLVE_PERSISTENT, "plain": _lib.SOLVE_PLAIN,
            "cooperative": _lib.SOLVE_COOPERATIVE}[args.form]

    ptr = state.pointers()
    n_y = ptr.n_groups * 64 * mech.n_species
    y_initial = torch.empty(n_y, dtype=torch.float64, device="cuda")
    _copy_device(y_initial.data_ptr(), ptr.variables, n_y * 8, stream)  # pristine initial state

    def step():
        _copy_device(ptr.variables, y_initial.data_ptr(), n_y * 8, stream)
        state.solve(dt, params, form)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    start_ev, end_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start_ev.record(stream)
    for _ in range(args.steps):
        step()
    end_ev.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = start_ev.elapsed_time(end_ev) / args.steps
    status, stats, _ = state.results()
    ms_per_step = wall * 1e3 / args.steps
    value = args.cells * args.steps / wall

    bytes_per_step = algorithmic_bytes(mech.info, mech.n_species, mech.n_reactions, stats)
    achieved = bytes_per_step / (kernel_ms * 1e-3) / 1e9
    result = {
        "metric": "cell-solves/s, 210-species x 100k cells, persistent batched Rosenbrock (FP64)",
        "value": round(value, 2),
        "unit": "cell-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mechanism and cells)",
        "config": {"workload": f"random 210-species/560-reaction mechanism, {args.cells} cells, 4-stage Rosenbrock, "
                               f"1 x 900 s, 50 % night cells", "form": args.form, "cells": args.cells,
                   "mechanism_seed": desc.seed, "cells_seed": cells.seed},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "solve kernels (k_persistent + k_coop), algorithmic bytes per solve "
                               f"{bytes_per_step}"},
        "attempts_per_step": int(stats[:, 2].sum()),
        "converged_fraction": round(float(np.mean(status == 2)), 4),
        "update_state_parameters_ms": round(update_ms, 3),
    }
    if args.cpu_sample > 0:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        kind, cps, el = cpu_baseline(desc, cells, order, dt, min(args.cpu_sample, args.cells), threads)
        result["cpu_baseline"] = {"value": round(cps, 2), "unit": "cell-solves/s", "cores": threads, "kind": kind,
                                  "sample": f"first {min(args.cpu_sample, args.cells)} cells of the same workload, "
                                            f"1-cell States, {el:.1f} s wall"}
    print(json.dumps(result), flush=True)


def _copy_device(dst, src, nbytes, stream):
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
    hip.hipMemcpyAsync.restype = ctypes.c_int
    rc = hip.hipMemcpyAsync(ctypes.c_void_p(dst), ctypes.c_void_p(src), nbytes, 3, ctypes.c_void_p(stream.cuda_stream))
    if rc != 0:
        raise RuntimeError(f"hipMemcpyAsync failed: {rc}")


if __name__ == "__main__":
    main()
