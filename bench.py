"""Throughput of the Rosenbrock path on one MI355X, in cell-solves per second.

Workload (default): seeded random mechanism of 210 species and 560 reactions
(75 % Arrhenius, 10 % Troe, 5 % ternary chemical activation, 10 % photolysis),
100,000 grid cells with T uniform in 180-320 K, P log-uniform in 1-101 kPa and
photolysis zero in a random 50 % of cells, FP64, 4-stage Rosenbrock, one
Solve of 900 s per step.  Inputs are resident in HBM before timing starts; a
step restores the initial concentrations (device-to-device copy) and runs
Solver::Solve with the per-cell form the autotuner picks (micm_amd/autotune.py;
every form gives bit-identical results).  `--workload ts1` runs the reference's
own benchmark instead: TS1, 10,000 cells, 3-stage, 30 s steps; `--workload packed`
three mechanisms of different sizes in one graph-captured launch (main_packed).

Prints one JSON line.  `roofline` relates the algorithmic bytes of the path --
per attempt F x forcing + Jacobian + LU + S x substitution + error, with the
per-cell byte counts of each kernel -- summed over the attempts the cells
actually took, to the measured time of the Solve launches (HIP events on the
solver's stream); `fp64` does the same with the FP64 operation count.
`plain_form` times the plain per-step form and the chosen form on the same
strided sample of cells.  `end_to_end` includes the PCIe
transfers of inputs and results.  `cpu_baseline` times the reference's own CPU
solver (or the CPU restatement, where the reference build is absent) on a
bounded sample of the same cells with one 1-cell State per cell, on every host
core and on one.
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
FP64_PEAK_TFLOPS = 78.6


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


def flops_per_invocation(desc, mech):
    """FP64 operations of one invocation of each phase, counted from the mechanism:
    forcing k * prod(y) plus one update per reactant and a multiply-add per product
    (process_set.hpp:524-590); the Jacobian's per-independent-reactant product and
    updates (:594-670); Doolittle's multiply-subtracts (one per L[i,j] * U[j,k] pair
    with j below both indices of the filled element) and divisions plus the alpha
    shift (lu_decomposition_doolittle.inl:325-403); both substitutions with their
    diagonal divisions (linear_solver.inl:235-294); the error norm's ~7 operations
    per species.  Stage combinations (a few N per stage) are left out."""
    nr = np.diff(desc.reactant_offsets)
    npr = np.diff(desc.product_offsets)
    N = mech.n_species
    pat = mech.patterns()
    lr, lc = pat["lower"]
    ur, uc = pat["upper"]
    strict_l = lr[: mech.nnz_lower] > lc[: mech.nnz_lower]
    strict_u = ur[: mech.nnz_upper] < uc[: mech.nnz_upper]
    lb = np.zeros((N, N), dtype=np.int64)
    ub = np.zeros((N, N), dtype=np.int64)
    lb[lr[: mech.nnz_lower][strict_l], lc[: mech.nnz_lower][strict_l]] = 1
    ub[ur[: mech.nnz_upper][strict_u], uc[: mech.nnz_upper][strict_u]] = 1
    pairs = lb @ ub  # pairs[i, k] = #{j < min(i, k): L[i, j] != 0 and U[j, k] != 0}
    filled = (lb + ub) > 0
    fma = int(pairs[filled].sum()) + int(np.trace(pairs))
    n_l, n_u = int(strict_l.sum()), int(strict_u.sum())
    return {"forcing": int(np.sum(2 * nr + 2 * npr)),
            "jacobian": int(np.sum(nr * (2 * nr - 1 + 2 * npr))),
            "lu": 2 * fma + n_l + N, "lu_multiply_subtracts": fma, "lu_divisions": n_l,
            "solve": 2 * (n_l + n_u) + 2 * N,
            "error": 7 * N}


def algorithmic_flops(per, stats):
    """FP64 operations summed over the invocations the cells performed (stats columns
    as in algorithmic_bytes)."""
    s = stats.sum(axis=0)
    return int(s[0] * per["forcing"] + s[1] * per["jacobian"] + s[5] * per["lu"] + s[6] * per["solve"] +
               s[2] * per["error"])


def cpu_baseline(desc, cells, order, dt, sample, threads, tableau=4, results=None, min_seconds=0.5):
    """Cell-solves/s of the reference CPU solver (per-cell States) on `sample` cells,
    the whole sample solved again until `min_seconds` have passed (a small sample is
    otherwise dominated by starting the threads).  With a dict `results`, its y, states
    and stats for those cells are stored there.  Returns (kind, rate, seconds, repeats)."""
    from oracle import Oracle, Reference, reference_available
    kind = "reference" if reference_available() else "port"
    idx = np.arange(sample)
    sub = cells.subset(idx)
    y = sub.variables[:, order]
    if kind == "port":
        solver = Oracle(desc, tableau)
        solver.lib.oracle_set_threads(int(threads))
        reps = 0
        t0 = time.perf_counter()
        while True:
            out = solver.solve(sub.conditions, sub.custom, y, dt, 1)
            reps += 1
            el = time.perf_counter() - t0
            if el >= min_seconds or reps >= 10000:
                break
        if results is not None:
            results.update(y=out[0], states=out[1], stats=out[2])
        return kind, sample * reps / el, el, reps
    solvers = [Reference(desc, tableau) for _ in range(threads)]
    chunks = np.array_split(idx, threads)
    outs = [None] * threads

    def work(k):
        c = chunks[k]
        if len(c):
            outs[k] = solvers[k].solve(sub.conditions[c], sub.custom[c], y[c], dt, 1)

    reps = 0
    t0 = time.perf_counter()
    while True:
        ts = [threading.Thread(target=work, args=(k,)) for k in range(threads)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or reps >= 10000:
            break
    if results is not None:
        done = [o for o in outs if o is not None]
        results.update(y=np.concatenate([o[0] for o in done]), states=np.concatenate([o[1] for o in done]),
                       stats=np.concatenate([o[2] for o in done]))
    return kind, sample * reps / el, el, reps


def parity(y_gpu, st_gpu, stats_gpu, ref):
    """The GPU's results against the reference's on the same cells
    (1e-8 relative on species above 1e-20 mol m-3), bitwise equality and SolverStats."""
    y_ref = ref["y"]
    n = y_ref.shape[0]
    g = y_gpu[:n]
    mask = np.abs(y_ref) > 1e-20
    rel = np.where(mask, np.abs(g - y_ref) / np.where(mask, np.abs(y_ref), 1.0), 0.0)
    return {"cells": int(n),
            "within_1e-8": round(float(np.mean(rel.max(axis=1) <= 1e-8)), 6),
            "bitwise_equal": round(float(np.mean(np.all(g.view(np.int64) == y_ref.view(np.int64), axis=1))), 6),
            "identical_stats": round(float(np.mean(np.all(stats_gpu[:n] == ref["stats"], axis=1))), 6),
            "identical_outcome": round(float(np.mean(st_gpu[:n] == ref["states"])), 6),
            "max_rel_error": float(rel.max())}


def accuracy(desc, cells, order, dt, y_gpu, tableau, n_sample, threads):
    """Error of the GPU's results against a tight-tolerance FP64 CPU solve (rtol 1e-12,
    atol 1e-20, step limit 10^7) on `n_sample` evenly spaced cells: the largest relative
    error on species above 1e-20 mol m-3 and the error in the solver's own norm
    (|y - y_tight| / (atol + rtol |y_tight|) with the run's tolerances)."""
    from oracle import Oracle
    idx = np.linspace(0, cells.n_cells - 1, n_sample).round().astype(np.int64)
    sub = cells.subset(idx)
    o = Oracle(desc, tableau)
    o.lib.oracle_set_threads(int(threads))
    o.set_max_steps(10 ** 7)
    t0 = time.perf_counter()
    y_t, st_t, _, _ = o.solve(sub.conditions, sub.custom, sub.variables[:, order], dt, 1, rtol=1e-12,
                              atol=np.full(o.n_species, 1e-20))
    el = time.perf_counter() - t0
    g = y_gpu[idx]
    mask = np.abs(y_t) > 1e-20
    rel = np.where(mask, np.abs(g - y_t) / np.where(mask, np.abs(y_t), 1.0), 0.0).max(axis=1)
    w = (np.abs(g - y_t) / (1e-3 + 1e-6 * np.abs(y_t))).max(axis=1)
    return {"cells": int(n_sample), "tight_solve": "rtol 1e-12, atol 1e-20", "seconds": round(el, 1),
            "tight_converged": int(np.sum(st_t == 2)),
            "max_rel_error_species_above_1e-20": {"median": float(np.median(rel)), "max": float(rel.max())},
            "weighted_error_atol_1e-3_rtol_1e-6": {"median": float(np.median(w)), "max": float(w.max())}}


def cpu_lockstep(desc, sub, order, dt, tableau):
    """Cell-solves/s of the reference CPU solver with all of `sub`'s cells in one
    State (one step size for every cell; a State runs on one core)."""
    from oracle import Oracle, Reference, reference_available
    kind = "reference" if reference_available() else "port"
    solver = (Reference if kind == "reference" else Oracle)(desc, tableau)
    if kind == "port":
        solver.lib.oracle_set_threads(1)
    t0 = time.perf_counter()
    solver.solve(sub.conditions, sub.custom, sub.variables[:, order], dt, 1, lockstep=True)
    el = time.perf_counter() - t0
    return kind, sub.n_cells / el, el


def kernel_rooflines(mech, state, stream, n_cells, repeats=5):
    """Per-kernel time of the single-phase entry points over the whole State, and their
    algorithmic bytes as a fraction of the HBM peak (the accumulating forcing and Jacobian
    entries also read their output: 8 N and 8 nnzJ bytes per cell beyond the count)."""
    import torch
    N, R, P = mech.n_species, mech.n_reactions, mech.n_custom
    i = mech.info
    ptr = state.pointers()
    groups = ptr.n_groups
    f = torch.zeros(groups * 64 * N, dtype=torch.float64, device="cuda")
    x = torch.zeros_like(f)
    err = torch.zeros(n_cells, dtype=torch.float64, device="cuda")
    st = stream.cuda_stream
    per_cell = {
        "rate_constants": 8 * (3 + P + R),
        "add_forcing": 8 * (2 * N + R),
        "fill_and_add_forcing": 8 * (2 * N + R),
        "subtract_jacobian": 8 * (N + R + i.nnz_jacobian),
        "fill_and_subtract_jacobian": 8 * (N + R + i.nnz_jacobian),
        "alpha_minus_jacobian": 8 * 2 * N,
        "lu_decompose": 8 * (i.nnz_jacobian + i.nnz_lower + i.nnz_upper),
        "lu_solve": 8 * (i.nnz_lower + i.nnz_upper),
        "normalized_error": 8 * 3 * N,
    }
    calls = {
        "rate_constants": lambda: mech.calculate_rate_constants(n_cells, ptr.conditions, ptr.custom_parameters,
                                                                ptr.rate_constants, st),
        "add_forcing": lambda: mech.add_forcing_terms(n_cells, ptr.rate_constants, ptr.variables, f.data_ptr(), st),
        "subtract_jacobian": lambda: mech.subtract_jacobian_terms(n_cells, ptr.rate_constants, ptr.variables,
                                                                  ptr.jacobian, st),
        # the pairs Solve issues at a step start: Fill(0) then AddForcingTerms /
        # SubtractJacobianTerms, the output only written
        "fill_and_add_forcing": lambda: mech.fill_and_add_forcing_terms(n_cells, ptr.rate_constants, ptr.variables,
                                                                        f.data_ptr(), st),
        "fill_and_subtract_jacobian": lambda: mech.fill_and_subtract_jacobian_terms(
            n_cells, ptr.rate_constants, ptr.variables, ptr.jacobian, st),
        "alpha_minus_jacobian": lambda: mech.alpha_minus_jacobian(n_cells, ptr.jacobian, 1.0e3, st),
        "lu_decompose": lambda: mech.lu_decompose(n_cells, ptr.jacobian, ptr.lower, ptr.upper, st),
        "lu_solve": lambda: mech.lu_solve(n_cells, ptr.lower, ptr.upper, x.data_ptr(), st),
        "normalized_error": lambda: mech.normalized_error(n_cells, ptr.variables, ptr.variables, f.data_ptr(),
                                                          ptr.absolute_tolerances, 1e-6, err.data_ptr(), st),
    }
    out = {}
    for name, call in calls.items():
        if name == "lu_solve":
            x.copy_(f)
        call()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        for _ in range(repeats):
            call()
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / repeats
        gbs = per_cell[name] * n_cells / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "bytes_per_cell": per_cell[name], "GB/s": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def measured_traffic(workload_name, n_cells, threads, stage):
    """HBM bytes per Solve launch from a committed rocprofv3 PMC summary
    (profiles/*_hbm_traffic.md, FETCH_SIZE and WRITE_SIZE runs scaled by the same
    counters on a 1 GiB copy) when one covers this workload, cell count and
    cooperative shape (threads and the requested stage, which fixes the build's
    register bound); None otherwise."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_hbm_traffic.md"))):
        text = open(path).read()
        for block in text.split("```json")[1:]:
            rec = json.loads(block.split("```")[0])
            if (rec.get("workload"), rec.get("cells"), rec.get("kernel"), rec.get("stage")) == \
                    (workload_name, n_cells, f"k_coop<{threads}>", stage):
                return rec["traffic_bytes"], os.path.relpath(path, ROOT)
    return None, None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def workload(name, n_cells):
    """(description, cell inputs, tableau preset, time step, label) of a named workload."""
    from micm_amd import synthetic
    if name == "210_560":
        desc = synthetic.mechanism_210_560()
        cells = synthetic.inputs_210_560(desc, n_cells)
        return desc, cells, 4, 900.0, (f"random 210-species/560-reaction mechanism (seed {desc.seed}), {n_cells} cells "
                                      f"(seed {cells.seed}), T 180-320 K, P 1-101 kPa, 50 % night, 4-stage Rosenbrock, "
                                      f"1 x 900 s")
    if name == "210_560_literal":
        # the configuration's text taken literally: products from every species
        desc = synthetic.mechanism_210_560(products_from="all")
        cells = synthetic.inputs_210_560(desc, n_cells)
        return desc, cells, 4, 900.0, (f"random 210-species/560-reaction mechanism with products drawn from every "
                                      f"species (seed {desc.seed}), {n_cells} cells (seed {cells.seed}), T 180-320 K, "
                                      f"P 1-101 kPa, 50 % night, 4-stage Rosenbrock, 1 x 900 s")
    if name == "60_200":
        desc = synthetic.mechanism_60_200()
        cells = synthetic.inputs_60_200(desc, n_cells)
        return desc, cells, 4, 300.0, (f"random 60-species/200-reaction mechanism (seed {desc.seed}), {n_cells} cells "
                                      f"(seed {cells.seed}), T 220-310 K, P 101325 Pa, 4-stage Rosenbrock, 1 x 300 s")
    if name == "readme":
        desc = synthetic.readme_mechanism()
        cells = synthetic.readme_inputs(n_cells)
        return desc, cells, 3, 500.0, (f"the reference README's 3-species system, {n_cells} cells, T 270-310 K, "
                                      f"P 90-105 kPa, 3-stage Rosenbrock, 1 x 500 s")
    if name == "ts1":
        from micm_amd import mechanism_io
        desc = mechanism_io.load(os.path.join(ROOT, "tests", "golden", "ts1_mechanism.json"))
        cond, y, custom = mechanism_io.initial_arrays(desc, n_cells)
        cells = synthetic.CellInputs(cond, y, custom, None)
        return desc, cells, 3, 30.0, (f"TS1 (209 species, 547 reactions), {n_cells} identical cells of the reference "
                                      f"benchmark's initial state, 3-stage Rosenbrock, 1 x 30 s")
    raise SystemExit(f"unknown workload {name}")


def log(msg):
    """Progress on stderr; stdout carries only the result line."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="210_560",
                    choices=["210_560", "210_560_literal", "60_200", "readme", "ts1", "packed", "adjoint",
                             "adjoint_210_560"])
    ap.add_argument("--cells", type=int, default=0,
                    help="cells (0: 100000 for 210_560, 10000 for 60_200 and ts1, 1000 for readme)")
    ap.add_argument("--form", default="autotune", choices=["autotune", "persistent", "plain", "cooperative"])
    ap.add_argument("--coop-threads", type=int, default=0, help="threads per cell of the cooperative form (0: library)")
    ap.add_argument("--coop-stage", type=int, default=-1,
                    help="table staging / layout of the cooperative form (micm_hip_state_set_cooperative_shape)")
    ap.add_argument("--handoff", type=int, default=-1, help="persistent form: hand-off step count (-1: library)")
    ap.add_argument("--compare-sample", type=int, default=8192,
                    help="cells on which the plain form and the chosen form are timed side by side (0: skip)")
    ap.add_argument("--no-kernels", action="store_true", help="skip the per-kernel roofline of the single-phase kernels")
    ap.add_argument("--cpu-sample", type=int, default=-1,
                    help="cells in the CPU-baseline sample (0: skip; -1: about 15 core-seconds of CPU work, "
                         "sized from the single-core leg)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU-baseline threads (0: all host cores)")
    ap.add_argument("--accuracy-cells", type=int, default=0,
                    help="cells whose results are compared with a tight-tolerance CPU solve (rtol 1e-12; "
                         "tens of CPU-seconds per 210-species cell)")
    ap.add_argument("--no-full-state-forms", action="store_true",
                    help="skip timing the persistent per-lane form on the whole state")
    ap.add_argument("--check-cells", type=int, default=4, help="adjoint workloads: cells of the gradient checks")
    ap.add_argument("--no-graph", action="store_true", help="launch the Solve directly instead of as a graph")
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
        print(f"bench.py measures one GPU (the per-cell forms need no exchange between GPUs; a multi-GPU run "
              f"shards cells, one process per GPU); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload == "packed":
        main_packed(args)
        return
    if args.workload.startswith("adjoint"):
        main_adjoint(args)
        return

    import torch

    from micm_amd import _lib, autotune

    n_cells = args.cells or {"210_560": 100000, "210_560_literal": 100000, "readme": 1000}.get(args.workload, 10000)
    torch.cuda.set_device(0)
    # a stream of our own: the Solve's graph is captured on it (the legacy default stream cannot be)
    torch.cuda.set_stream(torch.cuda.Stream())
    desc, cells, tableau, dt, label = workload(args.workload, n_cells)
    log(f"workload {args.workload}: {n_cells} cells generated")
    mech = _lib.Mechanism(desc)
    order = mech.species_order()
    params = _lib.rosenbrock_parameters(tableau)
    y_host = np.ascontiguousarray(cells.variables[:, order])

    # end to end from host buffers (PCIe included): inputs up, rate constants, Solve, results down
    state = _lib.DeviceState(mech, n_cells)
    stream = torch.cuda.current_stream()
    state.set_stream(stream.cuda_stream)

    t0 = time.perf_counter()
    state.set_variables(y_host)
    state.set_conditions(cells.conditions)
    state.set_custom_parameters(cells.custom)
    torch.cuda.synchronize()
    upload_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    state.update_state_parameters()
    torch.cuda.synchronize()
    update_ms = (time.perf_counter() - t0) * 1e3

    log(f"inputs uploaded ({upload_s:.2f} s), rate constants {update_ms:.1f} ms")
    if args.form == "autotune":
        choice = autotune.autotune(mech, state, dt, params, verbose=True)
        form = autotune.apply(state, choice)
        form_label = f"autotuned: {choice['label']}" + (f" (hand-off to {choice['handoff_to']})"
                                                         if choice.get("handoff_to") else "")
        coop_threads = choice["threads"] if form == _lib.SOLVE_COOPERATIVE else 0
        coop_stage = choice["stage"]
    else:
        form = {"persistent": _lib.SOLVE_PERSISTENT, "plain": _lib.SOLVE_PLAIN,
                "cooperative": _lib.SOLVE_COOPERATIVE}[args.form]
        state.set_cooperative_shape(args.coop_threads, args.coop_stage)
        state.set_handoff(args.handoff)
        form_label = args.form + (f" ({args.coop_threads} threads per cell)" if args.coop_threads else "")
        coop_threads = args.coop_threads if form == _lib.SOLVE_COOPERATIVE else 0
        coop_stage = args.coop_stage
        choice = None

    ptr = state.pointers()
    n_y = ptr.n_groups * 64 * mech.n_species
    y_initial = torch.empty(n_y, dtype=torch.float64, device="cuda")
    _copy_device(y_initial.data_ptr(), ptr.variables, n_y * 8, stream)  # pristine initial state

    # the per-cell forms run as one graph (the Solve's launches and the cost-order pass),
    # captured at the first step, again at the second (the cost order exists from then on), replayed after
    packed = None
    if form in (_lib.SOLVE_COOPERATIVE, _lib.SOLVE_PERSISTENT) and not args.no_graph:
        packed = _lib.PackedSolve([state], form)
        packed.set_stream(stream.cuda_stream)

    def step():
        _copy_device(ptr.variables, y_initial.data_ptr(), n_y * 8, stream)
        if packed is not None:
            packed.solve(dt, params)
        else:
            state.solve(dt, params, form)

    log(f"form: {form_label}")
    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"warm-up step {i + 1} done")
    start_ev, end_ev = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start_ev.record(stream)
    for _ in range(args.steps):
        step()
    end_ev.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"{args.steps} timed steps: {wall:.2f} s")
    kernel_ms = start_ev.elapsed_time(end_ev) / args.steps
    status, stats, _ = state.results()
    ms_per_step = wall * 1e3 / args.steps
    value = n_cells * args.steps / wall
    graph_replayed = packed.replayed if packed is not None else False

    t0 = time.perf_counter()
    y_out = state.get_variables()
    download_s = time.perf_counter() - t0
    end_to_end = n_cells / (upload_s + update_ms * 1e-3 + ms_per_step * 1e-3 + download_s)

    bytes_per_step = algorithmic_bytes(mech.info, mech.n_species, mech.n_reactions, stats)
    achieved = bytes_per_step / (kernel_ms * 1e-3) / 1e9
    flops_per_step = algorithmic_flops(flops_per_invocation(desc, mech), stats)
    tflops = flops_per_step / (kernel_ms * 1e-3) / 1e12
    traffic, traffic_source = (measured_traffic(args.workload, n_cells, coop_threads, coop_stage) if coop_threads
                               else (None, None))
    conv = status == 2
    result = {
        "metric": {"210_560": "cell-solves/s, 210-species x 100k cells, batched Rosenbrock (FP64)",
                   "210_560_literal": f"cell-solves/s, 210-species (products from every species) x {n_cells} cells, "
                                      "batched Rosenbrock (FP64)",
                   "60_200": f"cell-solves/s, 60-species x {n_cells} cells, Rosenbrock (FP64)",
                   "readme": f"cell-solves/s, README 3-species system x {n_cells} cells (FP64)",
                   "ts1": "cell-solves/s, TS1 x 10k cells (reference benchmark mechanism)"}[args.workload],
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
        "data": "the reference's TS1 benchmark mechanism and initial state" if args.workload == "ts1"
                else "synthetic (seeded mechanism and cells)",
        "config": {"workload": label, "form": form_label, "cells": n_cells, "tableau_stages": int(params.stages),
                   "time_step_s": dt, "graph_captured": packed is not None, "graph_replayed": graph_replayed},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": traffic_source,
                     "traffic_gbs": round(traffic / (kernel_ms * 1e-3) / 1e9, 2) if traffic else None,
                     "limiter": ("latency: the factorisation's and substitutions' level schedules, a few dependent "
                                 "LDS round trips per level, as many cells per CU as LDS holds" if coop_threads else
                                 "latency and HBM: one cell per lane over the grouped arrays"),
                     "kernel": "the Solve launch (one persistent or cooperative kernel per step); algorithmic bytes "
                               f"per step {bytes_per_step} summed over the attempts the cells took; `achieved` "
                               "prices them as per-kernel HBM streams, `traffic_gbs` is what the fused kernel "
                               "moved (rocprofv3 PMC)"},
        "fp64": {"achieved": round(tflops, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": round(tflops / FP64_PEAK_TFLOPS, 6), "flops_per_step": flops_per_step},
        "attempts_per_step": int(stats[:, 2].sum()),
        "attempts_by_outcome": {"converged": int(stats[conv, 2].sum()), "not_converged": int(stats[~conv, 2].sum())},
        "converged_fraction": round(float(np.mean(conv)), 4),
        "converged_value": {"value": round(float(conv.sum()) * args.steps / wall, 2), "unit": "cell-solves/s",
                            "note": "converged cells only, over the same wall time"},
        "update_state_parameters_ms": round(update_ms, 3),
        "end_to_end": {"value": round(end_to_end, 2), "unit": "cell-solves/s",
                       "note": "host inputs uploaded, rate constants, one Solve, results downloaded (PCIe included)"},
    }
    if choice is not None:
        result["autotune"] = {k: (round(v * 1e3, 2) if isinstance(v, float) else v)
                              for k, v in choice["seconds"].items()}
        result["autotune_sample_cells"] = choice["sample"]
    lock_cells = None
    if args.compare_sample > 0:
        m = min(args.compare_sample, n_cells)
        sub = cells.subset(np.linspace(0, n_cells - 1, m).round().astype(np.int64))
        probe = _lib.DeviceState(mech, m)
        probe.set_variables(sub.variables[:, order])
        probe.set_conditions(sub.conditions)
        probe.set_custom_parameters(sub.custom)
        probe.update_state_parameters()
        if choice is not None:
            autotune.apply(probe, choice)
        cmp = {}
        for name, f in (("plain", _lib.SOLVE_PLAIN), ("chosen", form)):
            probe.set_variables(sub.variables[:, order])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            probe.solve(dt, params, f)
            cmp[name] = m / (time.perf_counter() - t0)
        log("plain form compared")
        result["plain_form"] = {"value": round(cmp["plain"], 2), "unit": "cell-solves/s", "cells": m,
                                "chosen_form_same_cells": round(cmp["chosen"], 2),
                                "speedup": round(cmp["chosen"] / cmp["plain"], 3)}
        # the reference's own multi-cell State: one step size for every cell
        ml = min(m, 512)
        lock = _lib.DeviceState(mech, ml)
        subl = lock_cells = sub.subset(np.arange(ml))
        lock.set_variables(subl.variables[:, order])
        lock.set_conditions(subl.conditions)
        lock.set_custom_parameters(subl.custom)
        lock.update_state_parameters()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lock.solve(dt, params, _lib.SOLVE_LOCKSTEP)
        el = time.perf_counter() - t0
        st_l, stats_l, _ = lock.results()
        log("lockstep form compared")
        result["lockstep_form"] = {"value": round(ml / el, 2), "unit": "cell-solves/s", "cells": ml,
                                   "steps": int(stats_l[0, 2]), "converged": bool(np.all(st_l == 2)),
                                   "note": "one step size for all cells (the reference's multi-cell State)"}
    if not args.no_full_state_forms and form == _lib.SOLVE_COOPERATIVE and args.workload == "210_560":
        # the per-lane persistent form on every cell (the autotuner decides on samples)
        full = {}
        for lbl, handoff in (("persistent+handoff48", 48), ("persistent+handoff16", 16)):
            state.set_handoff(handoff)
            _copy_device(ptr.variables, y_initial.data_ptr(), n_y * 8, stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            state.solve(dt, params, _lib.SOLVE_PERSISTENT)
            full[lbl] = round(n_cells / (time.perf_counter() - t0), 2)
            log(f"full state, {lbl}: {full[lbl]} cell-solves/s")
        if choice is not None:
            autotune.apply(state, choice)
        result["forms_full_state"] = {**full, "chosen": round(value, 2), "unit": "cell-solves/s"}
    if not args.no_kernels:
        result["kernels"] = kernel_rooflines(mech, state, stream, n_cells)
        log("single-phase kernels timed")
    if args.accuracy_cells > 0:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        result["accuracy"] = accuracy(desc, cells, order, dt, y_out, tableau, args.accuracy_cells, threads)
        log("accuracy against the tight CPU solve done")
    if args.cpu_sample != 0:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        # one core first; its rate sizes the all-core sample when none is given
        sample1 = min(n_cells, 16 if args.cpu_sample < 0 else max(8, args.cpu_sample // threads))
        kind1, cps1, el1, reps1 = cpu_baseline(desc, cells, order, dt, sample1, 1, tableau)
        sample = min(n_cells, args.cpu_sample if args.cpu_sample > 0 else max(256, int(15.0 * cps1)))
        ref = {}
        kind, cps, el, reps = cpu_baseline(desc, cells, order, dt, sample, threads, tableau, results=ref)
        log(f"CPU baseline: {sample} cells x {reps} on {threads} cores in {el:.1f} s")
        # the timed Solve's own results on those cells against the reference's
        result["parity_sample"] = parity(y_out, status, stats, ref)
        result["cpu_baseline"] = {"value": round(cps, 2), "unit": "cell-solves/s", "cores": threads, "kind": kind,
                                  "cpu": cpu_model(),
                                  "sample": f"first {sample} cells of the same workload, one 1-cell State per "
                                            f"cell, solved {reps} time(s), {el:.1f} s wall on {threads} cores",
                                  "single_core_value": round(cps1, 2),
                                  "single_core_sample": f"first {sample1} cells, solved {reps1} time(s), {el1:.1f} s"}
        if args.workload == "readme" and lock_cells is not None:
            kind_l, cps_l, el_l = cpu_lockstep(desc, lock_cells, order, dt, tableau)
            result["cpu_baseline"]["lockstep"] = {
                "value": round(cps_l, 2), "unit": "cell-solves/s", "cores": 1, "kind": kind_l,
                "sample": f"the lockstep form's {lock_cells.n_cells} cells in one State, {el_l:.3f} s"}
    print(json.dumps(result), flush=True)


PACKED_PARTS = (("20/50", "mechanism_20_50", 200000), ("60/200", "mechanism_60_200", 50000),
                ("210/560", "mechanism_210_560", 20000))


def main_packed(args):
    """Three seeded mechanisms of different sizes (20/50 x 200k cells, 60/200 x 50k,
    210/560 x 20k; T 180-320 K, P log-uniform 1-101 kPa, 50 % night, 4-stage, 1 x 900 s)
    solved by one graph-captured launch (PackedSolve: the three cooperative launches back
    to back), each in the cooperative shape the autotuner picks for its mechanism.  `value` is
    all cells over the wall time of the packed solve; `separate` times the three
    solves one after another without the graph."""
    import torch

    from micm_amd import _lib, autotune, synthetic
    torch.cuda.set_device(0)
    dt, tableau = 900.0, 4
    params = _lib.rosenbrock_parameters(tableau)
    stream = torch.cuda.current_stream()
    coop = [c for c in autotune.CANDIDATES if c[1] == _lib.SOLVE_COOPERATIVE]
    parts = []
    for i, (label, maker, n) in enumerate(PACKED_PARTS):
        desc = getattr(synthetic, maker)()
        cells = synthetic.random_inputs(desc, n, 31 + i, (180.0, 320.0), log_pressure=(1.0e3, 1.01e5),
                                        night_fraction=0.5)
        mech = _lib.Mechanism(desc)
        order = mech.species_order()
        state = _lib.DeviceState(mech, n)
        state.set_variables(np.ascontiguousarray(cells.variables[:, order]))
        state.set_conditions(cells.conditions)
        state.set_custom_parameters(cells.custom)
        state.update_state_parameters()
        choice = autotune.autotune(mech, state, dt, params, candidates=coop)
        autotune.apply(state, choice)
        ptr = state.pointers()
        n_y = ptr.n_groups * 64 * mech.n_species
        y0 = torch.empty(n_y, dtype=torch.float64, device="cuda")
        _copy_device(y0.data_ptr(), ptr.variables, n_y * 8, stream)
        parts.append(dict(label=label, desc=desc, cells=cells, mech=mech, order=order, state=state, ptr=ptr, n_y=n_y,
                          y0=y0, choice=choice["label"]))
        log(f"{label}: {n} cells, {choice['label']}")
    torch.cuda.synchronize()

    def restore():
        for q in parts:
            _copy_device(q["ptr"].variables, q["y0"].data_ptr(), q["n_y"] * 8, stream)
        torch.cuda.synchronize()

    packed = _lib.PackedSolve([q["state"] for q in parts], _lib.SOLVE_COOPERATIVE)
    for _ in range(max(args.warmup, 1)):
        restore()
        packed.solve(dt, params)
    wall = 0.0
    for _ in range(args.steps):
        restore()
        t0 = time.perf_counter()
        packed.solve(dt, params)
        wall += time.perf_counter() - t0
    replayed = packed.replayed
    n_total = sum(q["state"].n_cells for q in parts)
    stats_all = [q["state"].results()[1] for q in parts]
    algo = sum(algorithmic_bytes(q["mech"].info, q["mech"].n_species, q["mech"].n_reactions, st)
               for q, st in zip(parts, stats_all))
    log(f"{args.steps} packed steps: {wall:.2f} s")
    restore()
    t0 = time.perf_counter()
    for q in parts:
        q["state"].solve(dt, params, _lib.SOLVE_COOPERATIVE)
    torch.cuda.synchronize()
    separate = time.perf_counter() - t0
    ms = wall * 1e3 / args.steps
    achieved = algo / (ms * 1e-3) / 1e9
    result = {
        "metric": "cell-solves/s, three mechanisms packed in one graph launch (20/50 x 200k, 60/200 x 50k, "
                  "210/560 x 20k cells, FP64)",
        "value": round(n_total * args.steps / wall, 2), "unit": "cell-solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded mechanisms and cells)",
        "config": {"workload": "packed: " + ", ".join(f"{q['label']} x {q['state'].n_cells} ({q['choice']})"
                                                      for q in parts) +
                               "; T 180-320 K, P 1-101 kPa, 50 % night, 4-stage Rosenbrock, 1 x 900 s",
                   "cells": n_total, "graph_replayed": replayed},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "the packed Solve (three cooperative launches in one graph), wall time "
                               f"of the graph launch; algorithmic bytes per step {algo}"},
        "attempts_per_step": int(sum(st[:, 2].sum() for st in stats_all)),
        "separate": {"value": round(n_total / separate, 2), "unit": "cell-solves/s",
                     "note": "the three solves one after another, no graph"},
    }
    if args.cpu_sample != 0:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        seconds = 0.0
        parts_cpu = []
        for q in parts:
            kind1, cps1, _, _ = cpu_baseline(q["desc"], q["cells"], q["order"], dt, 16, 1, tableau)
            sample = min(q["state"].n_cells, max(256, int(5.0 * cps1)))
            kind, cps, el, _ = cpu_baseline(q["desc"], q["cells"], q["order"], dt, sample, threads, tableau)
            seconds += q["state"].n_cells / cps
            parts_cpu.append(f"{q['label']}: {sample} cells {el:.1f} s ({cps:.0f} cell-solves/s)")
        result["cpu_baseline"] = {"value": round(n_total / seconds, 2), "unit": "cell-solves/s", "cores": threads,
                                  "kind": kind, "cpu": cpu_model(),
                                  "sample": "each mechanism's first cells, one 1-cell State per cell, combined as "
                                            "all cells over the summed times; " + "; ".join(parts_cpu)}
    print(json.dumps(result), flush=True)


ADJOINT_WORKLOADS = {
    # name: (mechanism, cells, Solves, time step, label)
    "adjoint": ("mechanism_60_200", 4096, 10, 300.0,
                "60/200 mechanism, 4096 cells (T 220-310 K, P 101325 Pa), 10 x 300 s, 4-stage, atol 1e-12"),
    "adjoint_210_560": ("mechanism_210_560", 1024, 1, 900.0,
                        "210/560 mechanism, 1024 cells of the packed workload's 210-species part (T 180-320 K, "
                        "P 1-101 kPa, 50 % night), 1 x 900 s, 4-stage, atol 1e-12"),
}
ADJOINT_TARGETS = (3, 10, 20, 33, 50)


def main_adjoint(args):
    """Gradients of loss = sum over cells of the squared final concentrations of five
    seeded species, with respect to every initial concentration and the log A of every
    reaction: a recorded Solve sequence (per-accepted-step checkpoints on the device) and
    one reverse sweep (micm_hip_adjoint).  `value` counts cell-solves with their gradients
    per second (forward and reverse).  atol 1e-12 throughout: the default 1e-3 exceeds
    every initial concentration of these workloads, and the map being differentiated is
    then not smooth at the scale of the difference checks."""
    import torch

    from micm_amd import _lib, synthetic
    from oracle import OraclePortableMath
    from oracle import Tape as HostTape
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    maker, n, n_solves, dt, label = ADJOINT_WORKLOADS[args.workload]
    desc = getattr(synthetic, maker)()
    if maker == "mechanism_60_200":
        cells = synthetic.inputs_60_200(desc, n)
    else:
        cells = synthetic.random_inputs(desc, n, 33, (180.0, 320.0), log_pressure=(1.0e3, 1.01e5),
                                        night_fraction=0.5)
    tableau = 4
    params = _lib.rosenbrock_parameters(tableau)
    mech = _lib.Mechanism(desc)
    order = mech.species_order()
    N, R = mech.n_species, mech.n_reactions
    y0 = np.ascontiguousarray(cells.variables[:, order])
    atol = np.full(N, 1e-12)
    state = _lib.DeviceState(mech, n)
    state.set_stream(stream.cuda_stream)
    state.set_variables(y0)
    state.set_conditions(cells.conditions)
    state.set_custom_parameters(cells.custom)
    state.set_tolerances(1e-6, atol)
    state.update_state_parameters()
    state.set_cooperative_shape(256, -1)
    max_records = 1002 * n_solves
    tape = _lib.Tape(mech, n, max_records)
    ptr = state.pointers()
    n_y = ptr.n_groups * 64 * N
    y_init = torch.empty(n_y, dtype=torch.float64, device="cuda")
    _copy_device(y_init.data_ptr(), ptr.variables, n_y * 8, stream)
    seed = torch.zeros(n_y, dtype=torch.float64, device="cuda")
    gy = torch.zeros_like(seed)
    gk = torch.zeros(ptr.n_groups * 64 * R, dtype=torch.float64, device="cuda")
    targets = [t for t in ADJOINT_TARGETS if t < N]

    def forward():
        _copy_device(ptr.variables, y_init.data_ptr(), n_y * 8, stream)
        tape.reset()
        for _ in range(n_solves):
            state.solve_recorded(dt, params, tape)

    def reverse():
        state.adjoint(params, tape, seed.data_ptr(), gy.data_ptr(), gk.data_ptr())

    forward()
    y_final = state.get_variables()
    lam = np.zeros((n, N))
    lam[:, targets] = 2.0 * y_final[:, targets]
    from micm_amd._lib import to_grouped
    seed.copy_(torch.from_numpy(to_grouped(lam)))
    reverse()
    log("warm-up forward and reverse done")
    fwd_ms, rev_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a, b, c = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        a.record(stream)
        forward()
        b.record(stream)
        reverse()
        c.record(stream)
        torch.cuda.synchronize()
        fwd_ms.append(a.elapsed_time(b))
        rev_ms.append(b.elapsed_time(c))
    wall = time.perf_counter() - t0
    log(f"{args.steps} timed forward+reverse: {wall:.2f} s")
    # the same Solves unrecorded: the cost of recording
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _copy_device(ptr.variables, y_init.data_ptr(), n_y * 8, stream)
    for _ in range(n_solves):
        state.solve(dt, params, _lib.SOLVE_COOPERATIVE)
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t1
    counts = tape.counts()
    status = state.results()[0]
    g_y = from_grouped_host(gy, n, N)
    g_k = from_grouped_host(gk, n, R)
    # gradient checks on sampled cells: directional central differences (h = 1e-6) of the
    # recorded map and the CPU tangent-linear model, three directions each in y0 and log A
    sample = np.linspace(0, n - 1, args.check_cells).astype(int)
    o = OraclePortableMath(desc, tableau)
    o.lib.oracle_set_threads(int(len(os.sched_getaffinity(0))))
    k = state.get_rate_constants()[sample]
    ht = HostTape(np.ascontiguousarray(tape.records(sample)), np.ascontiguousarray(counts[sample]))
    ys = y0[sample]
    rng = np.random.default_rng(17)
    h = 1e-6

    def loss(y):
        return float(np.sum(y[:, targets] ** 2))

    errs = {"fd_y0": [], "fd_logA": [], "fd_y0_h1e-4": [], "fd_logA_h1e-4": [], "tangent_y0": [],
            "tangent_logA": []}
    t2 = time.perf_counter()
    for _ in range(3):
        # unit directions (relative in y0), so that h is the size of the perturbation
        v = rng.standard_normal(ys.shape)
        v = v / np.linalg.norm(v) * ys
        w = rng.standard_normal(R)
        w = w / np.linalg.norm(w)
        ad_y = float(np.sum(g_y[sample] * v))
        ad_k = float(np.sum(g_k[sample] @ w))
        def central(step):
            return ((loss(o.replay(k, ys + step * v, ht)) - loss(o.replay(k, ys - step * v, ht))) / (2 * step),
                    (loss(o.replay(k * np.exp(step * w), ys, ht)) - loss(o.replay(k * np.exp(-step * w), ys, ht)))
                    / (2 * step))
        fd_y, fd_k = central(h)
        # at h = 1e-6 the differences' rounding (amplified through the recorded steps) dominates
        # their error; at 1e-4 their truncation error is still far smaller
        fd_y4, fd_k4 = central(1e-4)
        dyf = o.tangent(k, ys, ht, np.stack([v, np.zeros_like(v)], 1),
                        np.stack([np.zeros((len(sample), R)), np.broadcast_to(w, (len(sample), R))], 1))
        lam_s = lam[sample]
        tl_y = float(np.sum(dyf[:, 0] * lam_s))
        tl_k = float(np.sum(dyf[:, 1] * lam_s))
        # relative to the sum of the terms' magnitudes: a random direction's derivative can
        # cancel to far below its terms, and the differences' rounding scales with the terms
        sy = float(np.sum(np.abs(g_y[sample] * v)))
        sk = float(np.sum(np.abs(g_k[sample] * w)))
        errs["fd_y0"].append(abs(ad_y - fd_y) / sy)
        errs["fd_logA"].append(abs(ad_k - fd_k) / sk)
        errs["fd_y0_h1e-4"].append(abs(ad_y - fd_y4) / sy)
        errs["fd_logA_h1e-4"].append(abs(ad_k - fd_k4) / sk)
        errs["tangent_y0"].append(abs(ad_y - tl_y) / sy)
        errs["tangent_logA"].append(abs(ad_k - tl_k) / sk)
    log(f"gradient checks on {len(sample)} cells: {time.perf_counter() - t2:.1f} s")
    fwd, rev = float(np.mean(fwd_ms)), float(np.mean(rev_ms))
    rec_bytes = (N + 1) * 8
    result = {
        "metric": "cell-solves/s with gradients (recorded Solves + discrete adjoint, FP64)",
        "value": round(n * n_solves * args.steps / wall, 2), "unit": "cell-solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 1, "ms_per_step": round(wall * 1e3 / args.steps, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded mechanism and cells)",
        "config": {"workload": label, "cells": n, "solves": n_solves, "loss": "sum of squared final "
                   f"concentrations of species {targets} (state order)", "gradients": f"{N} initial concentrations "
                   f"and {R} log A per cell (summed over cells for a shared log A)"},
        "forward_recorded_ms": round(fwd, 2), "adjoint_ms": round(rev, 2),
        "forward_unrecorded_ms": round(plain_s * 1e3, 2),
        "memory": {"tape_allocated_bytes": int(n * max_records * rec_bytes + 8 * n),
                   "tape_used_bytes": int(counts.sum() * rec_bytes),
                   "records_per_cell": {"mean": round(float(counts.mean()), 1), "max": int(counts.max())},
                   "note": "one checkpoint (H, y) per accepted step and one per clamp; stage vectors recomputed"},
        "converged_fraction": round(float(np.mean(status == 2)), 4),
        "gradient_check": {"cells": int(len(sample)), "directions": 3, "h": h,
                           **{key: float(max(v)) for key, v in errs.items()},
                           "note": "largest error of a directional derivative of the loss over the sampled cells, "
                                   "relative to the sum of its terms' magnitudes: central differences of the recorded "
                                   "map (h = 1e-6 relative, and 1e-4: the error at 1e-6 is the differences' own "
                                   "rounding), and the CPU tangent-linear model (per-component checks of every "
                                   "gradient: tests/test_gpu_adjoint.py)"},
    }
    print(json.dumps(result), flush=True)


def from_grouped_host(t, n_cells, n_values):
    import torch
    from micm_amd._lib import from_grouped
    torch.cuda.synchronize()
    return from_grouped(t.cpu().numpy(), n_cells, n_values)


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
