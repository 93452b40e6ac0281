"""Launch-shape autotuner for the per-cell forms.

The fastest way to run a mechanism depends on its size and on how unevenly its
cells behave: the per-lane persistent form (one cell per lane, with or without
handing slow cells to the cooperative form), or the cooperative form (one
workgroup per cell) with 64, 128, 256 or 512 threads and with its constant tables
staged in LDS or read from global memory.  `autotune` times every candidate on a
strided sample of the state's own cells and returns the fastest.  Every
candidate produces bit-identical results, so the choice only affects speed.

Results are cached per (mechanism layout, tableau, time step, sample) in this
process.
"""
import sys
import time

import numpy as np

from . import _lib

# (label, form, handoff steps, cooperative threads, staging)
CANDIDATES = (
    ("persistent", _lib.SOLVE_PERSISTENT, 0, 0, -1),
    ("persistent+handoff16", _lib.SOLVE_PERSISTENT, 16, 0, -1),
    ("persistent+handoff48", _lib.SOLVE_PERSISTENT, 48, 0, -1),
    ("cooperative64", _lib.SOLVE_COOPERATIVE, -1, 64, -1),
    ("cooperative128", _lib.SOLVE_COOPERATIVE, -1, 128, -1),
    ("cooperative256", _lib.SOLVE_COOPERATIVE, -1, 256, -1),
    ("cooperative512", _lib.SOLVE_COOPERATIVE, -1, 512, -1),
    ("cooperative64-global-tables", _lib.SOLVE_COOPERATIVE, -1, 64, 0),
    ("cooperative128-global-tables", _lib.SOLVE_COOPERATIVE, -1, 128, 0),
)

_cache = {}


def _key(mech, tableau_stages, dt, n):
    i = mech.info
    return (i.n_species, i.n_reactions, i.nnz_jacobian, i.nnz_lower, i.nnz_upper, i.lu_multiply_adds,
            tableau_stages, float(dt), int(n))


def apply(state, choice):
    """Configure `state` with a choice returned by `autotune`."""
    state.set_handoff(choice["handoff"])
    state.set_cooperative_shape(choice["threads"], choice["stage"])
    return choice["form"]


def _time_candidates(mech, y, k, dt, params, candidates, repeats, verbose):
    probe = _lib.DeviceState(mech, y.shape[0])
    probe.set_rate_constants(k)
    times = {}
    for label, form, handoff, threads, stage in candidates:
        probe.set_handoff(handoff)
        probe.set_cooperative_shape(threads, stage)
        el = []
        for _ in range(repeats + 1):  # the first run warms up
            probe.set_variables(y)
            t0 = time.perf_counter()
            probe.solve(dt, params, form)
            el.append(time.perf_counter() - t0)
        times[label] = min(el[1:])
        if verbose:
            print(f"autotune: {label:32s} {y.shape[0]:6d} cells {times[label] * 1e3:9.2f} ms", file=sys.stderr,
                  flush=True)
    return times


def autotune(mech, state, dt, params, sample=2048, first_sample=512, finalists=3, repeats=1, candidates=CANDIDATES,
             verbose=False):
    """Pick the fastest per-cell form for `state` (whose inputs and rate constants
    are set): every candidate is timed on a strided sample of `first_sample` cells,
    the `finalists` fastest again on `sample` cells.  Returns a dict with keys label,
    form, handoff, threads, stage, sample, seconds (label -> seconds on the larger
    sample, or on the first one for candidates that did not go on)."""
    n = state.n_cells
    m = min(sample, n)
    key = _key(mech, params.stages, dt, m)
    if key in _cache:
        return dict(_cache[key])
    supported = mech.cooperative_supported(params.stages)
    cands = [c for c in candidates if c[1] != _lib.SOLVE_COOPERATIVE or supported]
    y_all = state.get_variables()
    k_all = state.get_rate_constants()

    def pick(count):
        idx = np.linspace(0, n - 1, count).round().astype(np.int64)
        return y_all[idx], k_all[idx]

    m1 = min(first_sample, m)
    times = _time_candidates(mech, *pick(m1), dt, params, cands, repeats, verbose)
    if m > m1:
        best = sorted(cands, key=lambda c: times[c[0]])[:finalists]
        times.update(_time_candidates(mech, *pick(m), dt, params, best, repeats, verbose))
        cands = best
    label, form, handoff, threads, stage = min(cands, key=lambda c: times[c[0]])
    choice = {"label": label, "form": form, "handoff": handoff, "threads": threads, "stage": stage,
              "sample": m, "seconds": times}
    _cache[key] = dict(choice)
    return choice
