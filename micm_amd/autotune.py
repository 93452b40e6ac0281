"""Launch-shape autotuner for the per-cell forms.

The fastest way to run a mechanism depends on its size and on how unevenly its
cells behave: the per-lane persistent form (one cell per lane, with or without
handing slow cells to the cooperative form), the cooperative form (one
workgroup per cell) with 64, 128, 256 or 512 threads and with its constant tables
staged in LDS or read from global memory, or the plain form (a kernel per phase
over every cell, fastest when the cells take the same steps).  `autotune` times every candidate on a
strided sample of the state's own cells and returns the fastest.  Every
candidate produces bit-identical results, so the choice only affects speed.

Results are cached per (mechanism layout, tableau, time step, sample) in this
process.
"""
import sys
import time

import numpy as np

from . import _lib

# (label, form, handoff steps, cooperative threads, staging); the shapes that are
# usually fast come first, so that the pilot can give up on slow ones early
# the same shape built with the larger register budget
WIDER_BUILD = {"cooperative256-compact-w3": "cooperative256-compact-w2"}

CANDIDATES = (
    ("cooperative128", _lib.SOLVE_COOPERATIVE, -1, 128, -1),
    ("cooperative256", _lib.SOLVE_COOPERATIVE, -1, 256, -1),
    ("cooperative64", _lib.SOLVE_COOPERATIVE, -1, 64, -1),
    ("cooperative512", _lib.SOLVE_COOPERATIVE, -1, 512, -1),
    # the compact layout's register budgets: three waves per SIMD (168 registers) or two (256)
    ("cooperative256-compact-w3", _lib.SOLVE_COOPERATIVE, -1, 256, 5),
    ("cooperative256-compact-w2", _lib.SOLVE_COOPERATIVE, -1, 256, 6),
    ("cooperative64-global-tables", _lib.SOLVE_COOPERATIVE, -1, 64, 0),
    ("cooperative128-global-tables", _lib.SOLVE_COOPERATIVE, -1, 128, 0),
    ("persistent+handoff16", _lib.SOLVE_PERSISTENT, 16, 0, -1),
    ("persistent+handoff48", _lib.SOLVE_PERSISTENT, 48, 0, -1),
    ("persistent", _lib.SOLVE_PERSISTENT, 0, 0, -1),
    # a kernel per phase over every cell: fastest when all cells take the same steps
    ("plain", _lib.SOLVE_PLAIN, -1, 0, -1),
)

# LU schedule blockings tried besides the library's default (shortest segment of an
# element's ordered sum; -1 keeps sums whole): micm_hip_state_set_lu_blocking
LU_BLOCKINGS = (1, 2, 8, -1)

_cache = {}


def _key(mech, tableau_stages, dt, n):
    i = mech.info
    return (i.n_species, i.n_reactions, i.nnz_jacobian, i.nnz_lower, i.nnz_upper, i.lu_multiply_adds,
            tableau_stages, float(dt), int(n))


def apply(state, choice):
    """Configure `state` with a choice returned by `autotune`."""
    state.set_handoff(choice["handoff"])
    state.set_cooperative_shape(choice["threads"], choice["stage"])
    state.set_lu_blocking(choice.get("lu_blocking", 0))
    return choice["form"]


def _time_candidates(mech, y, k, dt, params, candidates, repeats, verbose, give_up=None, lu_blocking=0):
    """Seconds per Solve of each candidate on these cells (best of `repeats` after a
    warm-up run); a candidate whose warm-up run takes more than `give_up` seconds is
    not repeated and keeps that time."""
    probe = _lib.DeviceState(mech, y.shape[0])
    probe.set_rate_constants(k)
    probe.set_lu_blocking(lu_blocking)
    times = {}
    for label, form, handoff, threads, stage in candidates:
        probe.set_handoff(handoff)
        probe.set_cooperative_shape(threads, stage)
        el = []
        for r in range(repeats + 1):  # the first run warms up
            probe.set_variables(y)
            t0 = time.perf_counter()
            probe.solve(dt, params, form)
            el.append(time.perf_counter() - t0)
            if r == 0 and give_up is not None and el[0] > give_up:
                break
        times[label] = min(el[1:]) if len(el) > 1 else el[0]
        if verbose:
            tag = label + (f" (LU blocking {lu_blocking})" if lu_blocking else "")
            print(f"autotune: {tag:44s} {y.shape[0]:6d} cells {times[label] * 1e3:9.2f} ms", file=sys.stderr,
                  flush=True)
    return times


def autotune(mech, state, dt, params, sample=65536, first_sample=2048, pilot=256, finalists=3, repeats=1,
             candidates=CANDIDATES, verbose=False):
    """Pick the fastest per-cell form for `state` (whose inputs and rate constants
    are set), on strided samples of its cells: every candidate on `pilot` cells,
    those within twice the best (the plain form: four times) on `first_sample` cells,
    the `finalists` fastest of those, with the persistent forms that hand slow cells
    on, on `sample` cells; a persistent form hands its slow cells to the fastest
    cooperative shape of that sample.  The last sample fills the chip for the per-lane forms (64
    lanes x 4 SIMDs x 256 CUs = 65,536 cells): on fewer cells a per-lane form waits for
    its slowest cell with most lanes idle, and the cooperative shapes' time is that of
    their slowest cells, which favours the shapes that run one cell fastest over those
    that run the most cells at once.  Returns a dict with keys label, form, handoff,
    threads, stage (for a persistent form: the hand-off's shape), sample, seconds (label -> seconds on the largest sample the
    candidate reached), handoff_to (the cooperative candidate a persistent choice hands its
    slow cells to), lu_blocking (micm_hip_state_set_lu_blocking, for the fastest
    cooperative shape of the last sample; 0 keeps the default) and lu_blocking_seconds
    (first-sample times of that shape with each blocking)."""
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

    def per_lane(c):
        return c[1] == _lib.SOLVE_PERSISTENT and c[2] > 0

    def handing_to(c, pool):
        """`c` with the hand-off shape of the fastest cooperative candidate of `pool` timed
        so far (the cells a persistent form parks are finished by that shape)."""
        coop = [d for d in pool if d[1] == _lib.SOLVE_COOPERATIVE and d[0] in times]
        if not per_lane(c) or not coop:
            return c
        best_coop = min(coop, key=lambda d: times[d[0]])
        hand_to[c[0]] = best_coop[0]
        return (c[0], c[1], c[2], best_coop[3], best_coop[4])

    m0 = min(pilot, m)
    times = {}
    hand_to = {}
    lu_blocking = 0
    blocking_times = {}
    best = None
    slow_handoff = False
    for c in cands:  # pilot: each candidate gives up once it is four times slower than the best so far
        if c[1] == _lib.SOLVE_PERSISTENT and c[2] == 0 and slow_handoff:
            continue  # without hand-off the slowest cells run even longer one per lane
        c = handing_to(c, cands)
        t = _time_candidates(mech, *pick(m0), dt, params, [c], repeats, verbose,
                             give_up=None if best is None else 4.0 * best)[c[0]]
        times[c[0]] = t
        if c[1] == _lib.SOLVE_PERSISTENT and c[2] > 0 and best is not None and t > 4.0 * best:
            slow_handoff = True
        best = t if best is None else min(best, t)
    # the plain form's time grows slowly with the cell count (its launches serve every
    # cell at once), so a pilot understates it: it stays in at four times the best.  The
    # persistent forms with a hand-off are timed on the last sample whatever the pilot
    # says: a few hundred cells leave almost every lane idle behind the slowest one.
    lane_forms = [c for c in cands if per_lane(c)]
    cands = [c for c in cands
             if c[0] in times and not per_lane(c) and times[c[0]] <= (4.0 if c[1] == _lib.SOLVE_PLAIN else 2.0) * best]
    m1 = min(first_sample, m)
    if m1 > m0 and len(cands) > 1:
        times.update(_time_candidates(mech, *pick(m1), dt, params, cands, repeats, verbose))
        ranked = sorted(cands, key=lambda c: times[c[0]])
        best1 = times[ranked[0][0]]
        # the plain form gains the most from more cells: it stays in at twice the best; a
        # finalist's build with the larger register budget stays in beside it (the last pick
        # prefers it on a near tie)
        top = ranked[:finalists]
        wider = [c for c in ranked[finalists:] if c[0] in {WIDER_BUILD.get(f[0]) for f in top}]
        cands = top + wider + [c for c in ranked[finalists:]
                               if c[1] == _lib.SOLVE_PLAIN and times[c[0]] <= 2.0 * best1 and c not in wider]
    first_times = dict(times)
    if m > m1:
        times.update(_time_candidates(mech, *pick(m), dt, params, cands, repeats, verbose))
    # the LU schedule's blocking, for the fastest cooperative shape of the last sample, timed
    # on the first sample; a blocking that is 2 % faster there is timed on the last sample too
    # and taken if it is faster there as well
    coop = sorted((c for c in cands if c[1] == _lib.SOLVE_COOPERATIVE and c[0] in times), key=lambda c: times[c[0]])
    if coop and m1 > m0 and coop[0][0] in first_times:
        top = coop[0]
        base = first_times[top[0]]
        best_t, best_b = base, 0
        for b in LU_BLOCKINGS:
            t = _time_candidates(mech, *pick(m1), dt, params, [top], repeats, verbose, lu_blocking=b)[top[0]]
            blocking_times[b] = t
            if t < 0.98 * base and t < best_t:
                best_t, best_b = t, b
        if best_b != 0:
            if m > m1:
                t = _time_candidates(mech, *pick(m), dt, params, [top], repeats, verbose, lu_blocking=best_b)[top[0]]
                if t < times[top[0]]:
                    times[top[0]], lu_blocking = t, best_b
            else:
                times[top[0]], lu_blocking = best_t, best_b
    # the hand-off goes to the fastest cooperative shape of the largest sample (with the LU
    # blocking chosen for it)
    lane_forms = [handing_to(c, cands) for c in lane_forms]
    if m > m0 and lane_forms:
        times.update(_time_candidates(mech, *pick(m), dt, params, lane_forms, repeats, verbose,
                                      lu_blocking=lu_blocking))
    last = cands + lane_forms
    cands = [c for c in last if c[0] in times]
    win = min(cands, key=lambda c: times[c[0]])
    # a build bounded to fewer registers wins only by more than 1 %: its extra waves per
    # SIMD help only where LDS holds that many cells per CU, and its spills cost HBM traffic
    wider = {c[0]: c for c in cands}.get(WIDER_BUILD.get(win[0]))
    if wider is not None and times[wider[0]] <= 1.01 * times[win[0]]:
        win = wider
    label, form, handoff, threads, stage = win
    choice = {"label": label, "form": form, "handoff": handoff, "threads": threads, "stage": stage,
              "handoff_to": hand_to.get(label) if handoff > 0 else None, "lu_blocking": lu_blocking,
              "lu_blocking_seconds": blocking_times, "sample": m, "seconds": times}
    _cache[key] = dict(choice)
    return choice
