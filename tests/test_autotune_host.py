"""The autotuner's selection logic on the host, with the GPU timings replaced by a
model: the pilot prunes candidates at twice the best (the plain form at four times),
the per-lane persistent form without hand-off is skipped once a hand-off variant
was four times slower than the best, the hand-off variants always reach the last
(chip-filling) sample, and the fastest there wins."""
import types

import numpy as np
import pytest

from micm_amd import _lib, autotune


class _Mech:
    def __init__(self, tag, supported=True):
        self.info = types.SimpleNamespace(n_species=tag, n_reactions=1, nnz_jacobian=1, nnz_lower=1, nnz_upper=1,
                                          lu_multiply_adds=1)
        self._supported = supported

    def cooperative_supported(self, stages):
        return self._supported


class _State:
    def __init__(self, n):
        self.n_cells = n

    def get_variables(self):
        return np.zeros((self.n_cells, 2))

    def get_rate_constants(self):
        return np.zeros((self.n_cells, 1))


def _run(monkeypatch, model, n_cells, tag, supported=True):
    """Runs autotune with seconds = model[label](cells); returns (choice, calls)."""
    calls = []

    def fake(mech, y, k, dt, params, candidates, repeats, verbose, give_up=None, lu_blocking=0):
        out = {}
        for c in candidates:
            calls.append((c[0], y.shape[0]) if not lu_blocking else (f"{c[0]} lu{lu_blocking}", y.shape[0]))
            out[c[0]] = model[c[0]](y.shape[0])
        return out

    monkeypatch.setattr(autotune, "_time_candidates", fake)
    params = types.SimpleNamespace(stages=3)
    return autotune.autotune(_Mech(tag, supported), _State(n_cells), 30.0, params), calls


def _linear(t0, per_cell):
    return lambda n: t0 + per_cell * n


def test_plain_form_survives_the_pilot_and_wins_at_scale(monkeypatch):
    # a launch-bound plain form: slow on 256 cells, fastest on 10,000
    model = {label: _linear(0.008, 1.8e-5) for label, *_ in autotune.CANDIDATES}
    model["cooperative256"] = _linear(0.006, 1.2e-5)
    model["cooperative128"] = _linear(0.006, 1.55e-5)
    model["plain"] = _linear(0.026, 0.66e-5)
    model["persistent+handoff16"] = model["persistent+handoff48"] = _linear(0.2, 1e-4)
    choice, calls = _run(monkeypatch, model, 10000, tag=101)
    assert choice["label"] == "plain" and choice["form"] == _lib.SOLVE_PLAIN
    assert ("plain", 10000) in calls
    # the persistent variants four times slower than the best: the one without hand-off is skipped
    assert ("persistent", 256) not in calls


def test_plain_form_pruned_beyond_four_times(monkeypatch):
    model = {label: _linear(0.5, 1e-4) for label, *_ in autotune.CANDIDATES}
    model["cooperative256"] = _linear(0.4, 0.5e-4)
    model["plain"] = _linear(2.5, 1e-4)  # over four times the best pilot
    choice, calls = _run(monkeypatch, model, 100000, tag=102)
    assert choice["label"] == "cooperative256"
    assert [n for label, n in calls if label == "plain"] == [256]
    assert max(n for _, n in calls) == 65536


def test_persistent_hand_off_forms_are_timed_at_chip_filling_size(monkeypatch):
    """A per-lane form that loses the 256-cell pilot (its lanes idle behind the slowest
    cell) but wins once 65,536 cells fill the chip is the choice."""
    model = {label: _linear(0.5, 1e-4) for label, *_ in autotune.CANDIDATES}
    model["cooperative256"] = _linear(0.3, 8.5e-5)
    # slow on few cells, faster per cell at scale
    model["persistent+handoff16"] = lambda n: 1.4 if n <= 2048 else 0.2 + 8.2e-5 * n
    choice, calls = _run(monkeypatch, model, 100000, tag=104)
    assert ("persistent+handoff16", 65536) in calls and ("persistent+handoff48", 65536) in calls
    assert choice["label"] == "persistent+handoff16"


def test_without_cooperative_support_only_per_lane_forms_compete(monkeypatch):
    model = {label: _linear(0.1, 1e-5) for label, *_ in autotune.CANDIDATES}
    choice, calls = _run(monkeypatch, model, 5000, tag=103, supported=False)
    assert all(not label.startswith("cooperative") for label, _ in calls)
    assert choice["form"] in (_lib.SOLVE_PERSISTENT, _lib.SOLVE_PLAIN)


@pytest.fixture(autouse=True)
def _fresh_cache():
    autotune._cache.clear()
    yield
    autotune._cache.clear()


def test_plain_form_reaches_the_last_sample_within_twice_the_best(monkeypatch):
    """Fourth at 2,048 cells but within twice the best, the plain form is timed on the
    last sample, where its launches amortise over every cell and it wins (TS1's identical
    cells)."""
    model = {label: _linear(0.5, 1e-4) for label, *_ in autotune.CANDIDATES}
    model["cooperative256"] = _linear(0.004, 1.3e-5)
    model["cooperative128"] = _linear(0.004, 1.7e-5)
    model["cooperative512"] = _linear(0.004, 1.5e-5)
    model["plain"] = lambda n: 0.03 + 0.9e-5 * n if n > 256 else 0.024
    choice, calls = _run(monkeypatch, model, 10000, tag=105)
    assert ("plain", 10000) in calls
    assert choice["label"] == "plain"


def test_hand_off_goes_to_the_fastest_cooperative_shape(monkeypatch):
    """The persistent forms finish their slow cells in the cooperative form: they are
    timed, and chosen, with the shape of the fastest cooperative candidate, not with the
    one-wavefront default."""
    seen = {}
    model = {label: _linear(0.5, 1e-4) for label, *_ in autotune.CANDIDATES}
    model["cooperative256-compact-w2"] = _linear(0.3, 8.5e-5)
    model["persistent+handoff16"] = lambda n: 1.4 if n <= 2048 else 0.2 + 8.2e-5 * n

    def fake(mech, y, k, dt, params, candidates, repeats, verbose, give_up=None, lu_blocking=0):
        out = {}
        for c in candidates:
            seen.setdefault(c[0], []).append((y.shape[0], c[3], c[4]))
            out[c[0]] = model[c[0]](y.shape[0])
        return out

    monkeypatch.setattr(autotune, "_time_candidates", fake)
    choice = autotune.autotune(_Mech(106), _State(100000), 30.0, types.SimpleNamespace(stages=4))
    w2 = next(c for c in autotune.CANDIDATES if c[0] == "cooperative256-compact-w2")
    assert (65536, w2[3], w2[4]) in seen["persistent+handoff16"]
    assert choice["label"] == "persistent+handoff16" and choice["handoff_to"] == "cooperative256-compact-w2"
    assert (choice["threads"], choice["stage"]) == (w2[3], w2[4])


def test_lu_blocking_is_chosen_for_the_fastest_cooperative_shape(monkeypatch):
    """The LU schedule's blocking is timed on the first sample with the fastest cooperative
    shape of the last sample, confirmed on the last sample when 2 % faster than the
    default, and carried into the hand-off timings and the choice."""
    calls = []

    def fake(mech, y, k, dt, params, candidates, repeats, verbose, give_up=None, lu_blocking=0):
        out = {}
        for c in candidates:
            calls.append((c[0], y.shape[0], lu_blocking))
            t = 0.5 + 1e-4 * y.shape[0]
            if c[0] == "cooperative256":
                t = 0.3 + 8.0e-5 * y.shape[0] * {0: 1.0, 1: 1.05, 2: 0.99, 8: 0.9, -1: 1.3}[lu_blocking]
            out[c[0]] = t
        return out

    monkeypatch.setattr(autotune, "_time_candidates", fake)
    choice = autotune.autotune(_Mech(107), _State(100000), 30.0, types.SimpleNamespace(stages=4))
    assert choice["lu_blocking"] == 8 and choice["label"] == "cooperative256"
    assert {b for (label, n, b) in calls if label == "cooperative256" and n == 2048} == {0, 1, 2, 8, -1}
    assert ("cooperative256", 65536, 8) in calls
    assert all(b == 8 for (label, n, b) in calls if n == 65536 and label.startswith("persistent"))
    assert set(choice["lu_blocking_seconds"]) == {1, 2, 8, -1}


def test_wider_register_budget_wins_a_near_tie(monkeypatch):
    # the three-wave compact build 0.5 % faster than the two-wave one: the two-wave build is
    # taken (no spills); 2 % faster: the three-wave build is taken
    for margin, expected in ((0.995, "cooperative256-compact-w2"), (0.98, "cooperative256-compact-w3")):
        model = {label: _linear(0.5, 1e-3) for label, *_ in autotune.CANDIDATES}
        model["cooperative256-compact-w2"] = _linear(0.01, 1e-5)
        model["cooperative256-compact-w3"] = lambda n, m=margin: m * (0.01 + 1e-5 * n)
        choice, _ = _run(monkeypatch, model, 100000, tag=110 + int(margin * 1000))
        assert choice["label"] == expected, (margin, choice["label"])


def test_wider_build_reaches_the_last_sample_beside_its_twin(monkeypatch):
    # the two-wave build ranks outside the three finalists on the first sample, its
    # three-wave twin inside: both reach the last sample, where the three-wave build is
    # fastest by 0.5 % and the near tie goes to the two-wave build
    base = _linear(0.0099, 0.99e-5)
    model = {label: _linear(0.5, 1e-3) for label, *_ in autotune.CANDIDATES}
    model["cooperative256"] = lambda n: (1.02 if n > 2048 else 0.95) * base(n)
    model["cooperative512"] = lambda n: (1.03 if n > 2048 else 0.97) * base(n)
    model["cooperative256-compact-w3"] = base
    model["cooperative256-compact-w2"] = lambda n: (1.005 if n > 2048 else 1.5) * base(n)
    choice, calls = _run(monkeypatch, model, 100000, tag=130)
    assert ("cooperative256-compact-w2", 65536) in calls
    assert choice["label"] == "cooperative256-compact-w2"
