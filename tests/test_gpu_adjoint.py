"""GPU discrete adjoint through the C-ABI (micm_hip_solve_recorded, micm_hip_adjoint).

* A recorded Solve is the ordinary per-cell Solve (bit for bit against the oracle) and
  its checkpoints are the oracle's, bit for bit.
* The adjoint on the GPU tape agrees with the oracle's adjoint (tests/test_adjoint_oracle.py
  pins that one against its tangent-linear model and finite differences).
* The differentiation workload: the 60-species mechanism over 4,096 cells, 10 Solves of
  300 s, loss = sum of squared final concentrations of five species; gradients with
  respect to every initial concentration and the log A of every reaction, checked on 16
  sampled cells against central finite differences (h = 1e-6 relative) and the CPU
  tangent-linear model.
"""
import numpy as np
import pytest

from helpers import bits_equal, dev, dev_zeros, host
from mechanisms import all_rate_types, all_types_inputs, ts1_inputs, ts1_mechanism
from micm_amd import _lib, synthetic
from oracle import OraclePortableMath, Tape

pytestmark = pytest.mark.gpu

TARGETS = [3, 10, 20, 33, 50]


CASES = {
    "readme": (synthetic.readme_mechanism(), lambda d, n: synthetic.readme_inputs(n, seed=3), 3, 500.0, 3),
    "all_types": (all_rate_types(), all_types_inputs, 4, 60.0, 2),
    "60_200": (synthetic.mechanism_60_200(), synthetic.inputs_60_200, 4, 300.0, 2),
    "60_200_2stage": (synthetic.mechanism_60_200(), synthetic.inputs_60_200, 2, 300.0, 1),
    "210_560": (synthetic.mechanism_210_560(), synthetic.inputs_210_560, 4, 900.0, 1),
    "ts1": (ts1_mechanism(), ts1_inputs, 3, 30.0, 2),
}


def _record(mech, cells, tableau, dt, n_solves, max_records, atol=None, coop_shape=None):
    n = cells.n_cells
    order = mech.species_order()
    state = _lib.DeviceState(mech, n)
    state.set_variables(cells.variables[:, order])
    state.set_conditions(cells.conditions)
    state.set_custom_parameters(cells.custom)
    if atol is not None:
        state.set_tolerances(1e-6, np.full(mech.n_species, atol))
    if coop_shape:
        state.set_cooperative_shape(*coop_shape)
    state.update_state_parameters()
    tape = _lib.Tape(mech, n, max_records)
    params = _lib.rosenbrock_parameters(tableau)
    for _ in range(n_solves):
        state.solve_recorded(dt, params, tape)
    return state, tape, params


def _adjoint(state, tape, params, lam):
    n, N, R = state.n_cells, state.mechanism.n_species, state.mechanism.n_reactions
    seed = dev(lam)
    gy, gk = dev_zeros(n, N), dev_zeros(n, R)
    state.adjoint(params, tape, seed.data_ptr(), gy.data_ptr(), gk.data_ptr())
    return host(gy, n, N), host(gk, n, R)


def _smooth_cells(tape, dt):
    """Cells none of whose recorded steps is below h_min (1e-15 dt, rosenbrock.inl:33)."""
    counts = tape.counts()
    out = np.zeros(tape.n_cells, dtype=bool)
    for c in range(tape.n_cells):
        h = tape.records([c])[0, :counts[c], 0]
        out[c] = not np.any((h > 0) & (h < 1e-15 * dt))
    return out


def _host_tape(tape, cells_idx=None):
    cnt = tape.counts()
    if cells_idx is not None:
        cnt = cnt[cells_idx]
    return Tape(np.ascontiguousarray(tape.records(cells_idx)), np.ascontiguousarray(cnt))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("coop_shape", [None, (256, 4)])
def test_recorded_solve_and_checkpoints_bitwise(name, coop_shape):
    desc, inputs, tableau, dt, n_solves = CASES[name]
    mech = _lib.Mechanism(desc)
    if not mech.cooperative_supported(tableau):
        pytest.skip("mechanism does not fit the cooperative form")
    n = 40 if name in ("210_560", "ts1") else 130
    print(f"recorded solve: {name} shape={coop_shape} n={n}", flush=True)
    cells = inputs(desc, n)
    state, tape, params = _record(mech, cells, tableau, dt, n_solves, 1024, coop_shape=coop_shape)
    o = OraclePortableMath(desc, tableau)
    y0 = cells.variables[:, mech.species_order()]
    k = o.rate_constants(cells.conditions, cells.custom)
    y_ref, st_ref, stats_ref, tape_ref = o.solve_record(k, y0, dt, n_solves=n_solves, max_records=1024)
    st, stats, _ = state.results()
    assert bits_equal(state.get_variables(), y_ref)
    assert np.array_equal(st, st_ref)
    counts, records = tape.counts(), tape.records()
    assert np.array_equal(counts, tape_ref.counts)
    for c in range(n):
        m = min(counts[c], 1024)
        assert bits_equal(records[c, :m], tape_ref.records[c, :m]), c


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_matches_oracle(name):
    desc, inputs, tableau, dt, n_solves = CASES[name]
    mech = _lib.Mechanism(desc)
    if not mech.cooperative_supported(tableau):
        pytest.skip("mechanism does not fit the cooperative form")
    n = 24 if name in ("210_560", "ts1") else 70
    print(f"adjoint vs oracle: {name} n={n}", flush=True)
    cells = inputs(desc, n)
    # room for every accepted step a Solve can take (max_number_of_steps) and its clamp
    state, tape, params = _record(mech, cells, tableau, dt, n_solves, 1002 * n_solves)
    assert tape.overflowed() == 0
    rng = np.random.default_rng(3)
    lam = rng.standard_normal((n, mech.n_species))
    gy, gk = _adjoint(state, tape, params, lam)
    o = OraclePortableMath(desc, tableau)
    k = state.get_rate_constants()
    gy_ref, gk_ref = o.adjoint(k, _host_tape(tape), lam)
    y0 = cells.variables[:, mech.species_order()]
    # relative sensitivities (d/dlog y0) and d/dlog k, normwise per cell.  A cell whose
    # step size fell below h_min had steps accepted whatever their error (rosenbrock.inl:
    # 229: `H < h_min`); its map is unstable (sensitivities of 1e30 and more, overflowing
    # on some TS1 cells, where the oracle's own tangent and adjoint part ways), so such
    # cells are left out.
    smooth = _smooth_cells(tape, dt)
    assert smooth.mean() >= 0.75
    for a, b in ((gy * y0, gy_ref * y0), (gk, gk_ref)):
        assert np.all(np.isfinite(a[smooth]))
        scale = np.maximum(np.abs(b).max(axis=1), 1e-300)
        err = np.abs(a - b).max(axis=1) / scale
        print(f"  max normwise difference {err[smooth].max():.2e} ({smooth.sum()} cells without forced steps), "
              f"{err.max():.2e} (all)", flush=True)
        assert np.all(err[smooth] <= 1e-9)


def test_overflowed_tape_gives_nan():
    desc = synthetic.mechanism_60_200()
    mech = _lib.Mechanism(desc)
    cells = synthetic.inputs_60_200(desc, 70)
    state, tape, params = _record(mech, cells, 4, 300.0, 1, 4)
    counts = tape.counts()
    assert np.all(counts > 4) and tape.overflowed() == 70
    gy, gk = _adjoint(state, tape, params, np.ones((70, mech.n_species)))
    assert np.all(np.isnan(gy)) and np.all(np.isnan(gk))


def _loss(y):
    return np.sum(y[:, TARGETS] ** 2, axis=1)


def test_differentiation_workload():
    """60/200, 4,096 cells, 10 Solves of 300 s, 4-stage; atol 1e-12 so that the solution
    (and so the map being differentiated) is accurate -- the default atol of 1e-3 exceeds
    every initial concentration of this workload (1e-12 .. 1e-3), its solutions go
    negative and are clamped, and central differences at h = 1e-6 stop resolving the
    derivative on a fifth of the cells."""
    desc = synthetic.mechanism_60_200()
    mech = _lib.Mechanism(desc)
    n, n_solves, dt = 4096, 10, 300.0
    cells = synthetic.inputs_60_200(desc, n)
    print("differentiation workload: recording", flush=True)
    state, tape, params = _record(mech, cells, 4, dt, n_solves, 2048, atol=1e-12)
    assert tape.overflowed() == 0
    st, stats, _ = state.results()
    y_final = state.get_variables()
    N, R = mech.n_species, mech.n_reactions
    lam = np.zeros((n, N))
    lam[:, TARGETS] = 2.0 * y_final[:, TARGETS]
    print("differentiation workload: adjoint", flush=True)
    gy, gk = _adjoint(state, tape, params, lam)
    assert np.all(np.isfinite(gy)) and np.all(np.isfinite(gk))
    sample = np.linspace(0, n - 1, 16).astype(int)
    o = OraclePortableMath(desc, 4)
    y0 = cells.variables[:, mech.species_order()][sample]
    k = state.get_rate_constants()[sample]
    tp = _host_tape(tape, sample)
    # the recorded map reproduces the GPU's final state
    y_rep = o.replay(k, y0, tp)
    assert np.max(np.abs(y_rep - y_final[sample]) / np.maximum(np.abs(y_final[sample]), 1e-30)) < 1e-8
    lam_s = lam[sample]
    s_adj = gy[sample] * y0
    g_adj = gk[sample]
    h = 1e-6
    print("differentiation workload: central differences", flush=True)
    s_fd = np.zeros_like(s_adj)
    for i in range(N):
        yp, ym = y0.copy(), y0.copy()
        yp[:, i] *= 1 + h
        ym[:, i] *= 1 - h
        s_fd[:, i] = (_loss(o.replay(k, yp, tp)) - _loss(o.replay(k, ym, tp))) / (2 * h)
    err_y = np.abs(s_adj - s_fd).max(axis=1) / np.abs(s_fd).max(axis=1)
    g_fd = np.zeros(R)
    for r in range(R):
        kp, km = k.copy(), k.copy()
        kp[:, r] *= np.exp(h)
        km[:, r] *= np.exp(-h)
        g_fd[r] = (_loss(o.replay(kp, y0, tp)).sum() - _loss(o.replay(km, y0, tp)).sum()) / (2 * h)
    err_k = np.abs(g_adj.sum(axis=0) - g_fd).max() / np.abs(g_fd).max()
    print(f"  FD: y0 normwise max {err_y.max():.2e}, log A normwise {err_k:.2e}", flush=True)
    assert np.all(err_y <= 1e-6) and err_k <= 1e-6
    print("differentiation workload: tangent-linear", flush=True)
    eye = np.zeros((16, N + R, N))
    dth = np.zeros((16, N + R, R))
    for i in range(N):
        eye[:, i, i] = y0[:, i]
    for r in range(R):
        dth[:, N + r, r] = 1.0
    dyf = o.tangent(k, y0, tp, eye, dth)
    tl = np.einsum("cdn,cn->cd", dyf, lam_s)
    err_tl_y = np.abs(tl[:, :N] - s_adj).max(axis=1) / np.abs(tl[:, :N]).max(axis=1)
    err_tl_k = np.abs(tl[:, N:] - g_adj).max(axis=1) / np.abs(tl[:, N:]).max(axis=1)
    print(f"  tangent-linear: y0 {err_tl_y.max():.2e}, log A {err_tl_k.max():.2e}", flush=True)
    assert np.all(err_tl_y <= 1e-9) and np.all(err_tl_k <= 1e-9)
