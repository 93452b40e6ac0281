"""The oracle's discrete adjoint, pinned without the GPU.

The reference has no derivatives of Solve, so nothing in it pins these functions;
they are pinned by mathematics instead:

* the adjoint is the transpose of the tangent-linear model (TLM) -- two separate
  formulations in oracle/micm_oracle.cpp (the TLM differentiates the assembled -J and
  the forcing and solves with A, the adjoint goes through per-reaction weights and
  solves with A^T), so <lam, TLM v> = <adjoint(lam), v> for random lam, v fails on a
  dropped term, a wrong sign or a transposed factor in either;
* both are derivatives of the recorded map: central finite differences of `replay`
  (the map itself, each accepted step with its H frozen);
* the recorded map reproduces the Solve it was recorded from.
"""
import numpy as np
import pytest

from micm_amd import synthetic
from oracle import Oracle, Tape, set_adjoint_fault

TARGETS = [3, 10, 20, 33, 50]


def _setup(desc, inputs, n_cells, tableau, dt, n_solves, atol=None):
    o = Oracle(desc, tableau)
    cells = inputs(desc, n_cells)
    y0 = cells.variables[:, o.species_order()]
    k = o.rate_constants(cells.conditions, cells.custom)
    yf, states, stats, tape = o.solve_record(k, y0, dt, n_solves=n_solves, max_records=4096,
                                             atol=None if atol is None else np.full(o.n_species, atol))
    return o, y0, k, yf, states, tape


CASES = {
    "readme_3stage": (synthetic.readme_mechanism(), lambda d, n: synthetic.readme_inputs(n, seed=3), 3, 500.0, 3),
    "60_200_4stage": (synthetic.mechanism_60_200(), synthetic.inputs_60_200, 4, 300.0, 3),
    "60_200_2stage": (synthetic.mechanism_60_200(), synthetic.inputs_60_200, 2, 300.0, 2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_recorded_map_reproduces_the_solve(name):
    desc, inputs, tableau, dt, n_solves = CASES[name]
    o, y0, k, yf, states, tape = _setup(desc, inputs, 8, tableau, dt, n_solves)
    # a recorded Solve gives the same results as an unrecorded one
    cells = inputs(desc, 8)
    y_plain, states_plain, _, _ = o.solve(cells.conditions, cells.custom, y0, dt, n_solves)
    assert np.array_equal(yf, y_plain) and np.array_equal(states, states_plain)
    # every Solve ends in a clamp record; steps carry positive H
    recs = [tape.records[c, :tape.counts[c]] for c in range(8)]
    assert all((r[:, 0] == Tape.CLAMP).sum() == n_solves for r in recs)
    assert all(np.all((r[:, 0] > 0) | (r[:, 0] == Tape.CLAMP)) for r in recs)
    # replaying the frozen steps lands on the Solve's result (steps accepted after a
    # rejection factor alpha in one go instead of in increments: rounding differs)
    y_rep = o.replay(k, y0, tape)
    scale = np.maximum(np.abs(yf), 1e-20)
    assert np.max(np.abs(y_rep - yf) / scale) < 1e-8


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_is_the_transpose_of_the_tangent(name):
    desc, inputs, tableau, dt, n_solves = CASES[name]
    n = 6
    o, y0, k, yf, states, tape = _setup(desc, inputs, n, tableau, dt, n_solves)
    rng = np.random.default_rng(11)
    N, R = o.n_species, o.n_reactions
    lam = rng.standard_normal((n, N))
    gy, gk = o.adjoint(k, tape, lam)
    nd = 3
    dy0 = rng.standard_normal((n, nd, N)) * y0[:, None, :]
    dth = rng.standard_normal((n, nd, R))
    dyf = o.tangent(k, y0, tape, dy0, dth)
    lhs = np.einsum("cn,cdn->cd", lam, dyf)
    rhs = np.einsum("cn,cdn->cd", gy, dy0) + np.einsum("cr,cdr->cd", gk, dth)
    scale = np.einsum("cn,cdn->cd", np.abs(gy), np.abs(dy0)) + np.einsum("cr,cdr->cd", np.abs(gk), np.abs(dth))
    assert np.all(np.abs(lhs - rhs) <= 1e-9 * scale)


def _loss(y):
    return np.sum(y[:, TARGETS] ** 2, axis=1)


def test_adjoint_matches_central_differences():
    """Loss = sum of squared final concentrations of five species; central differences with
    h = 1e-6 relative in each initial concentration and each log k.  With an accurate
    solve (atol 1e-12) the map is smooth at that scale and every cell agrees to 1e-6
    normwise (the error is dominated by the differences' own rounding)."""
    desc = synthetic.mechanism_60_200()
    n = 4
    o, y0, k, yf, states, tape = _setup(desc, synthetic.inputs_60_200, n, 4, 300.0, 2, atol=1e-12)
    N, R = o.n_species, o.n_reactions
    y_rep = o.replay(k, y0, tape)
    lam = np.zeros((n, N))
    lam[:, TARGETS] = 2.0 * y_rep[:, TARGETS]
    gy, gk = o.adjoint(k, tape, lam)
    h = 1e-6
    s_fd = np.zeros((n, N))
    for i in range(N):
        yp, ym = y0.copy(), y0.copy()
        yp[:, i] *= 1 + h
        ym[:, i] *= 1 - h
        s_fd[:, i] = (_loss(o.replay(k, yp, tape)) - _loss(o.replay(k, ym, tape))) / (2 * h)
    s_adj = gy * y0
    assert np.all(np.abs(s_adj - s_fd).max(axis=1) <= 1e-6 * np.abs(s_fd).max(axis=1))
    g_fd = np.zeros((n, R))
    for r in range(R):
        kp, km = k.copy(), k.copy()
        kp[:, r] *= np.exp(h)
        km[:, r] *= np.exp(-h)
        g_fd[:, r] = (_loss(o.replay(kp, y0, tape)) - _loss(o.replay(km, y0, tape))) / (2 * h)
    assert np.all(np.abs(gk - g_fd).max(axis=1) <= 1e-6 * np.abs(g_fd).max(axis=1))


@pytest.mark.parametrize("fault", [1, 2, 3])
def test_planted_mistakes_fail_the_transpose_check(fault):
    """Plausible mistakes planted in the oracle's adjoint -- the Jacobian's dependence on y
    and k dropped, A solved instead of A^T, a sign flipped in the forcing's k-derivative --
    each fail the dot-product test."""
    desc = synthetic.mechanism_60_200()
    n = 3
    o, y0, k, yf, states, tape = _setup(desc, synthetic.inputs_60_200, n, 4, 300.0, 1)
    rng = np.random.default_rng(5)
    N, R = o.n_species, o.n_reactions
    lam = rng.standard_normal((n, N))
    dy0 = rng.standard_normal((n, 1, N)) * y0[:, None, :]
    dth = rng.standard_normal((n, 1, R))
    dyf = o.tangent(k, y0, tape, dy0, dth)
    lhs = np.einsum("cn,cdn->cd", lam, dyf)
    set_adjoint_fault(fault)
    try:
        gy, gk = o.adjoint(k, tape, lam)
    finally:
        set_adjoint_fault(0)
    rhs = np.einsum("cn,cdn->cd", gy, dy0) + np.einsum("cr,cdr->cd", gk, dth)
    scale = np.einsum("cn,cdn->cd", np.abs(gy), np.abs(dy0)) + np.einsum("cr,cdr->cd", np.abs(gk), np.abs(dth))
    assert np.any(np.abs(lhs - rhs) > 1e-6 * scale)


def test_overflowed_tape_is_refused():
    desc = synthetic.mechanism_60_200()
    o = Oracle(desc, 4)
    cells = synthetic.inputs_60_200(desc, 2)
    y0 = cells.variables[:, o.species_order()]
    k = o.rate_constants(cells.conditions, cells.custom)
    _, _, _, tape = o.solve_record(k, y0, 300.0, n_solves=1, max_records=3)
    assert np.all(tape.counts > 3)
    with pytest.raises(RuntimeError, match="overflowed"):
        o.adjoint(k, tape, np.ones_like(y0))
