"""Packed solves: states of three mechanisms (20/50, 60/200 and TS1) solved by one
graph-captured launch give, on the capture and on every replay, exactly the
results of solving each state on its own.  Between replays the program does its
own eager work (resetting concentrations, reading results)."""
import numpy as np
import pytest

from helpers import bits_equal
from mechanisms import ts1_inputs, ts1_mechanism
from micm_amd import _lib, synthetic

pytestmark = pytest.mark.gpu


def _states():
    out = []
    for desc, inputs, n in ((synthetic.mechanism_20_50(), synthetic.inputs_60_200, 700),
                            (synthetic.mechanism_60_200(), synthetic.inputs_60_200, 300),
                            (ts1_mechanism(), ts1_inputs, 90)):
        mech = _lib.Mechanism(desc)
        cells = inputs(desc, n)
        y0 = cells.variables[:, mech.species_order()]
        st = _lib.DeviceState(mech, n)
        st.set_variables(y0)
        st.set_conditions(cells.conditions)
        st.set_custom_parameters(cells.custom)
        st.update_state_parameters()
        out.append((mech, st, y0))
    return out


@pytest.mark.parametrize("form", [_lib.SOLVE_COOPERATIVE, _lib.SOLVE_PERSISTENT])
def test_packed_solve_matches_separate_solves(form):
    print(f"packed: form={form}", flush=True)
    params = _lib.rosenbrock_parameters(4)
    dt = 300.0
    states = _states()
    expected = []
    for mech, st, y0 in states:
        st.set_handoff(4)
        st.solve(dt, params, form)
        expected.append((st.get_variables(), *st.results()))
        st.set_variables(y0)
    packed = _lib.PackedSolve([st for _, st, _ in states], form)
    for replay in range(4):
        packed.solve(dt, params)
        # the states were solved once above, so they run in a cost order from the first
        # packed Solve on: captured then, replayed after
        assert packed.replayed == (replay > 0)
        for (mech, st, y0), (y, status, stats, ft) in zip(states, expected):
            s2, stats2, ft2 = st.results()
            assert bits_equal(st.get_variables(), y)
            assert np.array_equal(s2, status) and np.array_equal(stats2, stats) and bits_equal(ft2, ft)
            st.set_variables(y0)
    # a different time step captures a new graph
    packed.solve(dt / 2, params)
    assert not packed.replayed


def test_packed_recaptures_when_rtol_or_scratch_change():
    """A changed relative tolerance (a kernel argument of the captured launches) and a
    scratch buffer reallocated by a direct Solve of another launch shape both force a new
    capture: results stay those of separate solves."""
    params = _lib.rosenbrock_parameters(4)
    dt = 300.0
    desc = synthetic.mechanism_20_50()
    mech = _lib.Mechanism(desc)
    cells = synthetic.inputs_60_200(desc, 500)
    y0 = cells.variables[:, mech.species_order()]

    def fresh(rtol, shape):
        st = _lib.DeviceState(mech, 500)
        st.set_variables(y0)
        st.set_conditions(cells.conditions)
        st.set_custom_parameters(cells.custom)
        st.set_tolerances(rtol)
        st.set_cooperative_shape(*shape)
        st.update_state_parameters()
        return st

    ref = {}
    for rtol in (1e-6, 1e-3):
        st = fresh(rtol, (256, 4))
        st.solve(dt, params, _lib.SOLVE_COOPERATIVE)
        ref[rtol] = st.get_variables()
    assert not bits_equal(ref[1e-6], ref[1e-3])
    st = fresh(1e-6, (256, 4))
    packed = _lib.PackedSolve([st], _lib.SOLVE_COOPERATIVE)
    packed.solve(dt, params)
    assert bits_equal(st.get_variables(), ref[1e-6])
    st.set_variables(y0)
    st.set_tolerances(1e-3)
    packed.solve(dt, params)
    assert not packed.replayed
    assert bits_equal(st.get_variables(), ref[1e-3])
    # another shape's direct Solve may reallocate the compact layout's scratch; back on the
    # captured shape the graph must not replay a freed buffer
    st.set_variables(y0)
    st.set_cooperative_shape(128, 4)
    st.solve(dt, params, _lib.SOLVE_COOPERATIVE)
    assert bits_equal(st.get_variables(), ref[1e-3])
    st.set_cooperative_shape(256, 4)
    st.set_variables(y0)
    packed.solve(dt, params)
    assert bits_equal(st.get_variables(), ref[1e-3])


PACKED_CONFIG = (("20/50", "mechanism_20_50", 200_000), ("60/200", "mechanism_60_200", 50_000),
                 ("210/560", "mechanism_210_560", 20_000))


def test_packed_workload_matches_the_reference_arithmetic():
    """The packed workload at full size (20/50 x 200k, 60/200 x 50k, 210/560 x 20k cells; T
    180-320 K, P log-uniform 1-101 kPa, 50 % night, 4-stage, 900 s) in one graph, captured
    and replayed: 64 evenly spaced cells of each mechanism equal the C-library oracle (the
    reference's own arithmetic) bit for bit -- outcome, counters and concentrations."""
    from oracle import Oracle
    params = _lib.rosenbrock_parameters(4)
    dt = 900.0
    parts = []
    for i, (label, maker, n) in enumerate(PACKED_CONFIG):
        desc = getattr(synthetic, maker)()
        cells = synthetic.random_inputs(desc, n, 31 + i, (180.0, 320.0), log_pressure=(1.0e3, 1.01e5),
                                        night_fraction=0.5)
        mech = _lib.Mechanism(desc)
        y0 = cells.variables[:, mech.species_order()]
        st = _lib.DeviceState(mech, n)
        st.set_variables(y0)
        st.set_conditions(cells.conditions)
        st.set_custom_parameters(cells.custom)
        st.set_cooperative_shape(256, 4)
        st.update_state_parameters()
        parts.append((label, desc, cells, mech, y0, st))
    packed = _lib.PackedSolve([p[-1] for p in parts], _lib.SOLVE_COOPERATIVE)
    for replay in range(3):
        print(f"packed workload: solve {replay}", flush=True)
        packed.solve(dt, params)
        assert packed.replayed == (replay > 1)
        for label, desc, cells, mech, y0, st in parts:
            sample = np.linspace(0, cells.n_cells - 1, 64).astype(int)
            y_ref, st_ref, stats_ref, _ = Oracle(desc, 4).solve(cells.conditions[sample], cells.custom[sample],
                                                                y0[sample], dt)
            status, stats, _ = st.results()
            y = st.get_variables()
            print(f"  {label}: converged {np.mean(status == 2):.3f} of {cells.n_cells}", flush=True)
            assert np.array_equal(status[sample], st_ref) and np.array_equal(stats[sample], stats_ref)
            assert bits_equal(y[sample], y_ref)
            st.set_variables(y0)
