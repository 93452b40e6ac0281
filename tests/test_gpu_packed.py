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
    for replay in range(3):
        packed.solve(dt, params)
        assert packed.replayed == (replay > 0)
        for (mech, st, y0), (y, status, stats, ft) in zip(states, expected):
            s2, stats2, ft2 = st.results()
            assert bits_equal(st.get_variables(), y)
            assert np.array_equal(s2, status) and np.array_equal(stats2, stats) and bits_equal(ft2, ft)
            st.set_variables(y0)
    # a different time step captures a new graph
    packed.solve(dt / 2, params)
    assert not packed.replayed
