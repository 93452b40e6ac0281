"""micm_hip_solve_noalloc: the cooperative Solve with its scratch in the library's static
device array instead of allocations.  Its results equal micm_hip_solve's bit for bit, on
two calls in a row, the second on a smaller State (the shared scratch keeps the first
call's contents)."""
import numpy as np
import pytest

from helpers import bits_equal
from micm_amd import _lib, synthetic

pytestmark = pytest.mark.gpu


def _state(mech, cells, shape, update=True):
    st = _lib.DeviceState(mech, cells.n_cells)
    st.set_variables(cells.variables[:, mech.species_order()])
    st.set_conditions(cells.conditions)
    st.set_custom_parameters(cells.custom)
    st.set_cooperative_shape(*shape)
    if update:
        st.update_state_parameters()
    return st


@pytest.mark.parametrize("shape", [(256, 4), (128, -1)])
def test_noalloc_solve_matches_solve(shape):
    import torch
    params = _lib.rosenbrock_parameters(4)
    runs = [(synthetic.mechanism_210_560(), synthetic.inputs_210_560, 900.0, 700),
            (synthetic.mechanism_60_200(), synthetic.inputs_60_200, 300.0, 130)]
    for desc, inputs, dt, n in runs:
        print(f"noalloc: {desc.n_species} species, {n} cells, shape {shape}", flush=True)
        mech = _lib.Mechanism(desc)
        cells = inputs(desc, n)
        ref = _state(mech, cells, shape)
        ref.solve(dt, params, _lib.SOLVE_COOPERATIVE)
        expected = (ref.get_variables(), *ref.results())
        for update in (False, True):
            st = _state(mech, cells, shape, update=not update)
            st.solve_noalloc(dt, params, update_rates=update)
            torch.cuda.synchronize()
            got = (st.get_variables(), *st.results())
            assert bits_equal(got[0], expected[0])
            assert np.array_equal(got[1], expected[1]) and np.array_equal(got[2], expected[2])
            assert bits_equal(got[3], expected[3])
