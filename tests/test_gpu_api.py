"""The reference's README example (README.md:103-166) written against the Python
mirror of MICM's API, solved on the GPU, reproducing the README's printed table
(README.md:171-181) -- the drop-in check of the builder / State /
UpdateStateParameters / Solve interface."""
import numpy as np
import pytest

from micm_amd import (ArrheniusRateConstantParameters, ChemicalReactionBuilder, HipSolverBuilder, Phase,
                      RosenbrockSolverParameters, SolverState, Species, StoichSpecies, System)

pytestmark = pytest.mark.gpu

README_TABLE = [
    (11.8, 5.90, 1.91), (6.79, 9.05, 3.32), (3.83, 10.7, 4.21), (2.14, 11.7, 4.74), (1.19, 12.2, 5.04),
    (0.658, 12.4, 5.21), (0.364, 12.6, 5.31), (0.201, 12.7, 5.36), (0.111, 12.7, 5.39), (0.0613, 12.8, 5.41),
]


def _readme_solver():
    foo, bar, baz = Species("Foo"), Species("Bar"), Species("Baz")
    gas = Phase("gas", [foo, bar, baz])
    r1 = (ChemicalReactionBuilder().SetReactants([foo])
          .SetProducts([StoichSpecies(bar, 0.8), StoichSpecies(baz, 0.2)])
          .SetRateConstant(ArrheniusRateConstantParameters(A=1.0e-3)).SetPhase(gas).Build())
    r2 = (ChemicalReactionBuilder().SetReactants([foo, bar]).SetProducts([StoichSpecies(baz, 1)])
          .SetRateConstant(ArrheniusRateConstantParameters(A=1.0e-5, C=110.0)).SetPhase(gas).Build())
    solver = (HipSolverBuilder(RosenbrockSolverParameters.ThreeStageRosenbrockParameters())
              .SetSystem(System(gas)).SetReactions([r1, r2]).Build())
    return solver, (foo, bar, baz)


@pytest.mark.parametrize("form", ["lockstep", "persistent", "cooperative", "auto"])
def test_readme_example_reproduces_the_printed_table(form):
    print(f"readme example, form={form}", flush=True)
    solver, (foo, bar, baz) = _readme_solver()
    state = solver.GetState(4, form=form)
    for c in range(4):
        state.conditions[c].temperature = 287.45
        state.conditions[c].pressure = 101319.9
        state.conditions[c].CalculateIdealAirDensity()
    state[foo] = 20.0
    for row in README_TABLE:
        solver.UpdateStateParameters(state)
        result = solver.Solve(500.0, state)
        assert result.state == SolverState.Converged
        for c in range(4):
            got = [float(f"{state[s][c]:.2e}") for s in (foo, bar, baz)]
            assert got == [float(f"{v:.2e}") for v in row]
