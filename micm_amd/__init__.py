"""micm_amd -- MI355X-native Rosenbrock chemistry solver with MICM's API.

Host-side mirror of the reference's System / Process / SolverBuilder / State /
UpdateStateParameters / Solve interface (include/micm/solver/solver_builder.hpp,
include/micm/solver/solver.hpp, include/micm/solver/state.hpp) in front of the
C-ABI of include/micm_hip.h.  `HipSolverBuilder` plays the role of the
reference's `CpuSolverBuilder` / `CudaSolverBuilderInPlace`: the same calls,
with every kernel of the path running as hand-written HIP code on the GPU.

    foo, bar, baz = Species("Foo"), Species("Bar"), Species("Baz")
    gas = Phase("gas", [foo, bar, baz])
    r1 = (ChemicalReactionBuilder().SetReactants([foo])
          .SetProducts([StoichSpecies(bar, 0.8), StoichSpecies(baz, 0.2)])
          .SetRateConstant(ArrheniusRateConstantParameters(A=1.0e-3)).SetPhase(gas).Build())
    solver = (HipSolverBuilder(RosenbrockSolverParameters.ThreeStageRosenbrockParameters())
              .SetSystem(System(gas)).SetReactions([r1]).Build())
    state = solver.GetState(1)
    state.conditions[0].temperature = 287.45
    ...
    solver.UpdateStateParameters(state)
    result = solver.Solve(500.0, state)
"""
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Optional

import numpy as np

from . import _desc
from ._desc import MechanismDesc
from ._lib import (SOLVE_COOPERATIVE, SOLVE_LOCKSTEP, SOLVE_PERSISTENT, SOLVE_PLAIN, DeviceState, Mechanism,
                   MicmHipError, rosenbrock_parameters)

__all__ = [
    "Species", "Phase", "System", "StoichSpecies", "Process", "ChemicalReactionBuilder",
    "ArrheniusRateConstantParameters", "TroeRateConstantParameters",
    "TernaryChemicalActivationRateConstantParameters", "BranchedRateConstantParameters",
    "TunnelingRateConstantParameters", "TaylorSeriesRateConstantParameters", "ReversibleRateConstantParameters",
    "UserDefinedRateConstantParameters", "SurfaceRateConstantParameters", "RosenbrockSolverParameters",
    "HipSolverBuilder", "Solver", "State", "SolverState", "SolverStats", "SolverResult", "MicmHipError",
    "MechanismDesc",
]


# ---------------------------------------------------------------- domain model (include/micm/system)

class Species:
    """A chemical species (system/species.hpp).  A parameterized species (e.g. a third
    body M, SetThirdBody) is not a state variable; its concentration is
    c0 + c_T*T + c_P*P + c_rho*rho (util/parameterized_function.hpp:36)."""

    def __init__(self, name, **properties):
        self.name = name
        self.properties = dict(properties)
        self.parameterize = None

    def SetThirdBody(self):
        self.parameterize = (0.0, 0.0, 0.0, 1.0)
        return self

    def SetParameterization(self, c0=0.0, c_T=0.0, c_P=0.0, c_rho=0.0):
        self.parameterize = (float(c0), float(c_T), float(c_P), float(c_rho))
        return self

    def IsParameterized(self):
        return self.parameterize is not None

    def SetProperty(self, key, value):
        self.properties[key] = value
        return self

    def HasProperty(self, key):
        return key in self.properties

    def GetProperty(self, key):
        return self.properties[key]

    def __repr__(self):
        return f"Species({self.name!r})"


class Phase:
    def __init__(self, name, species):
        self.name = name
        self.species = list(species)

    def StateSize(self):
        return sum(1 for s in self.species if not s.IsParameterized())

    def SpeciesNames(self):
        return [s.name for s in self.species if not s.IsParameterized()]


class System:
    def __init__(self, gas_phase):
        self.gas_phase = gas_phase

    def StateSize(self):
        return self.gas_phase.StateSize()

    def UniqueNames(self):
        return self.gas_phase.SpeciesNames()


@dataclass
class StoichSpecies:
    species: Species
    coefficient: float = 1.0


# ---------------------------------------------------------------- rate constants (process/rate_constant)

@dataclass
class ArrheniusRateConstantParameters:
    A: float = 1.0
    B: float = 0.0
    C: float = 0.0
    D: float = 300.0
    E: float = 0.0
    _type = _desc.ARRHENIUS

    def params(self):
        return [self.A, self.B, self.C, self.D, self.E]


@dataclass
class TroeRateConstantParameters:
    k0_A: float = 1.0
    k0_B: float = 0.0
    k0_C: float = 0.0
    kinf_A: float = 1.0
    kinf_B: float = 0.0
    kinf_C: float = 0.0
    Fc: float = 0.6
    N: float = 1.0
    _type = _desc.TROE

    def params(self):
        return [self.k0_A, self.k0_B, self.k0_C, self.kinf_A, self.kinf_B, self.kinf_C, self.Fc, self.N]


@dataclass
class TernaryChemicalActivationRateConstantParameters(TroeRateConstantParameters):
    _type = _desc.TERNARY


@dataclass
class BranchedRateConstantParameters:
    branch: str = "Alkoxy"  # or "Nitrate"
    X: float = 0.0
    Y: float = 0.0
    a0: float = 0.0
    n: int = 0
    _type = _desc.BRANCHED

    def params(self):
        return [0.0 if self.branch == "Alkoxy" else 1.0, self.X, self.Y, self.a0, float(int(self.n))]


@dataclass
class TunnelingRateConstantParameters:
    A: float = 1.0
    B: float = 0.0
    C: float = 0.0
    _type = _desc.TUNNELING

    def params(self):
        return [self.A, self.B, self.C]


@dataclass
class TaylorSeriesRateConstantParameters:
    A: float = 1.0
    B: float = 0.0
    C: float = 0.0
    D: float = 300.0
    E: float = 0.0
    coefficients: List[float] = field(default_factory=lambda: [1.0])
    _type = _desc.TAYLOR

    def params(self):
        c = list(self.coefficients)
        if len(c) > 16:
            raise ValueError("at most 16 Taylor coefficients")
        return [self.A, self.B, self.C, self.D, self.E, float(len(c))] + c


@dataclass
class ReversibleRateConstantParameters:
    A: float = 1.0
    C: float = 0.0
    k_r: float = 0.0
    _type = _desc.REVERSIBLE

    def params(self):
        return [self.A, self.C, self.k_r]


@dataclass
class UserDefinedRateConstantParameters:
    label: str = ""
    scaling_factor: float = 1.0
    _type = _desc.USER_DEFINED

    def params(self):
        return [self.scaling_factor]


@dataclass
class SurfaceRateConstantParameters:
    """species must carry 'diffusion coefficient [m2 s-1]' and 'molecular weight [kg mol-1]'."""
    label: str = ""
    species: Optional[Species] = None
    reaction_probability: float = 1.0
    _type = _desc.SURFACE

    def params(self):
        s = self.species
        return [s.GetProperty("diffusion coefficient [m2 s-1]"), s.GetProperty("molecular weight [kg mol-1]"),
                self.reaction_probability]


# ---------------------------------------------------------------- processes

class Process:
    def __init__(self, reactants, products, rate_constant, phase=None):
        self.reactants = list(reactants)
        self.products = list(products)
        self.rate_constant = rate_constant
        self.phase = phase


class ChemicalReactionBuilder:
    """process/chemical_reaction_builder.hpp:21-88"""

    def __init__(self):
        self._reactants, self._products, self._rate, self._phase = [], [], None, None

    def SetReactants(self, reactants):
        self._reactants = list(reactants)
        return self

    def SetProducts(self, products):
        self._products = [p if isinstance(p, StoichSpecies) else StoichSpecies(p, 1.0) for p in products]
        return self

    def SetRateConstant(self, rate):
        self._rate = rate
        return self

    def SetPhase(self, phase):
        self._phase = phase
        return self

    def Build(self):
        if self._rate is None:
            raise ValueError("reaction has no rate constant")
        return Process(self._reactants, self._products, self._rate, self._phase)


# ---------------------------------------------------------------- solver parameters and results

class RosenbrockSolverParameters:
    """micm::RosenbrockSolverParameters (solver/rosenbrock_solver_parameters.hpp:22-67)."""

    def __init__(self, preset):
        self._c = rosenbrock_parameters(preset)

    def __getattr__(self, name):
        c = object.__getattribute__(self, "_c")
        return getattr(c, name)

    def __setattr__(self, name, value):
        if name == "_c":
            object.__setattr__(self, name, value)
        else:
            setattr(self._c, name, value)

    @staticmethod
    def TwoStageRosenbrockParameters():
        return RosenbrockSolverParameters(2)

    @staticmethod
    def ThreeStageRosenbrockParameters():
        return RosenbrockSolverParameters(3)

    @staticmethod
    def FourStageRosenbrockParameters():
        return RosenbrockSolverParameters(4)

    @staticmethod
    def FourStageDifferentialAlgebraicRosenbrockParameters():
        return RosenbrockSolverParameters("4da")

    @staticmethod
    def SixStageDifferentialAlgebraicRosenbrockParameters():
        return RosenbrockSolverParameters("6da")


class SolverState(IntEnum):
    """solver/solver_result.hpp:13-35"""
    NotYetCalled = 0
    Running = 1
    Converged = 2
    ConvergenceExceededMaxSteps = 3
    StepSizeTooSmall = 4
    RepeatedlySingularMatrix = 5
    NaNDetected = 6
    InfDetected = 7


@dataclass
class SolverStats:
    """Sums over the state's cells of the per-cell counters (solver_result.hpp:37-57)."""
    function_calls: int = 0
    jacobian_updates: int = 0
    number_of_steps: int = 0
    accepted: int = 0
    rejected: int = 0
    decompositions: int = 0
    solves: int = 0
    final_time: float = 0.0


@dataclass
class SolverResult:
    """state: the worst per-cell outcome (Converged only if every cell converged);
    per-cell outcomes are in State.cell_state / cell_stats / cell_final_time."""
    state: SolverState = SolverState.NotYetCalled
    stats: SolverStats = field(default_factory=SolverStats)


# ---------------------------------------------------------------- builder, solver, state

_FORMS = {"lockstep": SOLVE_LOCKSTEP, "plain": SOLVE_PLAIN, "persistent": SOLVE_PERSISTENT,
          "cooperative": SOLVE_COOPERATIVE, "auto": None}


class HipSolverBuilder:
    """SolverBuilder (solver_builder.hpp:37-217) whose policies are the HIP kernels."""

    def __init__(self, parameters):
        self.parameters = parameters
        self.system = None
        self.reactions = []
        self.reorder_state = True
        self.ignore_unused_species = True

    def SetSystem(self, system):
        self.system = system
        return self

    def SetReactions(self, reactions):
        self.reactions = list(reactions)
        return self

    def SetReorderState(self, reorder):
        self.reorder_state = bool(reorder)
        return self

    def SetIgnoreUnusedSpecies(self, ignore):
        self.ignore_unused_species = bool(ignore)
        return self

    def description(self):
        """The flat description handed to the C-ABI."""
        if self.system is None:
            raise ValueError("Missing chemical system.")
        state_species = [s for s in self.system.gas_phase.species if not s.IsParameterized()]
        param_species = [s for s in self.system.gas_phase.species if s.IsParameterized()]
        index = {s.name: i for i, s in enumerate(state_species)}
        pindex = {s.name: i for i, s in enumerate(param_species)}

        def sid(s):
            if s.IsParameterized():
                if s.name not in pindex:
                    pindex[s.name] = len(param_species)
                    param_species.append(s)
                return -(pindex[s.name] + 1)
            if s.name not in index:
                raise ValueError(f"species '{s.name}' is not in the system")
            return index[s.name]

        rx = []
        labels = []
        for p in self.reactions:
            reactants = [sid(s) for s in p.reactants]
            products = [(sid(ps.species), ps.coefficient) for ps in p.products]
            rx.append((reactants, products, p.rate_constant._type, p.rate_constant.params()))
            labels.append(getattr(p.rate_constant, "label", None))
        atol = [s.GetProperty("absolute tolerance") if s.HasProperty("absolute tolerance") else 1e-3
                for s in state_species]
        desc = MechanismDesc(len(state_species), rx, [ps.parameterize for ps in param_species], atol,
                             self.reorder_state, self.ignore_unused_species, [s.name for s in state_species])
        return desc, labels

    def Build(self):
        desc, labels = self.description()
        return Solver(desc, labels, self.parameters)


class _Conditions:
    def __init__(self, arr, i):
        self._a, self._i = arr, i

    temperature = property(lambda s: s._a[s._i, 0], lambda s, v: s._a.__setitem__((s._i, 0), v))
    pressure = property(lambda s: s._a[s._i, 1], lambda s, v: s._a.__setitem__((s._i, 1), v))
    air_density = property(lambda s: s._a[s._i, 2], lambda s, v: s._a.__setitem__((s._i, 2), v))

    def CalculateIdealAirDensity(self):
        """conditions.hpp:20-23"""
        R = 1.380649e-23 * 6.02214076e23
        self.air_density = self.pressure / (R * self.temperature)


class _ConditionsView:
    def __init__(self, arr):
        self._a = arr

    def __getitem__(self, i):
        return _Conditions(self._a, i)

    def __len__(self):
        return self._a.shape[0]


class State:
    """Host mirror of one HIP state (state.hpp:47-416 / cuda_state.hpp:15-149).

    variables [n_cells, n_species] (state order), conditions [n_cells, 3],
    custom_rate_parameters [n_cells, n_custom]; inputs are copied to the device
    by UpdateStateParameters and Solve, results copied back after Solve."""

    def __init__(self, solver, n_cells, form):
        self.solver = solver
        self.number_of_grid_cells = n_cells
        self.form = form
        m = solver.mechanism
        self.variable_names = [solver.desc.species_names[i] for i in m.species_order()]
        self.variable_map = {n: i for i, n in enumerate(self.variable_names)}
        self.custom_rate_parameter_map = dict(solver.custom_parameter_map)
        self.variables = np.zeros((n_cells, m.n_species))
        self.conditions_array = np.zeros((n_cells, 3))
        self.custom_rate_parameters = np.zeros((n_cells, m.n_custom))
        self.relative_tolerance = 1e-6
        self.absolute_tolerance = None
        self.device = DeviceState(m, n_cells)
        self.cell_state = np.zeros(n_cells, dtype=np.int32)
        self.cell_stats = np.zeros((n_cells, 7), dtype=np.int64)
        self.cell_final_time = np.zeros(n_cells)

    @property
    def conditions(self):
        return _ConditionsView(self.conditions_array)

    def __setitem__(self, species, value):
        name = species.name if isinstance(species, Species) else species
        self.variables[:, self.variable_map[name]] = value

    def __getitem__(self, species):
        name = species.name if isinstance(species, Species) else species
        return self.variables[:, self.variable_map[name]]

    def SetConcentration(self, species, values):
        self[species] = values

    def SetCustomRateParameter(self, label, values):
        self.custom_rate_parameters[:, self.custom_rate_parameter_map[label]] = values

    def SetRelativeTolerance(self, rtol):
        self.relative_tolerance = float(rtol)

    def SetAbsoluteTolerances(self, atol):
        self.absolute_tolerance = np.asarray(atol, dtype=np.float64)

    def SyncInputsToDevice(self):
        self.device.set_variables(self.variables)
        self.device.set_conditions(self.conditions_array)
        self.device.set_custom_parameters(self.custom_rate_parameters)
        self.device.set_tolerances(self.relative_tolerance, self.absolute_tolerance)

    def SyncOutputsToHost(self):
        self.variables[:] = self.device.get_variables()
        self.cell_state, self.cell_stats, self.cell_final_time = self.device.results()


class Solver:
    """solver.hpp:22-298 with the HIP kernels behind every policy."""

    def __init__(self, desc, labels, parameters):
        self.desc = desc
        self.parameters = parameters
        self.mechanism = Mechanism(desc)
        self.custom_parameter_map = {}
        order = self.mechanism.reaction_order()
        p = 0
        for r in order:
            t = int(desc.rate_types[r])
            if t == _desc.USER_DEFINED:
                self.custom_parameter_map[labels[r]] = p
                p += 1
            elif t == _desc.SURFACE:
                self.custom_parameter_map[labels[r] + ".effective radius [m]"] = p
                self.custom_parameter_map[labels[r] + ".particle number concentration [# m-3]"] = p + 1
                p += 2

    def GetState(self, number_of_grid_cells=1, form="lockstep"):
        """form: "lockstep" (default) integrates all cells with one step size, as one
        reference State of that many cells does; "plain", "persistent",
        "cooperative" and "auto" (autotuned at the first Solve) give every cell its
        own step control, as one reference State per cell does.  The per-cell
        forms return bit-identical results."""
        return State(self, int(number_of_grid_cells), _FORMS[form])

    def UpdateStateParameters(self, state):
        state.SyncInputsToDevice()
        state.device.update_state_parameters()

    def Solve(self, time_step, state, parameters=None):
        if parameters is not None:
            self.parameters = parameters
        state.device.set_variables(state.variables)
        if state.form is None:
            from . import autotune
            choice = autotune.autotune(self.mechanism, state.device, time_step, self.parameters._c)
            state.form = autotune.apply(state.device, choice)
        state.device.solve(time_step, self.parameters._c, state.form)
        state.SyncOutputsToHost()
        stats = SolverStats()
        for i, f in enumerate(("function_calls", "jacobian_updates", "number_of_steps", "accepted", "rejected",
                               "decompositions", "solves")):
            setattr(stats, f, int(state.cell_stats[:, i].sum()))
        stats.final_time = float(state.cell_final_time.min()) if state.number_of_grid_cells else 0.0
        worst = SolverState.Converged
        for s in np.unique(state.cell_state):
            if int(s) != SolverState.Converged:
                worst = SolverState(int(s))
        return SolverResult(worst, stats)
