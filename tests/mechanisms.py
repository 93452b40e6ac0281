"""Mechanisms used by the tests besides the synthetic ones."""
import os

import numpy as np

from micm_amd import mechanism_io, synthetic
from micm_amd._desc import (ARRHENIUS, BRANCHED, REVERSIBLE, SURFACE, TAYLOR, TERNARY, TROE, TUNNELING,
                            USER_DEFINED, MechanismDesc)


def all_rate_types():
    """Eight species, one reaction of every rate-constant type, a third body M
    (c_rho = 1) on two reactions and a second parameterized species, declared in
    an order that the type sort has to rearrange."""
    M = -1          # third body
    Q = -2          # parameterized: 0.5 + 1e-3 T
    rx = [
        ([0], [(1, 1.0)], USER_DEFINED, [2.5]),
        ([1, 2], [(3, 0.7), (4, 0.3)], ARRHENIUS, [3.2e3, 1.1, -250.0, 280.0, 2.0e-6]),
        ([2, M], [(5, 1.0)], TROE, [6.5e4, -2.6, 40.0, 3.1e6, 0.3, -15.0, 0.55, 1.2]),
        ([3, 4], [(6, 1.0)], TERNARY, [1.2e5, -1.5, 0.0, 7.0e6, -0.4, 0.0, 0.6, 1.0]),
        ([4], [(7, 1.0)], BRANCHED, [0.0, 1.1e-2, 350.0, 0.15, 6]),
        ([5], [(7, 1.0)], BRANCHED, [1.0, 2.3e-2, 120.0, 0.4, 9]),
        ([6], [(0, 0.5)], TUNNELING, [4.0e-3, 150.0, 1.0e6]),
        ([7, Q], [(1, 1.0)], TAYLOR, [1.0e-3, 0.5, -30.0, 300.0, 1.0e-7, 3, 1.0, 1.0e-3, -2.0e-6]),
        ([0, M], [(2, 1.0)], REVERSIBLE, [2.0e-5, 800.0, 3.0e-2]),
        ([3], [(5, 1.0)], SURFACE, [1.5e-5, 0.046, 0.3]),
        ([1, 1], [(6, 1.0)], ARRHENIUS, [5.0e2, 0.0, 100.0, 300.0, 0.0]),
    ]
    return MechanismDesc(8, rx, parameterized=[(0.0, 0.0, 0.0, 1.0), (0.5, 1.0e-3, 0.0, 0.0)])


def all_types_inputs(desc, n, seed=5):
    """Cells for `all_rate_types`: T 200-320 K, P 5-110 kPa, log-uniform custom parameters
    (photolysis 1e-6..0.1 s-1, surface radius 1e-8..1e-5 m and number density 1e6..1e10)
    and concentrations 1e-8..1."""
    st = synthetic.CellStreams(seed, n)
    T = st.uniform(200.0, 320.0)
    P = st.uniform(5e3, 1.1e5)
    cond = np.stack([T, P, P / (synthetic.GAS_CONSTANT * T)], 1)
    custom = np.stack([st.log_uniform(1e-6, 1e-1), st.log_uniform(1e-8, 1e-5), st.log_uniform(1e6, 1e10)], 1)
    y = np.stack([st.log_uniform(1e-8, 1.0) for _ in range(desc.n_species)], 1)
    return synthetic.CellInputs(cond, y, custom, seed)


TS1_JSON = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ts1_mechanism.json")


def ts1_mechanism():
    """The reference's TS1 benchmark mechanism (test/benchmark/ts1_mechanism.hpp:
    209 species plus the third body M, 547 reactions), as exported from the
    reference's own header into tests/golden/ts1_mechanism.json."""
    return mechanism_io.load(TS1_JSON)


def ts1_inputs(desc, n, seed=13):
    """The reference benchmark's cell (bench::Ts1::InitState) in cell 0; the other
    cells vary its temperature, pressure and custom parameters."""
    cond0, y, custom = mechanism_io.initial_arrays(desc, n)
    st = synthetic.CellStreams(seed, n)
    first = np.arange(n) == 0
    T = np.where(first, cond0[:, 0], st.uniform(220.0, 310.0))
    P = np.where(first, cond0[:, 1], st.log_uniform(2e4, 1.01e5))
    scale = np.where(first, 1.0, st.log_uniform(0.1, 3.0))
    cond = np.stack([T, P, P / (synthetic.GAS_CONSTANT * T)], 1)
    cond[0] = cond0[0]
    return synthetic.CellInputs(cond, y, custom * scale[:, None], seed)
