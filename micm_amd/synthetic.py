"""Seeded synthetic mechanisms and grid-cell inputs.

Random numbers come from xoshiro256** seeded through SplitMix64 (Blackman &
Vigna, "Scrambled linear pseudorandom number generators", 2021), so the same
seed gives the same mechanism and inputs on every platform.  A uniform draw is
u = (x >> 11) * 2^-53 and a log-uniform draw on [a, b] is
exp(ln a + u * ln(b / a)).

Per-cell inputs use one independent stream per cell (stream c is seeded with
seed + (c + 1) * 0x9E3779B97F4A7C15) and are drawn vectorised across cells.

Units follow the reference: concentrations in mol m-3, rate constants for a
reaction with n reactants in (m3 mol-1)^(n-1) s-1.  Arrhenius pre-exponential
factors specified per molecule (m3 molecule-1 s-1) are multiplied by
N_A^(n-1); unimolecular Arrhenius reactions therefore keep A in 1e-20..1e-16
s-1 and are effectively inert.
"""
import math

import numpy as np

from ._desc import ARRHENIUS, TROE, TERNARY, USER_DEFINED, MechanismDesc

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
AVOGADRO = 6.02214076e23
BOLTZMANN = 1.380649e-23
GAS_CONSTANT = BOLTZMANN * AVOGADRO  # include/micm/util/constants.hpp:9-11


def splitmix64(x):
    """Returns (new_state, output)."""
    x = (x + GOLDEN) & M64
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return x, z ^ (z >> 31)


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M64


class Xoshiro256ss:
    """Scalar xoshiro256** stream."""

    def __init__(self, seed):
        x = int(seed) & M64
        self.s = []
        for _ in range(4):
            x, z = splitmix64(x)
            self.s.append(z)

    def next_u64(self):
        s = self.s
        result = (_rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self, a=0.0, b=1.0):
        u = (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)
        return a + u * (b - a)

    def log_uniform(self, a, b):
        return math.exp(math.log(a) + self.uniform() * math.log(b / a))

    def integer(self, n):
        """Uniform integer in [0, n)."""
        return min(int(self.uniform() * n), n - 1)


class CellStreams:
    """One xoshiro256** stream per cell, advanced together (numpy uint64 arithmetic wraps)."""

    def __init__(self, seed, n_cells):
        cells = np.arange(n_cells, dtype=np.uint64)
        with np.errstate(over="ignore"):
            x = np.uint64(int(seed) & M64) + (cells + np.uint64(1)) * np.uint64(GOLDEN)
        self.s = np.empty((4, n_cells), dtype=np.uint64)
        for i in range(4):
            x, z = self._splitmix(x)
            self.s[i] = z

    @staticmethod
    def _splitmix(x):
        with np.errstate(over="ignore"):
            x = x + np.uint64(GOLDEN)
            z = x
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x, z ^ (z >> np.uint64(31))

    @staticmethod
    def _rotl(x, k):
        return (x << np.uint64(k)) | (x >> np.uint64(64 - k))

    def next_u64(self):
        s = self.s
        with np.errstate(over="ignore"):
            result = self._rotl(s[1] * np.uint64(5), 7) * np.uint64(9)
        t = s[1] << np.uint64(17)
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def uniform(self, a=0.0, b=1.0):
        u = (self.next_u64() >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return a + u * (b - a)

    def log_uniform(self, a, b):
        return np.exp(math.log(a) + self.uniform() * math.log(b / a))


# ------------------------------------------------------------------ mechanisms

def readme_mechanism():
    """The reference README system (README.md:95-161): Foo -> 0.8 Bar + 0.2 Baz (A=1e-3);
    Foo + Bar -> Baz (A=1e-5, C=110).  Species Foo, Bar, Baz."""
    reactions = [
        ([0], [(1, 0.8), (2, 0.2)], ARRHENIUS, [1.0e-3, 0.0, 0.0, 300.0, 0.0]),
        ([0, 1], [(2, 1.0)], ARRHENIUS, [1.0e-5, 0.0, 110.0, 300.0, 0.0]),
    ]
    return MechanismDesc(3, reactions, species_names=["Foo", "Bar", "Baz"])


def random_mechanism(n_species, n_reactions, mix, seed, products_from="above"):
    """Seeded random mechanism.

    mix: dict rate_type -> exact count (sums to n_reactions).  Every reaction
    has 1-2 reactants drawn uniformly with replacement from species
    0..n_species-2 (user-defined reactions, first-order photolysis-like rates,
    have exactly one), and 1-3 products with yields uniform in [0.1, 2] drawn
    uniformly from the species above the highest reactant, so the reactant ->
    product graph is acyclic and the last species is a terminal product.  (With
    products drawn from all species, 60-80 % of cells of the 60- and
    210-species mechanisms end in NaN or exceed the step limit in the reference
    solver, because yields up to 2 make most reaction cycles net producers.)
    products_from="all" draws them from every species instead: the configurations'
    text taken literally.

      Arrhenius: A log-uniform in 1e-20..1e-16 m3 molecule-1 s-1 times
                 N_A^(n_reactants-1); C uniform in [-2000, 500] K; B=0, D=300, E=0.
      Troe / ternary chemical activation: k0_A log-uniform in 1e-33..1e-28
                 cm6 molecule-2 s-1 and kinf_A log-uniform in 1e-12..1e-10
                 cm3 molecule-1 s-1, converted to mol/m3 units; k0_B uniform in
                 [-5, -1], kinf_B uniform in [-1.5, 0.5], C=0, Fc=0.6, N=1.
      User-defined: scaling factor 1; the per-cell rate is a custom parameter.
    """
    rng = Xoshiro256ss(seed)
    types = []
    for t, n in sorted(mix.items()):
        types.extend([t] * n)
    if len(types) != n_reactions:
        raise ValueError("mix counts must sum to n_reactions")
    for i in range(len(types) - 1, 0, -1):  # Fisher-Yates
        j = rng.integer(i + 1)
        types[i], types[j] = types[j], types[i]
    reactions = []
    for t in types:
        n_react = 1 if t == USER_DEFINED else 1 + rng.integer(2)
        reactants = [rng.integer(n_species - 1) for _ in range(n_react)]
        lowest = max(reactants) + 1 if products_from == "above" else 0
        n_prod = 1 + rng.integer(3)
        products = [(lowest + rng.integer(n_species - lowest), rng.uniform(0.1, 2.0)) for _ in range(n_prod)]
        if t == ARRHENIUS:
            a = rng.log_uniform(1e-20, 1e-16) * AVOGADRO ** (n_react - 1)
            c = rng.uniform(-2000.0, 500.0)
            params = [a, 0.0, c, 300.0, 0.0]
        elif t in (TROE, TERNARY):
            k0_a = rng.log_uniform(1e-33, 1e-28) * AVOGADRO ** 2 * 1e-12
            k0_b = rng.uniform(-5.0, -1.0)
            kinf_a = rng.log_uniform(1e-12, 1e-10) * AVOGADRO * 1e-6
            kinf_b = rng.uniform(-1.5, 0.5)
            params = [k0_a, k0_b, 0.0, kinf_a, kinf_b, 0.0, 0.6, 1.0]
        elif t == USER_DEFINED:
            params = [1.0]
        else:
            raise ValueError("unsupported rate type in random mix")
        reactions.append((reactants, products, t, params))
    desc = MechanismDesc(n_species, reactions)
    desc.seed = seed
    return desc


def mechanism_20_50(seed=2020):
    return random_mechanism(20, 50, {ARRHENIUS: 40, TROE: 5, USER_DEFINED: 5}, seed)


def mechanism_60_200(seed=6020):
    """60 species / 200 reactions: 80 % Arrhenius, 10 % Troe, 10 % user-defined."""
    return random_mechanism(60, 200, {ARRHENIUS: 160, TROE: 20, USER_DEFINED: 20}, seed)


def mechanism_210_560(seed=21056, products_from="above"):
    """210 species / 560 reactions: 75 % Arrhenius, 10 % Troe, 5 % ternary, 10 % photolysis."""
    return random_mechanism(210, 560, {ARRHENIUS: 420, TROE: 56, TERNARY: 28, USER_DEFINED: 56}, seed,
                            products_from)


# ------------------------------------------------------------------ cell inputs

class CellInputs:
    """Per-cell inputs in declared species order, cell-major.

    conditions [n_cells, 3]: temperature [K], pressure [Pa], air density [mol m-3]
    variables  [n_cells, n_species]
    custom     [n_cells, n_custom] (the mechanism's custom-parameter order)
    """

    def __init__(self, conditions, variables, custom, seed=None, night=None):
        self.conditions = np.ascontiguousarray(conditions, dtype=np.float64)
        self.variables = np.ascontiguousarray(variables, dtype=np.float64)
        self.custom = np.ascontiguousarray(custom, dtype=np.float64)
        self.seed = seed
        self.night = night

    @property
    def n_cells(self):
        return self.conditions.shape[0]

    def subset(self, idx):
        idx = np.asarray(idx)
        return CellInputs(self.conditions[idx], self.variables[idx], self.custom[idx], self.seed,
                          None if self.night is None else self.night[idx])


def _conditions(T, P):
    rho = P / (GAS_CONSTANT * T)  # Conditions::CalculateIdealAirDensity, conditions.hpp:20-23
    return np.stack([T, P, rho], axis=1)


def readme_inputs(n_cells, seed=1):
    """T uniform 270-310 K, P uniform 90-105 kPa, Foo = 20 mol m-3, Bar = Baz = 0."""
    st = CellStreams(seed, n_cells)
    T = st.uniform(270.0, 310.0)
    P = st.uniform(90.0e3, 105.0e3)
    y = np.zeros((n_cells, 3))
    y[:, 0] = 20.0
    return CellInputs(_conditions(T, P), y, np.zeros((n_cells, 0)), seed)


def random_inputs(desc, n_cells, seed, t_range, pressure=None, log_pressure=None, night_fraction=0.0,
                  y_range=(1e-12, 1e-3), photolysis_range=(1e-8, 1e-2)):
    """Conditions, log-uniform initial concentrations and log-uniform photolysis rates
    (zero in a random `night_fraction` of cells)."""
    st = CellStreams(seed, n_cells)
    T = st.uniform(*t_range)
    if log_pressure is not None:
        P = st.log_uniform(*log_pressure)
    else:
        P = np.full(n_cells, float(pressure if pressure is not None else 101325.0))
    night = st.uniform() < night_fraction
    y = np.stack([st.log_uniform(*y_range) for _ in range(desc.n_species)], axis=1)
    n_custom = desc.n_custom_parameters
    custom = np.stack([st.log_uniform(*photolysis_range) for _ in range(n_custom)], axis=1) \
        if n_custom else np.zeros((n_cells, 0))
    custom[night] = 0.0
    return CellInputs(_conditions(T, P), y, custom, seed, night)


def inputs_60_200(desc, n_cells, seed=7):
    """T uniform 220-310 K, P = 101325 Pa, y log-uniform 1e-12..1e-3, J log-uniform 1e-8..1e-2."""
    return random_inputs(desc, n_cells, seed, (220.0, 310.0), pressure=101325.0)


def inputs_210_560(desc, n_cells, seed=11):
    """T uniform 180-320 K, P log-uniform 1-101 kPa, 50 % night cells with zero photolysis."""
    return random_inputs(desc, n_cells, seed, (180.0, 320.0), log_pressure=(1.0e3, 1.01e5), night_fraction=0.5)
