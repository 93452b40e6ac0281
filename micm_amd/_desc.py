"""Flat mechanism description shared by the solver, the oracle and the tests.

`MechanismDesc` holds a mechanism as numpy arrays and exposes it as the C
struct `micm_hip_mechanism_desc` of include/micm_hip.h.  Species are the
phase's non-parameterized species in declaration order; a negative species id
-(p+1) names parameterized species p (a third body such as M, whose
concentration is the linear function of (T, P, rho) of
include/micm/util/parameterized_function.hpp:36).
"""
import ctypes

import numpy as np

RATE_PARAM_STRIDE = 24

ARRHENIUS, TROE, TERNARY, BRANCHED, TUNNELING, TAYLOR, REVERSIBLE, USER_DEFINED, SURFACE = range(9)
RATE_TYPE_NAMES = ("Arrhenius", "Troe", "TernaryChemicalActivation", "Branched", "Tunneling",
                   "TaylorSeries", "Reversible", "UserDefined", "Surface")
# custom parameters consumed per reaction (reaction_rate_store.hpp:308-333)
CUSTOM_PARAMS_PER_TYPE = (0, 0, 0, 0, 0, 0, 0, 1, 2)


class MechanismDescC(ctypes.Structure):
    _fields_ = [
        ("n_species", ctypes.c_int64),
        ("n_reactions", ctypes.c_int64),
        ("reactant_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("reactants", ctypes.POINTER(ctypes.c_int64)),
        ("product_offsets", ctypes.POINTER(ctypes.c_int64)),
        ("products", ctypes.POINTER(ctypes.c_int64)),
        ("yields", ctypes.POINTER(ctypes.c_double)),
        ("rate_types", ctypes.POINTER(ctypes.c_int32)),
        ("rate_params", ctypes.POINTER(ctypes.c_double)),
        ("n_parameterized", ctypes.c_int64),
        ("parameterized", ctypes.POINTER(ctypes.c_double)),
        ("absolute_tolerances", ctypes.POINTER(ctypes.c_double)),
        ("reorder_state", ctypes.c_int32),
        ("ignore_unused_species", ctypes.c_int32),
    ]


def _ptr(a, ctype):
    if a is None:
        return ctypes.POINTER(ctype)()
    return a.ctypes.data_as(ctypes.POINTER(ctype))


class MechanismDesc:
    """A mechanism as flat arrays.

    reactions: sequence of (reactants, products, rate_type, params) where
      reactants is a list of species ids, products a list of (species id, yield),
      rate_type one of the module constants and params a sequence of at most
      RATE_PARAM_STRIDE floats laid out as documented in include/micm_hip.h.
    """

    def __init__(self, n_species, reactions, parameterized=(), absolute_tolerances=None,
                 reorder_state=True, ignore_unused_species=True, species_names=None):
        self.n_species = int(n_species)
        self.n_reactions = len(reactions)
        self.species_names = list(species_names) if species_names is not None else [f"S{i}" for i in range(n_species)]
        r_off, r_ids, p_off, p_ids, ylds, types = [0], [], [0], [], [], []
        params = np.zeros((max(self.n_reactions, 1), RATE_PARAM_STRIDE), dtype=np.float64)
        for r, (reactants, products, rtype, prm) in enumerate(reactions):
            r_ids.extend(int(x) for x in reactants)
            r_off.append(len(r_ids))
            for sid, y in products:
                p_ids.append(int(sid))
                ylds.append(float(y))
            p_off.append(len(p_ids))
            types.append(int(rtype))
            prm = list(prm)
            if len(prm) > RATE_PARAM_STRIDE:
                raise ValueError("too many rate parameters")
            params[r, :len(prm)] = prm
        self.reactant_offsets = np.asarray(r_off, dtype=np.int64)
        self.reactants = np.asarray(r_ids if r_ids else [0], dtype=np.int64)
        self.product_offsets = np.asarray(p_off, dtype=np.int64)
        self.products = np.asarray(p_ids if p_ids else [0], dtype=np.int64)
        self.yields = np.asarray(ylds if ylds else [0.0], dtype=np.float64)
        self.rate_types = np.asarray(types if types else [0], dtype=np.int32)
        self.rate_params = np.ascontiguousarray(params)
        self.parameterized = np.asarray(parameterized, dtype=np.float64).reshape(-1, 4) if len(parameterized) else None
        self.absolute_tolerances = (np.ascontiguousarray(absolute_tolerances, dtype=np.float64)
                                    if absolute_tolerances is not None else None)
        self.reorder_state = bool(reorder_state)
        self.ignore_unused_species = bool(ignore_unused_species)
        self.reactions = [(list(a), [(int(s), float(y)) for s, y in b], int(t), list(p)) for a, b, t, p in reactions]
        self._c = None

    @property
    def n_custom_parameters(self):
        return int(sum(CUSTOM_PARAMS_PER_TYPE[t] for t in self.rate_types[:self.n_reactions]))

    def sorted_reaction_order(self):
        """Declared reaction index of each rate-constant column (stable sort by type,
        include/micm/solver/solver_builder.inl:274-329)."""
        return sorted(range(self.n_reactions), key=lambda r: int(self.rate_types[r]))

    def as_c(self):
        """The C struct; valid while this object is alive."""
        if self._c is None:
            c = MechanismDescC()
            c.n_species = self.n_species
            c.n_reactions = self.n_reactions
            c.reactant_offsets = _ptr(self.reactant_offsets, ctypes.c_int64)
            c.reactants = _ptr(self.reactants, ctypes.c_int64)
            c.product_offsets = _ptr(self.product_offsets, ctypes.c_int64)
            c.products = _ptr(self.products, ctypes.c_int64)
            c.yields = _ptr(self.yields, ctypes.c_double)
            c.rate_types = _ptr(self.rate_types, ctypes.c_int32)
            c.rate_params = _ptr(self.rate_params, ctypes.c_double)
            c.n_parameterized = 0 if self.parameterized is None else len(self.parameterized)
            c.parameterized = _ptr(self.parameterized, ctypes.c_double)
            c.absolute_tolerances = _ptr(self.absolute_tolerances, ctypes.c_double)
            c.reorder_state = int(self.reorder_state)
            c.ignore_unused_species = int(self.ignore_unused_species)
            self._c = c
        return self._c
