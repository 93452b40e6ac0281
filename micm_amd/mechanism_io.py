"""On-disk mechanism description (JSON, versioned).

A file holds what `MechanismDesc` holds, with names::

    {"format": "micm_amd.mechanism", "version": 1,
     "species": ["O2", ...],                       # declared order
     "parameterized": [{"name": "M", "coefficients": [c0, cT, cP, crho]}],
     "absolute_tolerances": [...],                  # optional, declared order
     "reactions": [{"reactants": [i, ...],          # -(p+1): parameterized species p
                    "products": [[i, yield], ...],
                    "type": "arrhenius" | "troe" | "ternary" | "branched" | "tunneling" |
                            "taylor" | "reversible" | "user_defined" | "surface",
                    "params": [...],                # layout of include/micm_hip.h
                    "label": "jo2_b"}],             # user-defined / surface only
     "compiled": {"species_order": [...], "reaction_order": [...],
                  "jacobian": [[rows], [cols]], "lower": [...], "upper": [...]},  # optional
     "initial": {...}}                              # optional, free-form

`compiled`, when present, is checked on load against a fresh compilation of the
description, so a file written by one version of the mechanism compiler fails
loudly when read by a compiler that would order species or fill the LU
differently.  Custom rate parameters are ordered as the reference orders them:
one per user-defined reaction and two per surface reaction (effective radius,
particle number concentration), in the sorted reaction order
(reaction_rate_store.hpp:308-333); `custom_parameter_labels` lists them.
"""
import json

import numpy as np

from ._desc import (ARRHENIUS, BRANCHED, REVERSIBLE, SURFACE, TAYLOR, TERNARY, TROE, TUNNELING, USER_DEFINED,
                    MechanismDesc)

FORMAT = "micm_amd.mechanism"
VERSION = 1

TYPE_NAMES = {"arrhenius": ARRHENIUS, "troe": TROE, "ternary": TERNARY, "branched": BRANCHED,
              "tunneling": TUNNELING, "taylor": TAYLOR, "reversible": REVERSIBLE, "user_defined": USER_DEFINED,
              "surface": SURFACE}
_TYPE_OF = {v: k for k, v in TYPE_NAMES.items()}


class MechanismFileError(ValueError):
    pass


def from_dict(d):
    """MechanismDesc from a parsed file; `labels` (per declared reaction) and
    `initial` are attached to the returned object."""
    if d.get("format", FORMAT) != FORMAT or int(d.get("version", VERSION)) != VERSION:
        raise MechanismFileError(f"not a {FORMAT} version {VERSION} file")
    species = list(d["species"])
    n = len(species)
    params = [p["coefficients"] for p in d.get("parameterized", [])]
    reactions, labels = [], []
    for k, r in enumerate(d["reactions"]):
        t = r["type"]
        if t not in TYPE_NAMES:
            raise MechanismFileError(f"reaction {k}: unknown rate-constant type {t!r}")
        for s in r["reactants"]:
            if not (-len(params) <= int(s) < n):
                raise MechanismFileError(f"reaction {k}: reactant {s} out of range")
        for s, _ in r["products"]:
            if not (0 <= int(s) < n):
                raise MechanismFileError(f"reaction {k}: product {s} out of range")
        reactions.append((r["reactants"], [(int(s), float(y)) for s, y in r["products"]], TYPE_NAMES[t],
                          [float(x) for x in r["params"]]))
        labels.append(r.get("label"))
    desc = MechanismDesc(n, reactions, parameterized=params, absolute_tolerances=d.get("absolute_tolerances"),
                         species_names=species)
    desc.labels = labels
    desc.parameterized_names = [p["name"] for p in d.get("parameterized", [])]
    desc.initial = d.get("initial")
    compiled = d.get("compiled")
    if compiled is not None:
        mine = compiled_summary(desc)
        for key in ("species_order", "reaction_order", "jacobian", "lower", "upper"):
            if key in compiled and compiled[key] != mine[key]:
                raise MechanismFileError(f"stored {key} differs from this compiler's")
    return desc


def load(path):
    with open(path) as f:
        return from_dict(json.load(f))


def compiled_summary(desc):
    """Species order, reaction order and the Jacobian / L / U patterns the
    mechanism compiler produces for `desc` (needs the HIP library, not a GPU)."""
    from . import _lib
    m = _lib.Mechanism(desc)
    p = m.patterns()
    return {"species_order": [int(x) for x in m.species_order()],
            "reaction_order": [int(x) for x in m.reaction_order()],
            "jacobian": [[int(x) for x in a] for a in p["jacobian"]],
            "lower": [[int(x) for x in a] for a in p["lower"]],
            "upper": [[int(x) for x in a] for a in p["upper"]]}


def to_dict(desc, labels=None, include_compiled=False, initial=None):
    labels = labels if labels is not None else getattr(desc, "labels", None)
    out = {"format": FORMAT, "version": VERSION, "species": list(desc.species_names),
           "parameterized": [{"name": (getattr(desc, "parameterized_names", None) or [f"P{i}" for i in range(
               0 if desc.parameterized is None else len(desc.parameterized))])[i],
                              "coefficients": [float(x) for x in c]}
                             for i, c in enumerate([] if desc.parameterized is None else desc.parameterized)],
           "reactions": []}
    if desc.absolute_tolerances is not None:
        out["absolute_tolerances"] = [float(x) for x in desc.absolute_tolerances]
    for k, (reactants, products, t, prm) in enumerate(desc.reactions):
        r = {"reactants": [int(s) for s in reactants], "products": [[int(s), float(y)] for s, y in products],
             "type": _TYPE_OF[t], "params": [float(x) for x in prm]}
        if labels is not None and labels[k] is not None:
            r["label"] = labels[k]
        out["reactions"].append(r)
    if include_compiled:
        out["compiled"] = compiled_summary(desc)
    if initial is not None:
        out["initial"] = initial
    return out


def save(desc, path, labels=None, include_compiled=True, initial=None):
    with open(path, "w") as f:
        json.dump(to_dict(desc, labels, include_compiled, initial), f, separators=(",", ":"))


def custom_parameter_labels(desc):
    """Labels of the custom rate parameters in the solver's order."""
    labels = getattr(desc, "labels", None) or [None] * desc.n_reactions
    out = []
    for r in desc.sorted_reaction_order():
        t = int(desc.rate_types[r])
        name = labels[r] if labels[r] is not None else f"R{r}"
        if t == USER_DEFINED:
            out.append(name)
        elif t == SURFACE:
            out.append(f"{name}.effective radius [m]")
            out.append(f"{name}.particle number concentration [# m-3]")
    return out


def initial_arrays(desc, n_cells=1):
    """(conditions [n, 3], variables [n, N] declared order, custom [n, P]) from the
    file's `initial` block (concentrations and custom parameters by name)."""
    ini = desc.initial
    if ini is None:
        raise MechanismFileError("the file has no initial state")
    y = np.array([ini["concentrations"].get(s, 0.0) for s in desc.species_names], dtype=np.float64)
    custom = np.array([ini["custom"][lab] for lab in custom_parameter_labels(desc)], dtype=np.float64)
    cond = np.array([ini["temperature"], ini["pressure"], ini["air_density"]], dtype=np.float64)
    return np.tile(cond, (n_cells, 1)), np.tile(y, (n_cells, 1)), np.tile(custom, (n_cells, 1))
