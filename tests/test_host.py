"""Host-side checks that need no GPU: the C-ABI library loads and exports every
function of include/micm_hip.h, the mechanism compiler reproduces the
reference's species order and sparsity patterns, the API mirror builds the same
description as the flat generators, the seeded generators are stable, and bad
mechanisms are rejected with status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

from mechanisms import ts1_mechanism
from micm_amd import _lib, synthetic
from micm_amd import (ArrheniusRateConstantParameters, ChemicalReactionBuilder, HipSolverBuilder, Phase,
                      RosenbrockSolverParameters, Species, StoichSpecies, System)
from micm_amd._desc import ARRHENIUS, MechanismDesc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def header_functions():
    text = open(os.path.join(ROOT, "include", "micm_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|void|const char\*)\s+(micm_hip_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("name,fn", [("readme", synthetic.readme_mechanism), ("60_200", synthetic.mechanism_60_200),
                                     ("210_560", synthetic.mechanism_210_560), ("ts1", ts1_mechanism)])
def test_mechanism_compiler_matches_reference(name, fn):
    g = np.load(os.path.join(GOLDEN, f"{name}.npz"))
    m = _lib.Mechanism(fn())  # host compilation only; nothing touches the device
    assert np.array_equal(m.species_order(), g["species_order"])
    p = m.patterns()
    for key, (r, c) in (("jac", p["jacobian"]), ("lower", p["lower"]), ("upper", p["upper"])):
        assert np.array_equal(r, g[f"{key}_rows"]) and np.array_equal(c, g[f"{key}_cols"])
    d = m.jacobian_diagonal()
    assert np.array_equal(p["jacobian"][0][d], np.arange(m.n_species))
    assert np.array_equal(p["jacobian"][1][d], np.arange(m.n_species))


def test_reaction_order_is_stable_type_sort():
    desc = synthetic.mechanism_60_200()
    m = _lib.Mechanism(desc)
    assert list(m.reaction_order()) == desc.sorted_reaction_order()


def test_api_mirror_builds_the_readme_description():
    foo, bar, baz = Species("Foo"), Species("Bar"), Species("Baz")
    gas = Phase("gas", [foo, bar, baz])
    r1 = (ChemicalReactionBuilder().SetReactants([foo]).SetProducts([StoichSpecies(bar, 0.8), StoichSpecies(baz, 0.2)])
          .SetRateConstant(ArrheniusRateConstantParameters(A=1.0e-3)).SetPhase(gas).Build())
    r2 = (ChemicalReactionBuilder().SetReactants([foo, bar]).SetProducts([StoichSpecies(baz, 1)])
          .SetRateConstant(ArrheniusRateConstantParameters(A=1.0e-5, C=110.0)).SetPhase(gas).Build())
    builder = HipSolverBuilder(RosenbrockSolverParameters.ThreeStageRosenbrockParameters())
    desc, _ = builder.SetSystem(System(gas)).SetReactions([r1, r2]).description()
    ref = synthetic.readme_mechanism()
    for a in ("reactant_offsets", "reactants", "product_offsets", "products", "yields", "rate_types", "rate_params"):
        assert np.array_equal(getattr(desc, a), getattr(ref, a)), a


def test_tableau_presets():
    p3 = _lib.rosenbrock_parameters(3)
    assert p3.stages == 3 and p3.max_number_of_steps == 1000 and p3.factor_max == 6.0
    p4 = _lib.rosenbrock_parameters(4)
    assert p4.stages == 4 and list(p4.new_function_evaluation[:4]) == [1, 1, 1, 0]
    with pytest.raises(_lib.MicmHipError):
        _lib.check(_lib.lib().micm_hip_rosenbrock_parameters_preset(5, ctypes.byref(_lib.RosenbrockParameters())))


def test_synthetic_generators_are_stable():
    d = synthetic.mechanism_210_560()
    assert d.n_species == 210 and d.n_reactions == 560
    assert np.bincount(d.rate_types, minlength=8)[[0, 1, 2, 7]].tolist() == [420, 56, 28, 56]
    assert int(d.reactants.sum()) == 84397 and abs(float(d.yields.sum()) - 1165.48035429837) < 1e-9
    cells = synthetic.inputs_210_560(d, 1000)
    assert abs(cells.night.mean() - 0.5) < 0.05
    assert np.all(cells.custom[cells.night] == 0) and np.all(cells.custom[~cells.night] > 0)
    assert 180 <= cells.conditions[:, 0].min() and cells.conditions[:, 0].max() <= 320
    again = synthetic.inputs_210_560(d, 1000)
    assert np.array_equal(again.variables, cells.variables)


def test_xoshiro_matches_published_algorithm():
    """xoshiro256** seeded by SplitMix64 from 0: the first outputs of the published
    C implementation (Blackman & Vigna), and per-cell streams equal scalar streams."""
    r = synthetic.Xoshiro256ss(0)
    assert [r.next_u64() for _ in range(3)] == [11091344671253066420, 13793997310169335082, 1900383378846508768]
    cs = synthetic.CellStreams(77, 5)
    draws = np.stack([cs.next_u64() for _ in range(4)], 1)
    for c in range(5):
        sc = synthetic.Xoshiro256ss((77 + (c + 1) * synthetic.GOLDEN) & synthetic.M64)
        assert [int(v) for v in draws[c]] == [sc.next_u64() for _ in range(4)]


def test_bad_mechanisms_are_rejected():
    bad_species = MechanismDesc(2, [([0], [(5, 1.0)], ARRHENIUS, [1.0])])
    with pytest.raises(_lib.MicmHipError) as e:
        _lib.Mechanism(bad_species)
    assert e.value.code == -5
    unused = MechanismDesc(3, [([0], [(1, 1.0)], ARRHENIUS, [1.0])], ignore_unused_species=False)
    with pytest.raises(_lib.MicmHipError):
        _lib.Mechanism(unused)
    bad_type = MechanismDesc(2, [([0], [(1, 1.0)], 42, [1.0])])
    with pytest.raises(_lib.MicmHipError) as e:
        _lib.Mechanism(bad_type)
    assert e.value.code == -1


def test_mechanism_file_round_trip(tmp_path):
    """The on-disk description reproduces the mechanism exactly, and a file whose
    stored orders / patterns disagree with the compiler is rejected."""
    import json

    from micm_amd import mechanism_io
    desc = synthetic.mechanism_60_200()
    path = tmp_path / "m.json"
    mechanism_io.save(desc, str(path), include_compiled=True)
    back = mechanism_io.load(str(path))
    for key in ("reactant_offsets", "reactants", "product_offsets", "products", "yields", "rate_types", "rate_params"):
        assert np.array_equal(getattr(back, key), getattr(desc, key)), key
    a, b = _lib.Mechanism(desc), _lib.Mechanism(back)
    assert np.array_equal(a.species_order(), b.species_order())
    d = json.loads(path.read_text())
    d["compiled"]["species_order"][0], d["compiled"]["species_order"][1] = \
        d["compiled"]["species_order"][1], d["compiled"]["species_order"][0]
    with pytest.raises(mechanism_io.MechanismFileError):
        mechanism_io.from_dict(d)
    d = json.loads(path.read_text())
    d["reactions"][0]["type"] = "no_such_type"
    with pytest.raises(mechanism_io.MechanismFileError):
        mechanism_io.from_dict(d)


def test_ts1_file_matches_the_reference_benchmark():
    """tests/golden/ts1_mechanism.json was exported from the reference's TS1 header:
    its sizes are the reference's (test/benchmark/ts1_mechanism.hpp:9), its custom
    parameters are those bench::Ts1::InitState sets, and the compiler's species
    order and fill match the reference's (golden ts1.npz)."""
    from micm_amd import mechanism_io
    desc = ts1_mechanism()
    assert desc.n_species == 209 and desc.n_reactions == 547
    types = np.bincount(desc.rate_types[:desc.n_reactions], minlength=9)
    assert types[0] == 361 and types[1] == 31 and types[7] == 142 and types[8] == 13
    labels = mechanism_io.custom_parameter_labels(desc)
    assert len(labels) == 168 and set(labels) == set(desc.initial["custom"])
    m = _lib.Mechanism(desc)
    info = m.info
    assert (info.nnz_jacobian, info.nnz_lower, info.nnz_upper) == (1950, 1453, 1137)


def _doolittle_operation_count(lower, upper, n):
    """Walks the in-place Doolittle elimination of the filled pattern element by element
    (row i: U[i, k] for k >= i, then L[k, i] for k > i) and counts the products
    L[i, j] * U[j, k] it forms and the divisions by U[i, i]."""
    nz = set(zip(*lower)) | set(zip(*upper))
    fma = div = 0
    for i in range(n):
        for k in range(i, n):
            if (i, k) in nz:
                fma += sum(1 for j in range(i) if (i, j) in nz and (j, k) in nz)
        for k in range(i + 1, n):
            if (k, i) in nz:
                fma += sum(1 for j in range(i) if (k, j) in nz and (j, i) in nz)
                div += 1
    return fma, div


def test_fp64_operation_count():
    """bench.py's FP64 operation count: the Doolittle multiply-subtracts and divisions
    agree with an element-by-element walk of the elimination, and for TS1 with the
    counts measured on the reference's own factorisation (6,986 and 1,244)."""
    import bench
    from oracle import Oracle
    o = Oracle(synthetic.mechanism_60_200(), 3)
    pat = o.patterns()
    per = bench.flops_per_invocation(o.desc, o)
    assert (per["lu_multiply_subtracts"], per["lu_divisions"]) == \
        _doolittle_operation_count(pat["lower"], pat["upper"], o.n_species)
    t = Oracle(ts1_mechanism(), 3)
    per = bench.flops_per_invocation(t.desc, t)
    assert (per["lu_multiply_subtracts"], per["lu_divisions"]) == (6986, 1244)
    assert per["lu"] == 2 * 6986 + 1244 + 209
    assert per["solve"] == 2 * ((1453 - 209) + (1137 - 209)) + 2 * 209


def test_tableau_presets_equal_the_reference():
    """Every field of micm_hip_rosenbrock_parameters_preset for the five presets equals the
    reference's RosenbrockSolverParameters (tests/golden/tableaus.npz, written by the
    reference build; and the live reference where it is built), bit for bit."""
    import os

    import numpy as np

    from micm_amd import _lib
    from oracle import Reference, reference_available, tableau_fields
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "tableaus.npz"))
    for preset in (2, 3, 4, 41, 61):
        ours = tableau_fields(_lib.rosenbrock_parameters(preset))
        refs = [{k.split("_", 1)[1]: g[k] for k in g.files if k.split("_", 1)[0] == str(preset)}]
        if reference_available():
            refs.append(Reference.tableau(preset))
        for ref in refs:
            assert set(ref) == set(ours)
            for key, value in ours.items():
                assert np.array_equal(np.asarray(value).view(np.uint8), np.asarray(ref[key]).view(np.uint8)), \
                    (preset, key)
