"""The CPU restatement (oracle) against the reference: golden fixtures written by
the reference's own solver, the live reference build where present, the README
table, and closed forms that do not share the reference's formulas."""
import os

import numpy as np
import pytest

from helpers import bits_equal
from mechanisms import all_rate_types, all_types_inputs, ts1_inputs, ts1_mechanism
from micm_amd import synthetic
from oracle import Oracle, Reference, reference_available

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["readme", "all_types", "60_200", "210_560", "ts1"]
MECHS = {"readme": synthetic.readme_mechanism, "all_types": all_rate_types, "60_200": synthetic.mechanism_60_200,
         "210_560": synthetic.mechanism_210_560, "ts1": ts1_mechanism}


def golden(name):
    return np.load(os.path.join(GOLDEN, f"{name}.npz"))


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_golden_bitwise(name):
    g = golden(name)
    o = Oracle(MECHS[name](), int(g["tableau"][0]))
    assert np.array_equal(o.species_order(), g["species_order"])
    p = o.patterns()
    for key, (r, c) in (("jac", p["jacobian"]), ("lower", p["lower"]), ("upper", p["upper"])):
        assert np.array_equal(r, g[f"{key}_rows"]) and np.array_equal(c, g[f"{key}_cols"])
    assert bits_equal(o.rate_constants(g["conditions"], g["custom"]), g["k"])
    assert bits_equal(o.forcing(g["k"], g["y"]), g["forcing"])
    assert bits_equal(o.jacobian(g["k"], g["y"]), g["jacobian"])
    a, lo, up = o.lu(g["jacobian"], float(g["alpha"][0]))
    assert bits_equal(a, g["alpha_jacobian"]) and bits_equal(lo, g["lower"]) and bits_equal(up, g["upper"])
    assert bits_equal(o.lu_solve(g["lower"], g["upper"], g["forcing"]), g["solve"])
    y, st, stats, ft = o.solve(g["conditions"], g["custom"], g["y"], float(g["dt"][0]), int(g["n_solves"][0]))
    assert bits_equal(y, g["y_per_cell"])
    assert np.array_equal(st, g["state_per_cell"]) and np.array_equal(stats, g["stats_per_cell"])
    assert bits_equal(ft, g["final_time_per_cell"])


def test_cell_grouping_matches_reference():
    """Config-1 cells solved in one 1000-cell State (the reference's lockstep step
    control) and in 1000 one-cell States: both groupings reproduce the reference."""
    g = golden("readme")
    o = Oracle(synthetic.readme_mechanism(), 3)
    cond, y0 = g["lockstep_conditions"], g["lockstep_y0"]
    custom = np.zeros((len(y0), 0))
    y, _, stats, _ = o.solve(cond, custom, y0, 500.0, 10, lockstep=True)
    assert bits_equal(y, g["lockstep_y"]) and np.array_equal(stats, g["lockstep_stats"])
    y, _, stats, _ = o.solve(cond, custom, y0, 500.0, 10)
    assert bits_equal(y, g["percell_y"]) and np.array_equal(stats, g["percell_stats"])
    # the two groupings differ above the 1e-8 bar: the grouping is part of the contract
    rel = np.max(np.abs(g["lockstep_y"] - g["percell_y"]) / np.abs(g["percell_y"]))
    assert rel > 1e-8


@pytest.mark.skipif(not reference_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_live_reference(name):
    desc = MECHS[name]()
    tableau = 3 if name in ("readme", "ts1") else 4
    ref, o = Reference(desc, tableau), Oracle(desc, tableau)
    n = 40 if name not in ("210_560", "ts1") else 8
    if name == "readme":
        cells = synthetic.readme_inputs(n, seed=9)
    elif name == "all_types":
        cells = all_types_inputs(desc, n, seed=9)
    elif name == "ts1":
        cells = ts1_inputs(desc, n, seed=9)
    else:
        cells = synthetic.random_inputs(desc, n, 9, (190.0, 320.0), log_pressure=(2e3, 1.05e5), night_fraction=0.3)
    y = cells.variables[:, ref.species_order()]
    k = ref.rate_constants(cells.conditions, cells.custom)
    assert bits_equal(o.rate_constants(cells.conditions, cells.custom), k)
    j = ref.jacobian(k, y)
    assert bits_equal(o.jacobian(k, y), j)
    # AddForcingTerms / SubtractJacobianTerms accumulate into what the output holds
    rng = np.random.default_rng(5)
    f0 = rng.standard_normal(y.shape) * 1e-3
    assert bits_equal(o.forcing(k, y, f0), ref.forcing(k, y, f0))
    j0 = rng.standard_normal(j.shape)
    assert bits_equal(o.jacobian(k, y, j0), ref.jacobian(k, y, j0))
    for tab in ([2, 3, 4, 41, 61] if name in ("readme", "all_types") else [tableau]):
        r2, o2 = Reference(desc, tab), Oracle(desc, tab)
        a = r2.solve(cells.conditions, cells.custom, y, 120.0, 2)
        b = o2.solve(cells.conditions, cells.custom, y, 120.0, 2)
        assert bits_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


# README.md:171-181: Foo, Bar, Baz printed with %.2e after each 500 s solve
README_TABLE = [
    (11.8, 5.90, 1.91), (6.79, 9.05, 3.32), (3.83, 10.7, 4.21), (2.14, 11.7, 4.74), (1.19, 12.2, 5.04),
    (0.658, 12.4, 5.21), (0.364, 12.6, 5.31), (0.201, 12.7, 5.36), (0.111, 12.7, 5.39), (0.0613, 12.8, 5.41),
]


def test_readme_table():
    desc = synthetic.readme_mechanism()
    o = Oracle(desc, 3)
    T, P = 287.45, 101319.9
    cond = np.array([[T, P, P / (synthetic.GAS_CONSTANT * T)]])
    order = o.species_order()
    y = np.zeros((1, 3))
    y[0, list(order).index(0)] = 20.0
    for row in README_TABLE:
        y, st, _, _ = o.solve(cond, np.zeros((1, 0)), y, 500.0, 1)
        declared = np.empty(3)
        declared[order] = y[0]
        assert st[0] == 2
        assert [float(f"{v:.2e}") for v in declared] == [float(f"{v:.2e}") for v in row]


def _all_types_cells(n=32, seed=13):
    return all_types_inputs(all_rate_types(), n, seed)


def test_rate_constants_closed_forms():
    """Rate constants of every type against closed forms evaluated in log space
    (not the reference's expression), relative 1e-13."""
    desc = all_rate_types()
    o = Oracle(desc, 4)
    cells = _all_types_cells()
    k = o.rate_constants(cells.conditions, cells.custom)
    T, P, rho = cells.conditions.T
    sorted_rx = [desc.reactions[r] for r in desc.sorted_reaction_order()]
    cust = iter(range(cells.custom.shape[1]))
    for col, (reactants, _, rtype, p) in enumerate(sorted_rx):
        p = list(p) + [0.0] * 24
        if rtype == 0:
            lk = np.log(p[0]) + p[2] / T + p[1] * np.log(T / p[3]) + np.log1p(p[4] * P)
        elif rtype in (1, 2):
            lk0 = np.log(p[0]) + p[2] / T + p[1] * np.log(T / 300.0)
            lkinf = np.log(p[3]) + p[5] / T + p[4] * np.log(T / 300.0)
            x = np.exp(lk0 - lkinf) * rho
            lk = lk0 + (np.log(rho) if rtype == 1 else 0.0) - np.log1p(x) + np.log(p[6]) * p[7] / (
                p[7] + (np.log(x) / np.log(10.0)) ** 2)
        elif rtype == 3:
            k0 = 2.0e-22 * 6.02214076e23 * 1e-6 * np.exp(int(p[4]))
            air = 2.45e19 / 6.02214076e23 * 1e6
            a, b = k0 * air, 0.43 * (293.0 / 298.0) ** -8.0
            z = a / (1 + a / b) * 0.41 ** (1 / (1 + np.log10(a / b) ** 2)) * (1 - p[3]) / p[3]
            a, b = k0 * rho, 0.43 * (T / 298.0) ** -8.0
            A = a / (1 + a / b) * 0.41 ** (1 / (1 + np.log10(a / b) ** 2))
            frac = z / (z + A) if p[0] == 0.0 else A / (A + z)
            lk = np.log(p[1]) - p[2] / T + np.log(frac)
        elif rtype == 4:
            lk = np.log(p[0]) - p[1] / T + p[2] / T ** 3
        elif rtype == 5:
            poly = sum(p[6 + j] * T ** j for j in range(int(p[5])))
            lk = np.log(poly) + np.log(p[0]) + p[2] / T + p[1] * np.log(T / p[3]) + np.log1p(p[4] * P)
        elif rtype == 6:
            lk = np.log(p[0]) + p[1] / T + np.log(p[2])
        elif rtype == 7:
            lk = np.log(cells.custom[:, next(cust)] * p[0])
        elif rtype == 8:
            r_, n_ = cells.custom[:, next(cust)], cells.custom[:, next(cust)]
            speed = np.sqrt(8.0 * synthetic.GAS_CONSTANT * T / (np.pi * p[1]))
            lk = np.log(4.0 * n_ * np.pi * r_ ** 2 / (r_ / p[0] + 4.0 / (speed * p[2])))
        mult = np.ones_like(T)
        for s in reactants:
            if s < 0:
                c0, cT, cP, cr = desc.parameterized[-s - 1]
                mult = mult * (c0 + cT * T + cP * P + cr * rho)
        expected = np.exp(lk) * mult
        np.testing.assert_allclose(k[:, col], expected, rtol=1e-13, err_msg=f"column {col} type {rtype}")


def test_jacobian_is_derivative_of_forcing():
    """-J from the Jacobian kernel against central differences of the forcing
    (as test_process_set_policy.hpp:404-487 does with jacobian_verification.hpp).
    Mass-action forcing with at most two reactants is at most quadratic in each
    concentration, so a central difference with a large step is exact up to rounding."""
    desc = synthetic.mechanism_60_200()
    o = Oracle(desc, 4)
    cells = synthetic.inputs_60_200(desc, 3, seed=2)
    y = cells.variables[:, o.species_order()]
    k = o.rate_constants(cells.conditions, cells.custom)
    jac = o.jacobian(k, y)
    rows, cols = o.patterns()["jacobian"]
    n = y.shape[1]
    for c in range(3):
        dense = np.zeros((n, n))
        dense[rows, cols] = -jac[c]
        fd = np.zeros((n, n))
        noise = np.zeros((n, n))
        for j in range(n):
            h = max(abs(y[c, j]), 1e-3)
            yp, ym = y[c:c + 1].copy(), y[c:c + 1].copy()
            yp[0, j] += h
            ym[0, j] -= h
            fp, fm = o.forcing(k[c:c + 1], yp)[0], o.forcing(k[c:c + 1], ym)[0]
            fd[:, j] = (fp - fm) / (2 * h)
            noise[:, j] = 1e-13 * (np.abs(fp) + np.abs(fm)) / h
        assert np.all(np.abs(dense - fd) <= 1e-10 * np.abs(dense) + noise)
        assert np.all(fd[dense == 0] == 0)  # the pattern holds every non-zero


def _dense(rows, cols, vals, n):
    m = np.zeros((n, n))
    m[rows, cols] = vals
    return m


def test_lu_reconstructs_and_solves():
    """L U == alpha I - J and the substitution inverts it (test_lu_decomposition_policy.hpp,
    test_linear_solver_policy.hpp), checked with dense numpy algebra."""
    for name in ("all_types", "60_200", "210_560"):
        g = golden(name)
        o = Oracle(MECHS[name](), 4)
        p = o.patterns()
        n = len(g["species_order"])
        for c in range(min(3, g["k"].shape[0])):
            A = _dense(*p["jacobian"], g["alpha_jacobian"][c], n)
            L = _dense(*p["lower"], g["lower"][c], n)
            U = _dense(*p["upper"], g["upper"][c], n)
            assert np.allclose(np.diag(L), 1.0) and np.all(np.triu(L, 1) == 0) and np.all(np.tril(U, -1) == 0)
            assert np.max(np.abs(L @ U - A)) <= 1e-10 * np.max(np.abs(A))
            x = g["solve"][c]
            b = g["forcing"][c]
            assert np.max(np.abs(A @ x - b)) <= 1e-9 * (np.max(np.abs(A)) * np.max(np.abs(x)) + np.max(np.abs(b)))


def test_analytical_two_step_decay():
    """A -k1-> B -k2-> C against its closed form (test/integration/analytical_policy.hpp:138-276),
    at tight tolerances."""
    from micm_amd._desc import ARRHENIUS, MechanismDesc
    desc = MechanismDesc(3, [([0], [(1, 1.0)], ARRHENIUS, [4.0e-3, 0.0, 50.0, 300.0, 0.0]),
                             ([1], [(2, 1.0)], ARRHENIUS, [1.2e-3, 0.0, -75.0, 300.0, 0.0])],
                         absolute_tolerances=[1e-12] * 3)
    T = np.array([272.5, 254.7, 312.6])
    P = np.array([101253.3, 100672.5, 101319.8])
    cond = np.stack([T, P, P / (synthetic.GAS_CONSTANT * T)], 1)
    # (the 2-stage tableau is left out: its embedded error estimate, e = (1/2g, 1/2g)
    # at rosenbrock_solver_parameters.hpp:174-176, is the size of the step itself, so
    # the reference's own solver cannot grow the step at these tolerances)
    for tableau in (3, 4, 41, 61):
        o = Oracle(desc, tableau)
        order = o.species_order()
        y0 = np.zeros((3, 3))
        y0[:, 0] = [1.0, 0.6666666, 0.3333333]
        y = y0[:, order]
        k1 = 4.0e-3 * np.exp(50.0 / T)
        k2 = 1.2e-3 * np.exp(-75.0 / T)
        for step in range(1, 101):
            y, st, _, _ = o.solve(cond, np.zeros((3, 0)), y, 10.0, 1, rtol=1e-10)
            t = 10.0 * step
            A = y0[:, 0] * np.exp(-k1 * t)
            B = y0[:, 0] * k1 / (k2 - k1) * (np.exp(-k1 * t) - np.exp(-k2 * t))
            C = y0[:, 0] - A - B
            got = np.empty_like(y)
            got[:, order] = y
            assert np.all(st == 2)
            np.testing.assert_allclose(got, np.stack([A, B, C], 1), atol=2e-7, rtol=0)


@pytest.mark.skipif(not reference_available(), reason="reference build (oracle/_ref) not available")
@pytest.mark.parametrize("group", [1, 4])
def test_lockstep_orders_match_live_reference(group):
    """A multi-cell State over Matrix<Real> and over VectorMatrix<4>: the oracle's lockstep
    solve in the same summation order equals the reference's own State bit for bit, and
    the two orders give different step sequences on this mechanism (so the order is
    observable and pinned)."""
    desc = synthetic.mechanism_60_200()
    cells = synthetic.inputs_60_200(desc, 130)
    o, r = Oracle(desc, 4), Reference(desc, 4)
    y0 = cells.variables[:, o.species_order()]
    yo, so, sto, fto = o.solve(cells.conditions, cells.custom, y0, 300.0, 1, lockstep=group)
    yr, sr, str_, ftr = r.solve(cells.conditions, cells.custom, y0, 300.0, 1, lockstep=group)
    assert bits_equal(yo, yr) and np.array_equal(sto, str_) and np.array_equal(so, sr)
    other = 4 if group == 1 else 1
    _, _, st_other, _ = o.solve(cells.conditions, cells.custom, y0, 300.0, 1, lockstep=other)
    assert not np.array_equal(sto, st_other)


def test_grouped_norm_with_group_one_is_the_reference_norm():
    desc = synthetic.mechanism_60_200()
    o = Oracle(desc, 4)
    rng = np.random.default_rng(2)
    y = rng.lognormal(-8, 3, (77, o.n_species))
    ynew = y * 1.001
    err = 1e-6 * y * rng.standard_normal(y.shape)
    atol = np.full(o.n_species, 1e-12)
    assert o.normalized_error_grouped(y, ynew, err, 1e-4, atol, 1) == o.normalized_error(y, ynew, err, 1e-4, atol)
    if reference_available():
        assert o.normalized_error(y, ynew, err, 1e-4, atol) == Reference(desc, 4).normalized_error(y, ynew, err, 1e-4,
                                                                                                  atol)
