"""GPU parity of every kernel of the path and of full solves, through the C-ABI.

Two checkers, which are bitwise equal to each other and to the reference's own
compiled CPU solver (tests/test_oracle.py, tests/test_pmath.py):

* `Oracle` -- the CPU restatement with the C library's exp / log10 / pow, the
  reference's own arithmetic.
* `OraclePortableMath` -- the same restatement built with the device path's
  exp / log10 / pow (micm_amd/csrc/pmath.h), which restate the C library's
  algorithms operation for operation.

The HIP library is built without FMA contraction and every kernel keeps the
reference's operand order, so every kernel and every whole solve (all forms,
failed cells included) reproduces both bit for bit: outcomes, counters, final
times and concentrations -- the 1e-8 agreement bar for species above 1e-20 mol m-3
is met with zero difference.
"""
import numpy as np
import pytest

from helpers import bits_equal, dev, dev_zeros, host, stream
from mechanisms import all_rate_types, all_types_inputs, ts1_inputs, ts1_mechanism
from micm_amd import _lib, synthetic
from oracle import Oracle, OraclePortableMath

pytestmark = pytest.mark.gpu


def _cases():
    return {
        "readme": (synthetic.readme_mechanism(), lambda d, n: synthetic.readme_inputs(n, seed=3), 3, 500.0),
        "all_types": (all_rate_types(), all_types_inputs, 4, 60.0),
        "60_200": (synthetic.mechanism_60_200(), synthetic.inputs_60_200, 4, 300.0),
        "210_560": (synthetic.mechanism_210_560(), synthetic.inputs_210_560, 4, 900.0),
        # the reference's TS1 benchmark mechanism with its benchmark's 3-stage solver and 30 s step
        "ts1": (ts1_mechanism(), ts1_inputs, 3, 30.0),
    }


CASES = _cases()


@pytest.fixture(scope="module", params=list(CASES))
def case(request):
    desc, inputs, tableau, dt = CASES[request.param]
    mech = _lib.Mechanism(desc)
    return request.param, desc, inputs, tableau, dt, mech, Oracle(desc, tableau), OraclePortableMath(desc, tableau)


@pytest.mark.parametrize("n_cells", [1, 64, 301])
def test_kernels_match_oracle(case, n_cells):
    name, desc, inputs, tableau, dt, mech, oracle, oracle_pm = case
    print(f"kernels: {name} n_cells={n_cells}", flush=True)
    cells = inputs(desc, n_cells)
    order = mech.species_order()
    y = cells.variables[:, order]
    N, R, P = mech.n_species, mech.n_reactions, mech.n_custom
    s = stream()

    # rate constants: bitwise with both restatements (the reference's own arithmetic)
    # device inputs are kept referenced until the kernels that read them have run
    k_dev = dev_zeros(n_cells, R)
    cond_dev, custom_dev = dev(cells.conditions), dev(cells.custom) if P else None
    mech.calculate_rate_constants(n_cells, cond_dev.data_ptr(), custom_dev.data_ptr() if P else None,
                                  k_dev.data_ptr(), s)
    k_gpu = host(k_dev, n_cells, R)
    k_ref = oracle.rate_constants(cells.conditions, cells.custom)
    assert bits_equal(k_gpu, oracle_pm.rate_constants(cells.conditions, cells.custom))
    assert bits_equal(k_gpu, k_ref)

    # forcing from the oracle's rate constants: bitwise
    f_dev = dev_zeros(n_cells, N)
    kr_dev, y_dev = dev(k_ref), dev(y)
    mech.add_forcing_terms(n_cells, kr_dev.data_ptr(), y_dev.data_ptr(), f_dev.data_ptr(), s)
    f_ref = oracle.forcing(k_ref, y)
    assert bits_equal(host(f_dev, n_cells, N), f_ref)

    # accumulating into a forcing that holds values (AddForcingTerms without Fill(0)), and
    # Fill(0) with AddForcingTerms in one pass over a buffer of NaNs: bitwise
    rng = np.random.default_rng(n_cells)
    f0 = rng.standard_normal((n_cells, N)) * 1e-3
    fa_dev = dev(f0)
    mech.add_forcing_terms(n_cells, kr_dev.data_ptr(), y_dev.data_ptr(), fa_dev.data_ptr(), s)
    assert bits_equal(host(fa_dev, n_cells, N), oracle.forcing(k_ref, y, f0))
    fz_dev = dev(np.full((n_cells, N), np.nan))
    mech.fill_and_add_forcing_terms(n_cells, kr_dev.data_ptr(), y_dev.data_ptr(), fz_dev.data_ptr(), s)
    assert bits_equal(host(fz_dev, n_cells, N), f_ref)

    # -J: bitwise, from zero, accumulating, and Fill(0) with SubtractJacobianTerms in one pass
    nj = mech.nnz_jacobian
    j_dev = dev_zeros(n_cells, nj)
    mech.subtract_jacobian_terms(n_cells, kr_dev.data_ptr(), y_dev.data_ptr(), j_dev.data_ptr(), s)
    j_ref = oracle.jacobian(k_ref, y)
    assert bits_equal(host(j_dev, n_cells, nj), j_ref)
    j0 = rng.standard_normal((n_cells, nj))
    ja_dev = dev(j0)
    mech.subtract_jacobian_terms(n_cells, kr_dev.data_ptr(), y_dev.data_ptr(), ja_dev.data_ptr(), s)
    assert bits_equal(host(ja_dev, n_cells, nj), oracle.jacobian(k_ref, y, j0))
    jz_dev = dev(np.full((n_cells, nj), np.nan))
    mech.fill_and_subtract_jacobian_terms(n_cells, kr_dev.data_ptr(), y_dev.data_ptr(), jz_dev.data_ptr(), s)
    assert bits_equal(host(jz_dev, n_cells, nj), j_ref)

    # alpha - J and Doolittle LU: bitwise
    alpha = 1.0 / (37.5 * 0.5728200000000000)
    a_dev = dev(j_ref)
    mech.alpha_minus_jacobian(n_cells, a_dev.data_ptr(), alpha, s)
    l_dev, u_dev = dev_zeros(n_cells, mech.nnz_lower), dev_zeros(n_cells, mech.nnz_upper)
    mech.lu_decompose(n_cells, a_dev.data_ptr(), l_dev.data_ptr(), u_dev.data_ptr(), s)
    a_ref, l_ref, u_ref = oracle.lu(j_ref, alpha)
    assert bits_equal(host(a_dev, n_cells, nj), a_ref)
    assert bits_equal(host(l_dev, n_cells, mech.nnz_lower), l_ref)
    assert bits_equal(host(u_dev, n_cells, mech.nnz_upper), u_ref)

    # forward / back substitution: bitwise
    x_dev, lr_dev, ur_dev = dev(f_ref), dev(l_ref), dev(u_ref)
    mech.lu_solve(n_cells, lr_dev.data_ptr(), ur_dev.data_ptr(), x_dev.data_ptr(), s)
    assert bits_equal(host(x_dev, n_cells, N), oracle.lu_solve(l_ref, u_ref, f_ref))

    # per-cell normalized error: bitwise against the oracle's norm of a 1-cell State
    import torch
    atol = np.full(N, 1e-3)
    ynew = y * 1.01 + 1e-9
    yerr = f_ref * 1e-4
    e_dev = torch.zeros(n_cells, dtype=torch.float64, device="cuda")
    yn_dev, ye_dev, atol_dev = dev(ynew), dev(yerr), torch.from_numpy(atol).cuda()
    mech.normalized_error(n_cells, y_dev.data_ptr(), yn_dev.data_ptr(), ye_dev.data_ptr(), atol_dev.data_ptr(), 1e-6,
                          e_dev.data_ptr(), s)
    torch.cuda.synchronize()
    e_ref = np.array([oracle.normalized_error(y[c:c + 1], ynew[c:c + 1], yerr[c:c + 1], 1e-6, atol)
                      for c in range(n_cells)])
    assert bits_equal(e_dev.cpu().numpy(), e_ref)


def _solve(mech, cells, order, tableau, dt, form, n_solves=1, rate_constants=None, handoff=-1, coop_shape=(0, -1),
           lockstep_order=1, lu_blocking=0):
    """n_solves x (UpdateStateParameters, Solve); with rate_constants given, those are
    loaded instead of being computed on the device."""
    state = _lib.DeviceState(mech, cells.n_cells)
    state.set_lockstep_order(lockstep_order)
    state.set_handoff(handoff)
    state.set_cooperative_shape(*coop_shape)
    state.set_lu_blocking(lu_blocking)
    state.set_variables(cells.variables[:, order])
    state.set_conditions(cells.conditions)
    state.set_custom_parameters(cells.custom)
    params = _lib.rosenbrock_parameters(tableau)
    stats = np.zeros((cells.n_cells, 7), dtype=np.int64)
    for _ in range(n_solves):
        if rate_constants is None:
            state.update_state_parameters()
        else:
            state.set_rate_constants(rate_constants)
        state.solve(dt, params, form)
        st, s, ft = state.results()
        stats += s
    return state.get_variables(), st, stats, ft


def _n_cells(name):
    return 200 if name == "210_560" else 515


FORMS = [_lib.SOLVE_PLAIN, _lib.SOLVE_PERSISTENT, _lib.SOLVE_COOPERATIVE]


@pytest.mark.parametrize("form", FORMS)
def test_per_cell_solve_bitwise(case, form):
    """End to end (rate constants computed on the device): every cell, failed ones
    included, takes the portable-math restatement's exact step sequence."""
    name, desc, inputs, tableau, dt, mech, oracle, oracle_pm = case
    if form == _lib.SOLVE_COOPERATIVE and not mech.cooperative_supported(tableau):
        pytest.skip("mechanism does not fit the cooperative form's LDS budget")
    n_cells = _n_cells(name)
    n_solves = 3 if name in ("readme", "all_types", "ts1") else 1
    print(f"bitwise solve: {name} form={form} n_cells={n_cells}", flush=True)
    cells = inputs(desc, n_cells)
    order = mech.species_order()
    y_gpu, st_gpu, stats_gpu, ft_gpu = _solve(mech, cells, order, tableau, dt, form, n_solves)
    y_ref, st_ref, stats_ref, ft_ref = oracle_pm.solve(cells.conditions, cells.custom, cells.variables[:, order], dt,
                                                        n_solves)
    assert np.array_equal(st_gpu, st_ref)
    assert np.array_equal(stats_gpu, stats_ref)
    assert bits_equal(ft_gpu, ft_ref)
    assert bits_equal(y_gpu, y_ref)


@pytest.mark.parametrize("form", [_lib.SOLVE_PERSISTENT, _lib.SOLVE_COOPERATIVE])
def test_per_cell_solve_against_reference_math(case, form):
    """Against the C-library restatement (bitwise equal to the reference's compiled
    solver): every cell, converged or not, bit for bit -- outcome, counters, final time
    and concentrations.  A single differing bit anywhere on the path (a rate constant,
    a step-size factor, a summation order) changes some cell's adaptive step sequence
    on these mechanisms, so this fails on any arithmetic regression."""
    name, desc, inputs, tableau, dt, mech, oracle, oracle_pm = case
    if form == _lib.SOLVE_COOPERATIVE and not mech.cooperative_supported(tableau):
        pytest.skip("mechanism does not fit the cooperative form's LDS budget")
    n_cells = _n_cells(name)
    n_solves = 3 if name in ("readme", "all_types", "ts1") else 1
    print(f"solve vs reference math: {name} form={form} n_cells={n_cells}", flush=True)
    cells = inputs(desc, n_cells)
    order = mech.species_order()
    y_gpu, st_gpu, stats_gpu, ft_gpu = _solve(mech, cells, order, tableau, dt, form, n_solves)
    y_ref, st_ref, stats_ref, ft_ref = oracle.solve(cells.conditions, cells.custom, cells.variables[:, order], dt,
                                                     n_solves)
    print(f"  converged {np.mean(st_ref == 2):.4f}, bitwise-equal cells "
          f"{np.mean(np.all(y_gpu.view(np.int64) == y_ref.view(np.int64), axis=1)):.4f}", flush=True)
    assert np.array_equal(st_gpu, st_ref)
    assert np.array_equal(stats_gpu, stats_ref)
    assert bits_equal(ft_gpu, ft_ref)
    assert bits_equal(y_gpu, y_ref)


@pytest.mark.parametrize("name", ["60_200", "210_560", "ts1"])
def test_per_cell_forms_agree_bitwise(name):
    """Every per-cell form runs the same per-cell arithmetic (the cooperative form
    with its work spread over a workgroup without reordering any sum), so plain,
    persistent (with and without hand-off) and cooperative results (every launch
    shape, table staging and LU schedule blocking) are identical."""
    desc, inputs, tableau, dt = CASES[name]
    mech = _lib.Mechanism(desc)
    cells = inputs(desc, 700 if name == "60_200" else 130, seed=21)
    order = mech.species_order()
    print(f"forms agree: {name}", flush=True)
    ref = _solve(mech, cells, order, tableau, dt, _lib.SOLVE_PLAIN)
    runs = [_solve(mech, cells, order, tableau, dt, _lib.SOLVE_PERSISTENT, handoff=0),
            _solve(mech, cells, order, tableau, dt, _lib.SOLVE_PERSISTENT, handoff=3),
            # hand-off into the compact cooperative layout
            _solve(mech, cells, order, tableau, dt, _lib.SOLVE_PERSISTENT, handoff=3, coop_shape=(256, 4))]
    if mech.cooperative_supported(tableau):
        for threads, stage in ((64, -1), (128, 0), (256, 3), (64, 1), (512, -1), (1024, 2), (128, 4), (256, 4),
                              (256, 5), (256, 6), (256, 7)):
            runs.append(_solve(mech, cells, order, tableau, dt, _lib.SOLVE_COOPERATIVE, coop_shape=(threads, stage)))
        # every LU blocking, with the LU tables in global memory (compact) and staged in LDS
        for blocking in (1, 2, 8, -1):
            runs.append(_solve(mech, cells, order, tableau, dt, _lib.SOLVE_COOPERATIVE, coop_shape=(256, 6),
                               lu_blocking=blocking))
            runs.append(_solve(mech, cells, order, tableau, dt, _lib.SOLVE_COOPERATIVE, coop_shape=(128, 2),
                               lu_blocking=blocking))
    for r in runs:
        assert bits_equal(r[0], ref[0])
        assert np.array_equal(r[1], ref[1]) and np.array_equal(r[2], ref[2])
        assert bits_equal(r[3], ref[3])


def test_empty_and_negative_cell_counts():
    """Every kernel entry point accepts zero cells as a no-op that touches no buffer
    and rejects a negative count; a State needs at least one cell."""
    desc = synthetic.readme_mechanism()
    mech = _lib.Mechanism(desc)
    i = mech.info
    sizes = {"k": mech.n_reactions, "y": mech.n_species, "f": mech.n_species, "j": i.nnz_jacobian,
             "l": i.nnz_lower, "u": i.nnz_upper, "c": 3, "p": max(mech.n_custom, 1), "e": 1}
    buf = {key: dev_zeros(64, n) + 7.0 for key, n in sizes.items()}
    p = {key: t.data_ptr() for key, t in buf.items()}
    s = stream()
    calls = [
        lambda n: mech.calculate_rate_constants(n, p["c"], p["p"], p["k"], s),
        lambda n: mech.add_forcing_terms(n, p["k"], p["y"], p["f"], s),
        lambda n: mech.subtract_jacobian_terms(n, p["k"], p["y"], p["j"], s),
        lambda n: mech.alpha_minus_jacobian(n, p["j"], 2.0, s),
        lambda n: mech.lu_decompose(n, p["j"], p["l"], p["u"], s),
        lambda n: mech.lu_solve(n, p["l"], p["u"], p["f"], s),
        lambda n: mech.normalized_error(n, p["y"], p["y"], p["f"], p["y"], 1e-6, p["e"], s),
    ]
    for call in calls:
        call(0)
        with pytest.raises(_lib.MicmHipError):
            call(-1)
    import torch
    torch.cuda.synchronize()
    for key, t in buf.items():
        assert bool((t == 7.0).all()), key
    for n in (0, -5):
        with pytest.raises(_lib.MicmHipError):
            _lib.DeviceState(mech, n)


def test_cooperative_shape_arguments_checked():
    """Launch-shape requests outside the supported set fail with an error; a compact
    request the mechanism cannot take (fewer threads than species) falls back to the
    full layout and still solves."""
    desc, inputs, tableau, dt = CASES["210_560"]
    mech = _lib.Mechanism(desc)
    cells = inputs(desc, 70, seed=3)
    order = mech.species_order()
    state = _lib.DeviceState(mech, cells.n_cells)
    for threads, stage in ((96, -1), (2048, -1), (256, 8), (256, -2)):
        with pytest.raises(_lib.MicmHipError):
            state.set_cooperative_shape(threads, stage)
    for blocking in (3, 16, -2):
        with pytest.raises(_lib.MicmHipError):
            state.set_lu_blocking(blocking)
    ref = _solve(mech, cells, order, tableau, dt, _lib.SOLVE_PLAIN)
    run = _solve(mech, cells, order, tableau, dt, _lib.SOLVE_COOPERATIVE, coop_shape=(128, 4))
    assert bits_equal(run[0], ref[0]) and np.array_equal(run[2], ref[2])


@pytest.mark.parametrize("name", ["60_200", "210_560"])
def test_repeated_solves_in_cost_order_agree_bitwise(name):
    """From the second Solve of a state on, the per-cell forms take cells in the order
    of the previous Solve's attempt counts (most first); results do not change."""
    desc, inputs, tableau, dt = CASES[name]
    mech = _lib.Mechanism(desc)
    cells = inputs(desc, 300 if name == "60_200" else 130, seed=5)
    order = mech.species_order()
    print(f"repeated solves: {name}", flush=True)
    ref = _solve(mech, cells, order, tableau, dt, _lib.SOLVE_PLAIN, n_solves=3)
    runs = [_solve(mech, cells, order, tableau, dt, _lib.SOLVE_PERSISTENT, n_solves=3, handoff=2),
            _solve(mech, cells, order, tableau, dt, _lib.SOLVE_COOPERATIVE, n_solves=3),
            _solve(mech, cells, order, tableau, dt, _lib.SOLVE_COOPERATIVE, n_solves=3, coop_shape=(256, 4))]
    for r in runs:
        assert bits_equal(r[0], ref[0])
        assert np.array_equal(r[1], ref[1]) and np.array_equal(r[2], ref[2])
        assert bits_equal(r[3], ref[3])


@pytest.mark.parametrize("form, coop_shape", [(_lib.SOLVE_PLAIN, (0, -1)), (_lib.SOLVE_PERSISTENT, (0, -1)),
                                              (_lib.SOLVE_COOPERATIVE, (0, -1)), (_lib.SOLVE_PERSISTENT, (128, 4)),
                                              (_lib.SOLVE_COOPERATIVE, (128, 4))])
def test_fused_rate_constants_match_separate_update(form, coop_shape):
    """UpdateStateParameters fused into the Solve launch gives the same rate constants
    and the same solve as the separate call (hand-off and the compact cooperative
    layout, whose rate constants go through global memory, included)."""
    desc, inputs, tableau, dt = CASES["60_200"]
    mech = _lib.Mechanism(desc)
    cells = inputs(desc, 300, seed=17)
    order = mech.species_order()
    print(f"fused rates: form={form} shape={coop_shape}", flush=True)
    ref = _solve(mech, cells, order, tableau, dt, form, handoff=2)
    state = _lib.DeviceState(mech, cells.n_cells)
    state.set_handoff(2)
    state.set_cooperative_shape(*coop_shape)
    state.set_variables(cells.variables[:, order])
    state.set_conditions(cells.conditions)
    state.set_custom_parameters(cells.custom)
    state.solve(dt, _lib.rosenbrock_parameters(tableau), form | _lib.SOLVE_UPDATE_RATES)
    st, stats, ft = state.results()
    assert bits_equal(state.get_variables(), ref[0])
    assert np.array_equal(st, ref[1]) and np.array_equal(stats, ref[2]) and bits_equal(ft, ref[3])
    other = _lib.DeviceState(mech, cells.n_cells)
    other.set_conditions(cells.conditions)
    other.set_custom_parameters(cells.custom)
    other.update_state_parameters()
    assert bits_equal(state.get_rate_constants(), other.get_rate_constants())


@pytest.mark.parametrize("name,n_cells", [("all_types", 130), ("60_200", 130), ("readme", 1024), ("readme", 1500)])
def test_lockstep_solve_bitwise(name, n_cells):
    """One step size for all cells (one multi-cell State): bitwise with the portable-math
    restatement, including the global error norm's summation order.  States of up to
    1,024 cells of a small mechanism (all_types, readme at 1,024) run the whole loop in
    one workgroup; the others launch a kernel per phase."""
    desc, inputs, tableau, dt = CASES[name]
    mech = _lib.Mechanism(desc)
    cells = inputs(desc, n_cells)
    order = mech.species_order()
    print(f"lockstep bitwise: {name} n_cells={n_cells}", flush=True)
    y_gpu, st_gpu, stats_gpu, ft_gpu = _solve(mech, cells, order, tableau, dt, _lib.SOLVE_LOCKSTEP, n_solves=2)
    y_ref, st_ref, stats_ref, ft_ref = OraclePortableMath(desc, tableau).solve(
        cells.conditions, cells.custom, cells.variables[:, order], dt, 2, lockstep=True)
    assert np.array_equal(st_gpu, st_ref)
    assert np.array_equal(stats_gpu, stats_ref)
    assert bits_equal(ft_gpu, ft_ref)
    assert bits_equal(y_gpu, y_ref)


def test_lockstep_solve_matches_oracle():
    """README system, 1000 cells in one State, 10 x 500 s, 3-stage: bit for bit against
    the reference's math."""
    desc = synthetic.readme_mechanism()
    mech = _lib.Mechanism(desc)
    oracle = Oracle(desc, 3)
    cells = synthetic.readme_inputs(1000, seed=1)
    order = mech.species_order()
    print("lockstep: readme 1000 cells", flush=True)
    y_gpu, st_gpu, stats_gpu, _ = _solve(mech, cells, order, 3, 500.0, _lib.SOLVE_LOCKSTEP, n_solves=10)
    y_ref, st_ref, stats_ref, _ = oracle.solve(cells.conditions, cells.custom, cells.variables[:, order], 500.0, 10,
                                               lockstep=True)
    assert np.all(st_gpu == 2) and np.all(st_ref == 2)
    assert np.array_equal(stats_gpu, stats_ref)
    assert bits_equal(y_gpu, y_ref)


@pytest.mark.parametrize("name,n_cells,group", [("60_200", 130, 4), ("readme", 1000, 4), ("all_types", 70, 16),
                                                ("60_200", 130, 64)])
def test_lockstep_solve_vector_matrix_order(name, n_cells, group):
    """A multi-cell State over VectorMatrix<L> sums its error norm group by group; the
    lockstep form in that order equals the oracle's State of the same order bit for bit
    (itself bitwise with the reference's own VectorMatrix<4> State, tests/test_oracle.py),
    and differs from the Matrix<Real> order."""
    desc, inputs, tableau, dt = CASES[name]
    mech = _lib.Mechanism(desc)
    cells = inputs(desc, n_cells)
    order = mech.species_order()
    print(f"lockstep VectorMatrix<{group}> order: {name} n_cells={n_cells}", flush=True)
    y_gpu, st_gpu, stats_gpu, ft_gpu = _solve(mech, cells, order, tableau, dt, _lib.SOLVE_LOCKSTEP, n_solves=2,
                                              lockstep_order=group)
    y_ref, st_ref, stats_ref, ft_ref = OraclePortableMath(desc, tableau).solve(
        cells.conditions, cells.custom, cells.variables[:, order], dt, 2, lockstep=group)
    assert np.array_equal(st_gpu, st_ref) and np.array_equal(stats_gpu, stats_ref)
    assert bits_equal(ft_gpu, ft_ref) and bits_equal(y_gpu, y_ref)


@pytest.mark.parametrize("n_cells", [1, 64, 301, 5000])
@pytest.mark.parametrize("group", [1, 4, 64, 100])
def test_state_normalized_error(n_cells, group):
    """micm_hip_normalized_error_state: one scalar per State, bit for bit with the
    oracle's norm in the same order (Matrix<Real> for 1: the reference's own norm,
    tests/test_oracle.py)."""
    desc = synthetic.mechanism_60_200()
    mech = _lib.Mechanism(desc)
    o = OraclePortableMath(desc, 4)
    rng = np.random.default_rng(n_cells + group)
    N = mech.n_species
    y = rng.lognormal(-8, 3, (n_cells, N))
    ynew = y * (1 + 1e-3 * rng.standard_normal((n_cells, N)))
    err = 1e-6 * rng.standard_normal((n_cells, N)) * y
    atol = np.full(N, 1e-12)
    e_dev = torch_zeros(1)
    bufs = [dev(a) for a in (y, ynew, err)]
    atol_dev = torch_tensor(atol)
    mech.normalized_error_state(n_cells, *[b.data_ptr() for b in bufs], atol_dev.data_ptr(), 1e-4, group,
                                e_dev.data_ptr(), stream())
    e_gpu = float(e_dev.cpu().numpy()[0])
    e_ref = o.normalized_error_grouped(y, ynew, err, 1e-4, atol, group)
    assert bits_equal(np.array([e_gpu]), np.array([e_ref]))
    if group == 1:
        assert e_ref == o.normalized_error(y, ynew, err, 1e-4, atol)


def torch_zeros(n):
    import torch
    return torch.zeros(n, dtype=torch.float64, device="cuda")


def torch_tensor(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda")


ALL_FORMS = [_lib.SOLVE_PLAIN, _lib.SOLVE_PERSISTENT, _lib.SOLVE_COOPERATIVE, _lib.SOLVE_LOCKSTEP]


@pytest.mark.parametrize("tableau", [2, 3, 4, 41, 61])
@pytest.mark.parametrize("form", ALL_FORMS)
def test_every_tableau_matches_reference_math(tableau, form):
    """The five presets (2-, 3-, 4-stage, 4-stage DA, 6-stage DA; fields equal the
    reference's, tests/test_host.py) drive every form to the C-library oracle's result
    bit for bit: 60/200, 130 cells, 2 x 300 s (lockstep: one State of all cells)."""
    desc = synthetic.mechanism_60_200()
    mech = _lib.Mechanism(desc)
    if form == _lib.SOLVE_COOPERATIVE and not mech.cooperative_supported(6 if tableau == 61 else 4):
        pytest.skip("mechanism does not fit the cooperative form")
    cells = synthetic.inputs_60_200(desc, 130, seed=41)
    order = mech.species_order()
    print(f"tableau {tableau} form {form}", flush=True)
    y, st, stats, ft = _solve(mech, cells, order, tableau, 300.0, form, n_solves=2)
    y_ref, st_ref, stats_ref, ft_ref = Oracle(desc, tableau).solve(
        cells.conditions, cells.custom, cells.variables[:, order], 300.0, 2, lockstep=form == _lib.SOLVE_LOCKSTEP)
    assert np.array_equal(st, st_ref) and np.array_equal(stats, stats_ref)
    assert bits_equal(ft, ft_ref) and bits_equal(y, y_ref)


@pytest.mark.parametrize("name", ["60_200", "210_560", "ts1"])
@pytest.mark.parametrize("form", [_lib.SOLVE_PERSISTENT, _lib.SOLVE_COOPERATIVE])
def test_fused_rate_constants_match_reference_math(name, form):
    """UpdateStateParameters fused into the Solve launch (MICM_HIP_SOLVE_UPDATE_RATES):
    rate constants, outcomes, counters and concentrations equal the C-library oracle's."""
    desc, inputs, tableau, dt = CASES[name]
    mech = _lib.Mechanism(desc)
    if form == _lib.SOLVE_COOPERATIVE and not mech.cooperative_supported(tableau):
        pytest.skip("mechanism does not fit the cooperative form")
    cells = inputs(desc, 150)
    order = mech.species_order()
    print(f"fused rates vs oracle: {name} form {form}", flush=True)
    state = _lib.DeviceState(mech, cells.n_cells)
    state.set_variables(cells.variables[:, order])
    state.set_conditions(cells.conditions)
    state.set_custom_parameters(cells.custom)
    params = _lib.rosenbrock_parameters(tableau)
    state.solve(dt, params, form | _lib.SOLVE_UPDATE_RATES)
    o = Oracle(desc, tableau)
    y_ref, st_ref, stats_ref, ft_ref = o.solve(cells.conditions, cells.custom, cells.variables[:, order], dt, 1)
    st, stats, ft = state.results()
    assert bits_equal(state.get_rate_constants(), o.rate_constants(cells.conditions, cells.custom))
    assert np.array_equal(st, st_ref) and np.array_equal(stats, stats_ref)
    assert bits_equal(ft, ft_ref) and bits_equal(state.get_variables(), y_ref)


@pytest.mark.parametrize("form", [_lib.SOLVE_COOPERATIVE, _lib.SOLVE_PERSISTENT])
def test_literal_210_species_mechanism_matches_reference_math(form):
    """The 210/560 configuration with its products drawn from every species (the text taken
    literally: cyclic, with a denser factor -- 13,996 in-place entries against 8,548): every
    cell, converged or not, takes the reference's exact step sequence.  Most of its cells
    fail in the reference too; the test checks that some converge and some do not."""
    desc = synthetic.mechanism_210_560(products_from="all")
    mech = _lib.Mechanism(desc)
    cells = synthetic.inputs_210_560(desc, 48)
    order = mech.species_order()
    print(f"literal 210/560: form {form}", flush=True)
    y_gpu, st_gpu, stats_gpu, ft_gpu = _solve(mech, cells, order, 4, 900.0, form)
    y_ref, st_ref, stats_ref, ft_ref = Oracle(desc, 4).solve(cells.conditions, cells.custom,
                                                              cells.variables[:, order], 900.0, 1)
    print(f"  converged {np.mean(st_ref == 2):.3f}", flush=True)
    assert np.any(st_ref == 2) and np.any(st_ref != 2)
    assert np.array_equal(st_gpu, st_ref) and np.array_equal(stats_gpu, stats_ref)
    assert bits_equal(ft_gpu, ft_ref) and bits_equal(y_gpu, y_ref)


def test_ten_step_sequence_of_the_60_species_workload():
    """The 60/200 workload's sequence -- 10 consecutive Solves of 300 s, rate constants
    updated before each -- on 515 cells, cooperative and persistent forms: every cell bit
    for bit against the C-library oracle after all ten."""
    desc = synthetic.mechanism_60_200()
    mech = _lib.Mechanism(desc)
    cells = synthetic.inputs_60_200(desc, 515)
    order = mech.species_order()
    y_ref, st_ref, stats_ref, ft_ref = Oracle(desc, 4).solve(cells.conditions, cells.custom,
                                                             cells.variables[:, order], 300.0, 10)
    for form in (_lib.SOLVE_COOPERATIVE, _lib.SOLVE_PERSISTENT):
        print(f"10 x 300 s: form {form}", flush=True)
        y, st, stats, ft = _solve(mech, cells, order, 4, 300.0, form, n_solves=10)
        assert np.array_equal(st, st_ref) and np.array_equal(stats, stats_ref)
        assert bits_equal(ft, ft_ref) and bits_equal(y, y_ref)


def _forced(case):
    """(mechanism, cells, y0 mutator, rtol, atol, parameter overrides, expected state) of
    the inputs that force each of the reference's failure states (solver_result.hpp:13-35)."""
    readme = synthetic.readme_mechanism()
    if case == "nan":
        def mutate(y):
            y[1, 0] = np.nan
            return y
        return readme, synthetic.readme_inputs(70, seed=3), mutate, None, None, {}, 6
    if case == "inf":  # the scaled errors overflow
        return readme, synthetic.readme_inputs(70, seed=3), None, 1e-300, 1e-300, {}, 7
    if case == "max_steps":
        return synthetic.mechanism_60_200(), synthetic.inputs_60_200(synthetic.mechanism_60_200(), 70), None, \
            None, None, {"max_number_of_steps": 1}, 3
    # step size below the round-off
    return readme, synthetic.readme_inputs(70, seed=3), None, None, None, {"h_max": 1e-300}, 4


@pytest.mark.parametrize("case", ["nan", "inf", "max_steps", "step_size"])
@pytest.mark.parametrize("form", ALL_FORMS)
def test_forced_failure_states_match_reference_math(case, form):
    """NaNDetected (a NaN initial concentration), InfDetected (tolerances of 1e-300, whose
    scaled errors overflow), ConvergenceExceededMaxSteps (max_number_of_steps = 1) and
    StepSizeTooSmall (h_max below the round-off): every form reports the oracle's state,
    counters and concentrations for every cell."""
    desc, cells, mutate, rtol, atol, overrides, expected = _forced(case)
    mech = _lib.Mechanism(desc)
    order = mech.species_order()
    y0 = np.array(cells.variables[:, order])
    if mutate:
        y0 = mutate(y0)
    print(f"forced {case}: form {form}", flush=True)
    params = _lib.rosenbrock_parameters(4 if desc.n_species > 3 else 3)
    for key, value in overrides.items():
        setattr(params, key, value)
    state = _lib.DeviceState(mech, cells.n_cells)
    state.set_variables(y0)
    state.set_conditions(cells.conditions)
    state.set_custom_parameters(cells.custom)
    if rtol is not None:
        state.set_tolerances(rtol, np.full(mech.n_species, atol))
    state.update_state_parameters()
    state.solve(300.0 if desc.n_species > 3 else 500.0, params, form)
    st, stats, ft = state.results()
    o = Oracle(desc, int(params.stages))
    o.set_control(int(params.max_number_of_steps), params.h_min, params.h_max, params.h_start)
    y_ref, st_ref, stats_ref, ft_ref = o.solve(
        cells.conditions, cells.custom, y0, 300.0 if desc.n_species > 3 else 500.0, 1,
        lockstep=form == _lib.SOLVE_LOCKSTEP, rtol=rtol if rtol is not None else 1e-6,
        atol=None if atol is None else np.full(mech.n_species, atol))
    assert np.any(st_ref == expected)
    assert np.array_equal(st, st_ref) and np.array_equal(stats, stats_ref)
    assert bits_equal(ft, ft_ref) and bits_equal(state.get_variables(), y_ref)
