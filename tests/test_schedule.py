"""The cooperative form's level schedules (LU, forward and back substitution),
executed on the CPU with every level reading the values as they stood when the
level began: they must reproduce the sequential Doolittle factorisation and
substitution of the oracle bit for bit.  A task scheduled at the level of one of
its inputs would read a stale value and fail this."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from helpers import bits_equal
from mechanisms import all_rate_types, ts1_inputs, ts1_mechanism
from micm_amd import synthetic
from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("harness") / "libschedule_harness.so")
    subprocess.run(["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-ffp-contract=off",
                    os.path.join(ROOT, "tests", "schedule_harness.cpp"),
                    os.path.join(ROOT, "micm_amd", "csrc", "mechanism.cpp"), "-o", out], check=True)
    lib = ctypes.CDLL(out)
    lib.harness_sizes.argtypes = [ctypes.c_void_p, _ip]
    lib.harness_factor_solve.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp, _dp, _dp]
    lib.harness_factor_blocked.argtypes = [ctypes.c_void_p, ctypes.c_int32, _dp, ctypes.c_double, _dp, _dp]
    lib.harness_factor_capped.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp, _dp,
                                          ctypes.POINTER(ctypes.c_int32)]
    lib.harness_blocked_conflict.argtypes = [ctypes.c_void_p, ctypes.c_int32]
    lib.harness_solve_swept.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp]
    lib.harness_chain_check.argtypes = [ctypes.c_void_p]
    return lib


CASES = {
    "readme": (synthetic.readme_mechanism, lambda d, n: synthetic.readme_inputs(n)),
    "all_types": (all_rate_types, None),
    "20_50": (synthetic.mechanism_20_50, synthetic.inputs_60_200),
    "60_200": (synthetic.mechanism_60_200, synthetic.inputs_60_200),
    "210_560": (synthetic.mechanism_210_560, synthetic.inputs_210_560),
    "210_560_literal": (lambda: synthetic.mechanism_210_560(products_from="all"), synthetic.inputs_210_560),
    "ts1": (ts1_mechanism, ts1_inputs),
}


@pytest.mark.parametrize("name", list(CASES))
def test_level_schedules_reproduce_doolittle(harness, name):
    make, inputs = CASES[name]
    desc = make()
    cdesc = desc.as_c()
    sizes = np.zeros(9, dtype=np.int64)
    harness.harness_sizes(ctypes.byref(cdesc), sizes.ctypes.data_as(_ip))
    n, nnz_j, nnz_alu, nnz_l, nnz_u, lv_lu, lv_f, lv_b, ok = (int(v) for v in sizes)
    assert ok == 1
    # levels never exceed the sequential step counts
    assert lv_f <= n and lv_b <= n and lv_lu <= nnz_alu
    o = Oracle(desc, 4)
    rng = np.random.default_rng(7)
    if inputs is None:
        y = rng.uniform(1e-6, 1.0, (3, n))
        cond = np.tile([280.0, 9e4, 38.0], (3, 1))
        custom = rng.uniform(1e-6, 1e-2, (3, o.n_custom))
    else:
        cells = inputs(desc, 3)
        y, cond, custom = cells.variables[:, o.species_order()], cells.conditions, cells.custom
    k = o.rate_constants(cond, custom)
    jac = o.jacobian(k, y)
    f = o.forcing(k, y)
    # (the literal 210/560 mechanism, the largest, on one cell: each harness call compiles it)
    for c in range(1 if name == "210_560_literal" else 3):
        alpha = [1.0 / (300.0 * 0.57282), 3.7e2, 1e-3][c]
        _, lo_ref, up_ref = o.lu(jac[c:c + 1], alpha)
        x_ref = o.lu_solve(lo_ref, up_ref, f[c:c + 1])
        lo, up, x = np.zeros(nnz_l), np.zeros(nnz_u), f[c].copy()
        harness.harness_factor_solve(ctypes.byref(cdesc), np.ascontiguousarray(jac[c]).ctypes.data_as(_dp), alpha,
                                     lo.ctypes.data_as(_dp), up.ctypes.data_as(_dp), x.ctypes.data_as(_dp))
        assert bits_equal(lo, lo_ref[0]) and bits_equal(up, up_ref[0])
        assert bits_equal(x, x_ref[0])
        # the forward substitution's last levels as the single-wavefront column sweep
        xs = f[c].copy()
        first = harness.harness_solve_swept(ctypes.byref(cdesc), np.ascontiguousarray(jac[c]).ctypes.data_as(_dp),
                                            alpha, xs.ctypes.data_as(_dp))
        if name in ("210_560", "60_200", "ts1"):
            assert first > 0, name  # these mechanisms sweep the forward tail
        if first >= 0:
            assert bits_equal(xs, x_ref[0])
        if c == 0:
            # the back substitution's task chain reproduces its levels' tasks, and no
            # product formed a task ahead reads a value the previous task writes
            assert harness.harness_chain_check(ctypes.byref(cdesc)) == 0, name
        # every LU blocking (shortest segment 1, 2, 4, 8, whole sums) gives the same factors;
        # the blocking changes how a sum is split over the levels, not the level at which
        # the element is final, so every blocking has as many levels
        levels = []
        for b in range(5):
            lo_b, up_b = np.zeros(nnz_l), np.zeros(nnz_u)
            nl = harness.harness_factor_blocked(ctypes.byref(cdesc), b,
                                                np.ascontiguousarray(jac[c]).ctypes.data_as(_dp), alpha,
                                                lo_b.ctypes.data_as(_dp), up_b.ctypes.data_as(_dp))
            assert nl > 0
            levels.append(nl)
            assert bits_equal(lo_b, lo_ref[0]) and bits_equal(up_b, up_ref[0])
        assert len(set(levels)) == 1
        # the schedules re-levelled to at most 64 ... 512 tasks per level: same factors,
        # no level above its cap, no two tasks of a level touching one position, and at
        # least as many levels as the uncapped schedule
        if c == 0:
            for b in range(5):
                assert harness.harness_blocked_conflict(ctypes.byref(cdesc), b) == 0
        slots = 5 * 4
        lo_c, up_c = np.zeros((slots, nnz_l)), np.zeros((slots, nnz_u))
        nl_c = np.zeros(slots, dtype=np.int32)
        assert harness.harness_factor_capped(ctypes.byref(cdesc), np.ascontiguousarray(jac[c]).ctypes.data_as(_dp),
                                             alpha, lo_c.ctypes.data_as(_dp), up_c.ctypes.data_as(_dp),
                                             nl_c.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))) == slots
        for k in range(slots):
            assert nl_c[k] == 0 or nl_c[k] >= levels[0], (k, nl_c[k])
            if nl_c[k] > 0:
                assert bits_equal(lo_c[k], lo_ref[0]) and bits_equal(up_c[k], up_ref[0]), k
        if name == "210_560":
            assert (nl_c > 0).all()  # every cap binds somewhere for this mechanism
        assert harness.harness_factor_blocked(ctypes.byref(cdesc), 5, np.ascontiguousarray(jac[c]).ctypes.data_as(_dp),
                                              alpha, lo.ctypes.data_as(_dp), up.ctypes.data_as(_dp)) == -1
