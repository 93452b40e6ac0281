"""The portable exp / log / log10 / pow of micm_amd/csrc/pmath.h (host build, via
the portable-math oracle library).

They restate the C library's exp / log / pow / log10 operation for operation, so
they are checked bit for bit against it (Python's math module calls it) on the
whole domain and on the ranges the path uses.  Independently of any C library,
accuracy is measured against values computed to 60 significant digits with
Python's `decimal` module: a dropped series term, a wrong reduction constant or a
wrong scaling exponent each push the error far beyond these bounds.  The device
build of the same header is pinned to this host build bit for bit by the GPU
rate-constant and whole-solve tests."""
import ctypes
import decimal
import math
from decimal import Decimal

import numpy as np
import pytest

from micm_amd import synthetic
from oracle import ORACLE_PM_LIB, Oracle, OraclePortableMath

decimal.getcontext().prec = 60
_dp = ctypes.POINTER(ctypes.c_double)


def _lib():
    lib = ctypes.CDLL(ORACLE_PM_LIB)
    lib.oracle_pmath.argtypes = [ctypes.c_int32, ctypes.c_int64, _dp, _dp, _dp]
    lib.oracle_pmath.restype = ctypes.c_int
    return lib


def pm(which, x, y=None):
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    y = np.ascontiguousarray(np.atleast_1d(y if y is not None else x), dtype=np.float64)
    out = np.zeros_like(x)
    assert _lib().oracle_pmath({"exp": 0, "log": 1, "log10": 2, "pow": 3}[which], len(x), x.ctypes.data_as(_dp),
                               y.ctypes.data_as(_dp), out.ctypes.data_as(_dp)) == 0
    return out


def ulp_error(got, exact):
    e = float(exact)
    return float(abs(Decimal(float(got)) - exact) / Decimal(math.ulp(e)))


_LN10 = Decimal(10).ln()


def test_exp_accuracy():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(-745, 709, 1500), rng.uniform(-1, 1, 500), rng.uniform(-1e-8, 1e-8, 100)])
    got = pm("exp", x)
    worst = max(ulp_error(g, Decimal(v).exp()) for g, v in zip(got, x) if abs(v) < 708)
    assert worst <= 0.52, worst


def test_log_and_log10_accuracy():
    rng = np.random.default_rng(2)
    x = np.concatenate([np.exp(rng.uniform(-700, 700, 1500)), rng.uniform(0.5, 2.0, 500), [5e-324, 1e-310]])
    got = pm("log", x)
    worst = max(ulp_error(g, Decimal(v).ln()) for g, v in zip(got, x) if v != 1.0)
    assert worst <= 0.52, worst
    got = pm("log10", x)
    worst = max(ulp_error(g, Decimal(v).ln() / _LN10) for g, v in zip(got, x) if v != 1.0)
    assert worst <= 1.5, worst
    powers = np.array([10.0 ** k for k in range(23)])
    assert np.array_equal(pm("log10", powers), np.arange(23.0))


def test_pow_accuracy_in_rate_constant_ranges():
    """Relative error of pow over the arguments the rate constants and the step
    controller produce: (T/300)^B with B in [-10, 2], Fc^(...) and error^(-1/ELO)."""
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(0.5, 1.2, 1000), rng.uniform(0.1, 1.0, 300), np.exp(rng.uniform(-23, 23, 300))])
    y = np.concatenate([rng.uniform(-10, 2, 1000), rng.uniform(0, 1, 300), rng.choice([-1 / 3, -1 / 4, -1 / 2], 300)])
    got = pm("pow", x, y)
    worst = max(ulp_error(g, (Decimal(b) * Decimal(a).ln()).exp()) for g, a, b in zip(got, x, y))
    assert worst <= 0.6, worst


def _libm_sample(rng, n):
    """Arguments over the whole domain and over the ranges the path uses: exp(C/T),
    (T/D)^B, Fc^(N/(N + log10^2)), error^(1/ELO), log10 of falloff ratios; random
    bit patterns (subnormals, infinities and NaNs included) and special values."""
    bits = rng.integers(0, 2 ** 63, n, dtype=np.int64).view(np.float64) * rng.choice([-1.0, 1.0], n)
    special = np.array([0.0, -0.0, 1.0, -1.0, 2.0, 0.5, math.inf, -math.inf, math.nan, 5e-324, -5e-324,
                        2.2250738585072014e-308, 1e308, 709.78, -745.13, -708.4, 3.0, -3.0, 1e-300, 10.0])
    return bits, special


def _same(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return (a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b))


def test_bitwise_equal_to_the_c_library():
    """exp, log, log10 and pow return exactly what the C library returns -- the
    reference's own arithmetic (std::exp / std::pow / std::log10 on this platform)."""
    rng = np.random.default_rng(4)
    n = 60000
    bits, special = _libm_sample(rng, n)
    for name, fn, args in [
        ("exp", math.exp, [bits, rng.uniform(-750, 750, n), rng.uniform(-2000, 500, n) / rng.uniform(180, 320, n),
                           special]),
        ("log", math.log, [np.abs(bits), np.exp(rng.uniform(-700, 700, n)), rng.uniform(0.9, 1.1, n), special]),
        ("log10", math.log10, [np.abs(bits), np.exp(rng.uniform(-60, 60, n)), special]),
    ]:
        x = np.concatenate(args)
        want = np.empty_like(x)
        for i, v in enumerate(x):
            try:
                want[i] = fn(v)
            except (ValueError, OverflowError):  # CPython raises where the C library returns NaN / inf
                want[i] = np.nan
        got = pm(name, x)
        ok = _same(got, want) | np.isnan(want)
        assert ok.all(), (name, x[~ok][:5], got[~ok][:5], want[~ok][:5])
    a = np.concatenate([np.abs(bits[:n // 2]), rng.uniform(0.5, 1.2, n), np.exp(rng.uniform(-23, 14, n)),
                        rng.uniform(0.1, 1.0, n), -np.exp(rng.uniform(-5, 5, n))])
    b = np.concatenate([bits[n // 2:], rng.uniform(-10, 2, n), rng.choice([0.25, 1 / 3, 0.5, 1 / 6], n),
                        1.0 / (1.0 + rng.uniform(0, 30, n)), np.round(rng.uniform(-30, 30, n))])
    want = np.empty_like(a)
    for i, (p, q) in enumerate(zip(a, b)):
        try:
            want[i] = math.pow(p, q)
        except (ValueError, OverflowError):
            want[i] = np.nan
    got = pm("pow", a, b)
    ok = _same(got, want) | np.isnan(want)
    assert ok.all(), (a[~ok][:5], b[~ok][:5], got[~ok][:5], want[~ok][:5])
    sa, sb = np.repeat(special, len(special)), np.tile(special, len(special))
    want = np.empty_like(sa)
    for i, (p, q) in enumerate(zip(sa, sb)):
        try:
            want[i] = math.pow(p, q)
        except (ValueError, OverflowError):
            want[i] = np.nan
    got = pm("pow", sa, sb)
    assert (_same(got, want) | np.isnan(want)).all()


def test_special_values():
    inf, nan = math.inf, math.nan
    assert pm("exp", 0.0)[0] == 1.0
    assert pm("exp", 710.0)[0] == inf and pm("exp", -746.0)[0] == 0.0
    assert math.isnan(pm("exp", nan)[0])
    assert 0.0 < pm("exp", -740.0)[0] < 1e-320  # subnormal result
    assert pm("exp", 709.7)[0] == pytest.approx(math.exp(709.7), rel=1e-15)
    assert pm("log", 1.0)[0] == 0.0 and pm("log", 0.0)[0] == -inf and pm("log", inf)[0] == inf
    assert math.isnan(pm("log", -1.0)[0])
    assert pm("log", 2.0)[0] == pytest.approx(math.log(2.0), rel=2e-16)
    assert pm("log10", 1000.0)[0] == pytest.approx(3.0, rel=1e-15)
    cases = [((2.0, 0.0), 1.0), ((1.0, nan), 1.0), ((-2.0, 3.0), -8.0), ((-2.0, 2.0), 4.0), ((0.0, 2.5), 0.0),
             ((0.0, -1.0), inf), ((-0.0, -3.0), -inf), ((inf, -1.0), 0.0), ((0.5, inf), 0.0), ((2.0, -inf), 0.0),
             ((3.0, 1.0), 3.0), ((1.5, 2.0), 2.25)]
    for (a, b), want in cases:
        got = pm("pow", a, b)[0]
        assert got == want and math.copysign(1, got) == math.copysign(1, want), (a, b, got)
    assert math.isnan(pm("pow", -2.0, 0.5)[0])
    assert pm("pow", -2.0, 3.0)[0] == -8.0


def test_portable_oracle_equals_c_library_oracle():
    """The two restatements differ only in where exp / log10 / pow come from, so rate
    constants and whole solves (step sequences, outcomes, final states) are bit for
    bit the same."""
    desc = synthetic.mechanism_60_200()
    cells = synthetic.inputs_60_200(desc, 257)
    a, b = Oracle(desc, 4), OraclePortableMath(desc, 4)
    ka, kb = a.rate_constants(cells.conditions, cells.custom), b.rate_constants(cells.conditions, cells.custom)
    assert np.array_equal(ka.view(np.int64), kb.view(np.int64))
    y = cells.variables[:, a.species_order()]
    assert np.array_equal(a.forcing(ka, y), b.forcing(ka, y))
    assert np.array_equal(a.jacobian(ka, y), b.jacobian(ka, y))
    for desc, inputs, tableau, dt, n, ns in [(desc, synthetic.inputs_60_200, 4, 300.0, 300, 10),
                                             (synthetic.mechanism_210_560(), synthetic.inputs_210_560, 4, 900.0, 64, 1)]:
        a, b = Oracle(desc, tableau), OraclePortableMath(desc, tableau)
        cells = inputs(desc, n, seed=29)
        y0 = cells.variables[:, a.species_order()]
        ra = a.solve(cells.conditions, cells.custom, y0, dt, ns)
        rb = b.solve(cells.conditions, cells.custom, y0, dt, ns)
        assert np.array_equal(ra[1], rb[1]) and np.array_equal(ra[2], rb[2])
        assert np.array_equal(ra[0].view(np.int64), rb[0].view(np.int64))
        assert np.array_equal(ra[3].view(np.int64), rb[3].view(np.int64))
