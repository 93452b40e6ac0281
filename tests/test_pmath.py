"""The portable exp / log / log10 / pow of micm_amd/csrc/pmath.h (host build, via
the portable-math oracle library).

Accuracy is measured against values computed to 60 significant digits with
Python's `decimal` module (correctly rounded exp and ln), independently of any C
library.  A dropped series term, a wrong reduction constant or a wrong scaling
exponent each push the error far beyond these bounds.  The device build of the
same header is pinned to this host build bit for bit by the GPU rate-constant
tests."""
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


def test_agreement_with_the_c_library():
    """The reference computes rate constants with the C library; where that library is
    accurate to about half an ulp (glibc's exp / log / pow, and log10 by the fdlibm
    method), these functions return its results in nearly every case."""
    rng = np.random.default_rng(4)
    x = rng.uniform(-20.0, 5.0, 20000)
    assert np.mean(pm("exp", x) == np.array([math.exp(v) for v in x])) >= 0.99
    x = np.exp(rng.uniform(-30.0, 30.0, 20000))
    assert np.mean(pm("log", x) == np.array([math.log(v) for v in x])) >= 0.99
    assert np.mean(pm("log10", x) == np.array([math.log10(v) for v in x])) >= 0.99
    a, b = rng.uniform(0.6, 1.1, 20000), rng.uniform(-5.0, 1.0, 20000)
    assert np.mean(pm("pow", a, b) == np.array([math.pow(p, q) for p, q in zip(a, b)])) >= 0.99


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


def test_portable_oracle_differs_only_in_rate_constant_bits():
    """The two restatements share every step but exp / log10 / pow: their rate
    constants agree to a few ulp, and given the same rate constants the forcing,
    Jacobian and LU are bit for bit the same."""
    desc = synthetic.mechanism_60_200()
    cells = synthetic.inputs_60_200(desc, 257)
    a, b = Oracle(desc, 4), OraclePortableMath(desc, 4)
    ka, kb = a.rate_constants(cells.conditions, cells.custom), b.rate_constants(cells.conditions, cells.custom)
    nz = ka != 0
    assert np.array_equal(nz, kb != 0)
    assert np.max(np.abs(ka - kb)[nz] / np.abs(ka)[nz]) <= 4e-15
    assert np.mean(ka == kb) >= 0.99
    y = cells.variables[:, a.species_order()]
    assert np.array_equal(a.forcing(ka, y), b.forcing(ka, y))
    assert np.array_equal(a.jacobian(ka, y), b.jacobian(ka, y))
