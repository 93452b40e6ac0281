"""TEST INFRASTRUCTURE ONLY -- the oracle side.

Two CPU solvers with one ctypes interface:

* `Oracle`     -- liboracle.so, this project's CPU restatement of the
                  reference's Rosenbrock path (oracle/micm_oracle.cpp).  It is
                  the checker on the GPU box and the "port" CPU baseline.
* `OraclePortableMath` -- liboracle_pm.so, the restatement with the device
                  path's own exp / log10 / pow: the checker for whole solves.
* `Reference`  -- oracle/_ref/libmicm_ref.so, the reference's own header-only
                  solver compiled from its sources (oracle/ref_driver.cpp).  It
                  exists only where the reference headers were available at
                  build time and pins the restatement.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this package, and only as the checker or the CPU baseline.  Every per-cell
array is cell-major in the solver's state order ([cell][state index]).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
ORACLE_PM_LIB = os.path.join(HERE, "liboracle_pm.so")
REF_LIB = os.path.join(HERE, "_ref", "libmicm_ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else _dp()


def _i(a):
    return a.ctypes.data_as(_ip) if a is not None else _ip()


def _i32(a):
    return a.ctypes.data_as(_i32p) if a is not None else _i32p()


def reference_available():
    return os.path.exists(REF_LIB)


class _Base:
    prefix = None
    path = None

    _libs = {}

    @classmethod
    def _lib(cls):
        if cls.path not in cls._libs:
            if not os.path.exists(cls.path):
                raise FileNotFoundError(f"{cls.path} is not built (python -c 'import __graft_entry__ as g; g.build()')")
            lib = ctypes.CDLL(cls.path)
            p = cls.prefix
            getattr(lib, p + "solver_create").restype = _vp
            getattr(lib, p + "solver_create").argtypes = [_vp, ctypes.c_int32]
            getattr(lib, p + "solver_destroy").argtypes = [_vp]
            getattr(lib, p + "last_error").restype = ctypes.c_char_p
            sig = {
                "sizes": [_vp, _ip],
                "species_order": [_vp, _ip],
                "patterns": [_vp, _ip, _ip, _ip, _ip, _ip, _ip],
                "rate_constants": [_vp, ctypes.c_int64, _dp, _dp, _dp],
                "forcing": [_vp, ctypes.c_int64, _dp, _dp, _dp],
                "jacobian": [_vp, ctypes.c_int64, _dp, _dp, _dp],
                "add_forcing": [_vp, ctypes.c_int64, _dp, _dp, _dp],
                "subtract_jacobian": [_vp, ctypes.c_int64, _dp, _dp, _dp],
                "lu": [_vp, ctypes.c_int64, _dp, ctypes.c_double, _dp, _dp, _dp],
                "lu_solve": [_vp, ctypes.c_int64, _dp, _dp, _dp],
                "normalized_error": [_vp, ctypes.c_int64, _dp, _dp, _dp, ctypes.c_double, _dp, _dp],
                "solve": [_vp, ctypes.c_int64, ctypes.c_int32, _dp, _dp, _dp, ctypes.c_double, ctypes.c_int64,
                          ctypes.c_double, _dp, _i32p, _ip, _dp],
            }
            if p == "oracle_":
                sig.update({
                    "solve_record": [_vp, ctypes.c_int64, _dp, _dp, ctypes.c_double, ctypes.c_int64, ctypes.c_double,
                                     _dp, _i32p, _ip, ctypes.c_int64, _dp, _ip],
                    "replay": [_vp, ctypes.c_int64, _dp, _dp, ctypes.c_int64, _dp, _ip],
                    "adjoint": [_vp, ctypes.c_int64, _dp, ctypes.c_int64, _dp, _ip, _dp, _dp, _dp],
                    "tangent": [_vp, ctypes.c_int64, _dp, _dp, ctypes.c_int64, _dp, _ip, ctypes.c_int64, _dp, _dp,
                                _dp],
                    "normalized_error_grouped": [_vp, ctypes.c_int64, _dp, _dp, _dp, ctypes.c_double, _dp,
                                                 ctypes.c_int64, _dp],
                })
            for name, args in sig.items():
                f = getattr(lib, p + name)
                f.argtypes = args
                f.restype = ctypes.c_int
            cls._libs[cls.path] = lib
        return cls._libs[cls.path]

    def __init__(self, desc, tableau=3):
        self.desc = desc
        self.lib = self._lib()
        self.h = getattr(self.lib, self.prefix + "solver_create")(ctypes.byref(desc.as_c()), int(tableau))
        if not self.h:
            raise RuntimeError(getattr(self.lib, self.prefix + "last_error")().decode())
        sizes = np.zeros(6, dtype=np.int64)
        self._call("sizes", _i(sizes))
        self.n_species, self.n_reactions, self.n_custom, self.nnz_jacobian, self.nnz_lower, self.nnz_upper = \
            (int(x) for x in sizes)

    def __del__(self):
        if getattr(self, "h", None):
            getattr(self.lib, self.prefix + "solver_destroy")(self.h)
            self.h = None

    def _call(self, name, *args):
        rc = getattr(self.lib, self.prefix + name)(self.h, *args)
        if rc != 0:
            raise RuntimeError(f"{self.prefix}{name}: " + getattr(self.lib, self.prefix + "last_error")().decode())

    # ---------------------------------------------------------------- symbolic
    def species_order(self):
        """state_to_declared[i] = declared species index held in state column i."""
        out = np.zeros(self.n_species, dtype=np.int64)
        self._call("species_order", _i(out))
        return out

    def patterns(self):
        arrs = [np.zeros(n, dtype=np.int64) for n in (self.nnz_jacobian, self.nnz_jacobian, self.nnz_lower,
                                                     self.nnz_lower, self.nnz_upper, self.nnz_upper)]
        self._call("patterns", *[_i(a) for a in arrs])
        return {"jacobian": (arrs[0], arrs[1]), "lower": (arrs[2], arrs[3]), "upper": (arrs[4], arrs[5])}

    # ---------------------------------------------------------------- kernels
    def rate_constants(self, conditions, custom):
        conditions = np.ascontiguousarray(conditions, dtype=np.float64)
        n = conditions.shape[0]
        custom = np.ascontiguousarray(custom, dtype=np.float64).reshape(n, -1)
        k = np.zeros((n, self.n_reactions))
        self._call("rate_constants", n, _d(conditions), _d(custom), _d(k))
        return k

    def forcing(self, k, y, f0=None):
        """f(y) (Fill(0) then AddForcingTerms); with f0, AddForcingTerms into a copy of f0."""
        k = np.ascontiguousarray(k, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if f0 is None:
            f = np.zeros_like(y)
            self._call("forcing", y.shape[0], _d(k), _d(y), _d(f))
        else:
            f = np.array(f0, dtype=np.float64, copy=True, order="C")
            self._call("add_forcing", y.shape[0], _d(k), _d(y), _d(f))
        return f

    def jacobian(self, k, y, jac0=None):
        """-J(y) (Fill(0) then SubtractJacobianTerms); with jac0, SubtractJacobianTerms into
        a copy of jac0."""
        k = np.ascontiguousarray(k, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if jac0 is None:
            jac = np.zeros((y.shape[0], self.nnz_jacobian))
            self._call("jacobian", y.shape[0], _d(k), _d(y), _d(jac))
        else:
            jac = np.array(jac0, dtype=np.float64, copy=True, order="C")
            self._call("subtract_jacobian", y.shape[0], _d(k), _d(y), _d(jac))
        return jac

    def lu(self, jac, alpha):
        """Returns (jacobian with alpha on its diagonal, L, U)."""
        jac = np.ascontiguousarray(jac, dtype=np.float64)
        n = jac.shape[0]
        jout = np.zeros_like(jac)
        lo = np.zeros((n, self.nnz_lower))
        up = np.zeros((n, self.nnz_upper))
        self._call("lu", n, _d(jac), float(alpha), _d(jout), _d(lo), _d(up))
        return jout, lo, up

    def lu_solve(self, lower, upper, b):
        x = np.array(b, dtype=np.float64, copy=True, order="C")
        self._call("lu_solve", x.shape[0], _d(np.ascontiguousarray(lower)), _d(np.ascontiguousarray(upper)), _d(x))
        return x

    def normalized_error_grouped(self, y, ynew, yerr, rtol, atol, group):
        """The norm over all cells of one State, summed in VectorMatrix<group> order (oracle only)."""
        y, ynew, yerr = (np.ascontiguousarray(a, dtype=np.float64) for a in (y, ynew, yerr))
        atol = np.ascontiguousarray(atol, dtype=np.float64)
        e = np.zeros(1)
        self._call("normalized_error_grouped", y.shape[0], _d(y), _d(ynew), _d(yerr), float(rtol), _d(atol),
                   int(group), _d(e))
        return float(e[0])

    def normalized_error(self, y, ynew, yerr, rtol, atol):
        """The reference's norm over all cells of one State."""
        y, ynew, yerr = (np.ascontiguousarray(a, dtype=np.float64) for a in (y, ynew, yerr))
        atol = np.ascontiguousarray(atol, dtype=np.float64)
        e = np.zeros(1)
        self._call("normalized_error", y.shape[0], _d(y), _d(ynew), _d(yerr), float(rtol), _d(atol), _d(e))
        return float(e[0])

    def solve(self, conditions, custom, y, dt, n_solves=1, lockstep=False, rtol=1e-6, atol=None):
        """n_solves x (UpdateStateParameters, Solve(dt)); per-cell States unless lockstep
        (True or 1: one State of every cell over Matrix<Real>; an int L > 1: over
        VectorMatrix<L>, whose error norm is summed group by group -- the reference build
        offers L = 4).  Returns (y_final, states, stats summed over solves, final_time of
        the last solve)."""
        conditions = np.ascontiguousarray(conditions, dtype=np.float64)
        n = conditions.shape[0]
        custom = np.ascontiguousarray(custom, dtype=np.float64).reshape(n, -1)
        y = np.array(y, dtype=np.float64, copy=True, order="C")
        states = np.zeros(n, dtype=np.int32)
        stats = np.zeros((n, 7), dtype=np.int64)
        ft = np.zeros(n)
        atol = None if atol is None else np.ascontiguousarray(atol, dtype=np.float64)
        self._call("solve", n, int(lockstep), _d(conditions), _d(custom), _d(y), float(dt), int(n_solves),
                   float(rtol), _d(atol), _i32(states), _i(stats), _d(ft))
        return y, states, stats, ft


class Tape:
    """Per-cell records of a recorded Solve sequence: records[c, r] = (H, y at the step's
    start) of an accepted step, or (-1, y before the clamp) at the end of a Solve."""

    CLAMP = -1.0

    def __init__(self, records, counts):
        self.records = records   # [cells, max_records, N + 1]
        self.counts = counts     # [cells]

    @property
    def max_records(self):
        return self.records.shape[1]


class _AdjointMixin:
    def set_control(self, max_steps=1000, h_min=0.0, h_max=0.0, h_start=0.0):
        """max_number_of_steps, h_min, h_max and h_start of this solver's parameters."""
        f = self.lib.oracle_set_control
        f.argtypes = [_vp, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        if f(self.h, int(max_steps), float(h_min), float(h_max), float(h_start)) != 0:
            raise ValueError(max_steps)

    def set_max_steps(self, n):
        """max_number_of_steps of this solver's Solves (for tight-tolerance references)."""
        self.lib.oracle_set_max_steps.argtypes = [_vp, ctypes.c_int64]
        if self.lib.oracle_set_max_steps(self.h, int(n)) != 0:
            raise ValueError(n)

    """The discrete adjoint of the recorded map (each accepted step with its H frozen),
    its tangent-linear model and the map itself, for finite differences."""

    def solve_record(self, k, y, dt, n_solves=1, max_records=4096, rtol=1e-6, atol=None):
        k = np.ascontiguousarray(k, dtype=np.float64)
        y = np.array(y, dtype=np.float64, copy=True, order="C")
        n = y.shape[0]
        states = np.zeros(n, dtype=np.int32)
        stats = np.zeros((n, 7), dtype=np.int64)
        tape = np.zeros((n, max_records, self.n_species + 1))
        counts = np.zeros(n, dtype=np.int64)
        atol = None if atol is None else np.ascontiguousarray(atol, dtype=np.float64)
        self._call("solve_record", n, _d(k), _d(y), float(dt), int(n_solves), float(rtol), _d(atol), _i32(states),
                   _i(stats), int(max_records), _d(tape), _i(counts))
        return y, states, stats, Tape(tape, counts)

    def replay(self, k, y0, tape):
        y = np.array(y0, dtype=np.float64, copy=True, order="C")
        self._call("replay", y.shape[0], _d(np.ascontiguousarray(k, dtype=np.float64)), _d(y), tape.max_records,
                   _d(tape.records), _i(tape.counts))
        return y

    def adjoint(self, k, tape, lam_final):
        lam_final = np.ascontiguousarray(lam_final, dtype=np.float64)
        n = lam_final.shape[0]
        gy = np.zeros_like(lam_final)
        gk = np.zeros((n, self.n_reactions))
        self._call("adjoint", n, _d(np.ascontiguousarray(k, dtype=np.float64)), tape.max_records, _d(tape.records),
                   _i(tape.counts), _d(lam_final), _d(gy), _d(gk))
        return gy, gk

    def tangent(self, k, y0, tape, dy0, dlogk):
        """dy0 [cells, n_dir, N], dlogk [cells, n_dir, R] -> dy_final [cells, n_dir, N]."""
        dy0 = np.ascontiguousarray(dy0, dtype=np.float64)
        dlogk = np.ascontiguousarray(dlogk, dtype=np.float64)
        n, nd = dy0.shape[:2]
        out = np.zeros_like(dy0)
        self._call("tangent", n, _d(np.ascontiguousarray(k, dtype=np.float64)),
                   _d(np.ascontiguousarray(y0, dtype=np.float64)), tape.max_records, _d(tape.records),
                   _i(tape.counts), nd, _d(dy0), _d(dlogk), _d(out))
        return out


def set_adjoint_fault(which):
    """Plants a mistake in the oracle's adjoint (0 none, 1 no second-derivative term,
    2 A instead of A^T, 3 a sign flip in the forcing's k-derivative) -- only for the tests
    that show the adjoint's checks catch such mistakes."""
    for cls in (Oracle, OraclePortableMath):
        lib = cls._lib()
        lib.oracle_set_adjoint_fault.argtypes = [ctypes.c_int32]
        if lib.oracle_set_adjoint_fault(int(which)) != 0:
            raise ValueError(which)


class Reference(_Base):
    """The reference's compiled CPU solver (needs oracle/_ref)."""
    prefix = "ref_"
    path = REF_LIB

    @classmethod
    def tableau(cls, preset):
        """Every field of the reference's RosenbrockSolverParameters preset, as arrays."""
        from micm_amd._lib import RosenbrockParameters
        lib = cls._lib()
        lib.ref_tableau.argtypes = [ctypes.c_int32, ctypes.POINTER(RosenbrockParameters)]
        p = RosenbrockParameters()
        if lib.ref_tableau(int(preset), ctypes.byref(p)) != 0:
            raise ValueError(preset)
        return tableau_fields(p)


def tableau_fields(p):
    """The fields of a micm_hip_rosenbrock_parameters struct as numpy arrays."""
    out = {}
    for name, _ in p._fields_:
        v = getattr(p, name)
        out[name] = np.array(list(v)) if hasattr(v, "_length_") else np.array([v])
    return out


class Oracle(_Base, _AdjointMixin):
    """The CPU restatement."""
    prefix = "oracle_"
    path = ORACLE_LIB


class OraclePortableMath(_Base, _AdjointMixin):
    """The CPU restatement with the HIP path's exp / log10 / pow (micm_amd/csrc/pmath.h)."""
    prefix = "oracle_"
    path = ORACLE_PM_LIB
