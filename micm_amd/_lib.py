"""ctypes binding of the C-ABI in include/micm_hip.h (libmicm_hip.so, built in-tree).

There is no fallback: if the HIP library is not built, loading fails loudly.
"""
import ctypes
import os

import numpy as np

from ._desc import MechanismDescC

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmicm_hip.so")

GROUP = 64

OK = 0
STATE_NAMES = {0: "NotYetCalled", 1: "Running", 2: "Converged", 3: "ConvergenceExceededMaxSteps",
               4: "StepSizeTooSmall", 5: "RepeatedlySingularMatrix", 6: "NaNDetected", 7: "InfDetected"}
SOLVE_LOCKSTEP, SOLVE_PLAIN, SOLVE_PERSISTENT, SOLVE_COOPERATIVE = 0, 1, 2, 3
SOLVE_UPDATE_RATES = 0x100  # OR-ed into a form: UpdateStateParameters fused into the Solve launch
PRESETS = {2: 2, 3: 3, 4: 4, "4da": 41, "6da": 61, 41: 41, 61: 61}

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32


class MechanismInfo(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "n_species", "n_reactions", "n_custom_parameters", "nnz_jacobian", "nnz_lower", "nnz_upper",
        "n_process_infos", "n_jacobian_terms", "lu_multiply_adds", "n_reactant_slots", "n_product_slots")]


class RosenbrockParameters(ctypes.Structure):
    _fields_ = [
        ("stages", ctypes.c_int64), ("max_number_of_steps", ctypes.c_int64),
        ("round_off", ctypes.c_double), ("factor_min", ctypes.c_double), ("factor_max", ctypes.c_double),
        ("rejection_factor_decrease", ctypes.c_double), ("safety_factor", ctypes.c_double),
        ("h_min", ctypes.c_double), ("h_max", ctypes.c_double), ("h_start", ctypes.c_double),
        ("new_function_evaluation", ctypes.c_int32 * 6), ("estimator_of_local_order", ctypes.c_double),
        ("a", ctypes.c_double * 15), ("c", ctypes.c_double * 15), ("m", ctypes.c_double * 6),
        ("e", ctypes.c_double * 6), ("alpha", ctypes.c_double * 6), ("gamma", ctypes.c_double * 6),
    ]


class StatePointers(ctypes.Structure):
    _fields_ = [
        ("variables", _vp), ("conditions", _vp), ("custom_parameters", _vp), ("rate_constants", _vp),
        ("jacobian", _vp), ("lower", _vp), ("upper", _vp), ("absolute_tolerances", _vp),
        ("cell_state", _vp), ("cell_stats", _vp), ("cell_final_time", _vp),
        ("n_cells", ctypes.c_int64), ("n_groups", ctypes.c_int64),
    ]


# name -> argtypes (all return int unless listed in _VOID)
_SIGNATURES = {
    "micm_hip_rosenbrock_parameters_preset": [ctypes.c_int32, ctypes.POINTER(RosenbrockParameters)],
    "micm_hip_mechanism_create": [ctypes.POINTER(MechanismDescC), ctypes.POINTER(_vp)],
    "micm_hip_mechanism_destroy": [_vp],
    "micm_hip_mechanism_get_info": [_vp, ctypes.POINTER(MechanismInfo)],
    "micm_hip_mechanism_get_species_order": [_vp, _ip],
    "micm_hip_mechanism_get_reaction_order": [_vp, _ip],
    "micm_hip_mechanism_get_custom_parameter_owner": [_vp, _ip],
    "micm_hip_mechanism_get_patterns": [_vp, _ip, _ip, _ip, _ip, _ip, _ip],
    "micm_hip_mechanism_get_jacobian_diagonal": [_vp, _ip],
    "micm_hip_calculate_rate_constants": [_vp, _i64, _vp, _vp, _vp, _vp],
    "micm_hip_add_forcing_terms": [_vp, _i64, _vp, _vp, _vp, _vp],
    "micm_hip_fill_and_add_forcing_terms": [_vp, _i64, _vp, _vp, _vp, _vp],
    "micm_hip_subtract_jacobian_terms": [_vp, _i64, _vp, _vp, _vp, _vp],
    "micm_hip_fill_and_subtract_jacobian_terms": [_vp, _i64, _vp, _vp, _vp, _vp],
    "micm_hip_alpha_minus_jacobian": [_vp, _i64, _vp, ctypes.c_double, _vp],
    "micm_hip_lu_decompose": [_vp, _i64, _vp, _vp, _vp, _vp],
    "micm_hip_lu_solve": [_vp, _i64, _vp, _vp, _vp, _vp],
    "micm_hip_normalized_error": [_vp, _i64, _vp, _vp, _vp, _vp, ctypes.c_double, _vp, _vp],
    "micm_hip_state_create": [_vp, _i64, ctypes.POINTER(_vp)],
    "micm_hip_state_destroy": [_vp],
    "micm_hip_state_get_pointers": [_vp, ctypes.POINTER(StatePointers)],
    "micm_hip_state_set_variables": [_vp, _dp],
    "micm_hip_state_get_variables": [_vp, _dp],
    "micm_hip_state_set_conditions": [_vp, _dp],
    "micm_hip_state_set_custom_parameters": [_vp, _dp],
    "micm_hip_state_get_rate_constants": [_vp, _dp],
    "micm_hip_state_set_rate_constants": [_vp, _dp],
    "micm_hip_state_set_tolerances": [_vp, ctypes.c_double, _dp],
    "micm_hip_state_get_results": [_vp, _i32p, _ip, _dp],
    "micm_hip_state_set_stream": [_vp, _vp],
    "micm_hip_state_set_handoff": [_vp, _i64],
    "micm_hip_state_set_cooperative_shape": [_vp, _i32, _i32],
    "micm_hip_state_set_lu_blocking": [_vp, ctypes.c_int32],
    "micm_hip_packed_create": [_i32, ctypes.POINTER(_vp), _i32, ctypes.POINTER(_vp)],
    "micm_hip_packed_solve": [_vp, ctypes.c_double, ctypes.POINTER(RosenbrockParameters)],
    "micm_hip_packed_replayed": [_vp],
    "micm_hip_packed_set_stream": [_vp, _vp],
    "micm_hip_packed_destroy": [_vp],
    "micm_hip_mechanism_cooperative_supported": [_vp, ctypes.c_int32, _i32p],
    "micm_hip_update_state_parameters": [_vp, _vp],
    "micm_hip_solve": [_vp, _vp, ctypes.c_double, ctypes.POINTER(RosenbrockParameters), ctypes.c_int32],
    "micm_hip_normalized_error_state": [_vp, _i64, _vp, _vp, _vp, _vp, ctypes.c_double, _i64, _vp, _vp],
    "micm_hip_state_set_lockstep_order": [_vp, _i64],
    "micm_hip_solve_noalloc": [_vp, _vp, ctypes.c_double, ctypes.POINTER(RosenbrockParameters), ctypes.c_int32],
    "micm_hip_state_set_variables_device": [_vp, _vp],
    "micm_hip_state_get_variables_device": [_vp, _vp],
    "micm_hip_state_set_rate_constants_device": [_vp, _vp],
    "micm_hip_state_get_rate_constants_device": [_vp, _vp],
    "micm_hip_tape_create": [_vp, _i64, _i64, ctypes.POINTER(_vp)],
    "micm_hip_tape_destroy": [_vp],
    "micm_hip_tape_reset": [_vp],
    "micm_hip_tape_get": [_vp, _i64, _i64, _dp, _ip],
    "micm_hip_tape_set": [_vp, _dp, _ip],
    "micm_hip_tape_overflowed": [_vp, _ip],
    "micm_hip_solve_recorded": [_vp, _vp, ctypes.c_double, ctypes.POINTER(RosenbrockParameters), ctypes.c_int32, _vp],
    "micm_hip_adjoint": [_vp, _vp, ctypes.POINTER(RosenbrockParameters), _vp, _vp, _vp, _vp],
}
_VOID = {"micm_hip_mechanism_destroy", "micm_hip_state_destroy", "micm_hip_packed_destroy", "micm_hip_tape_destroy"}

_lib = None


def lib():
    """The loaded HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
        # PyTorch bundles its own HIP runtime under the same soname as the system one this
        # library links; whichever loads first serves the whole process, and PyTorch only
        # works with its own.  Load it first so that both share it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        L.micm_hip_last_error.restype = ctypes.c_char_p
        L.micm_hip_last_error.argtypes = []
        for name, args in _SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None if name in _VOID else ctypes.c_int
        _lib = L
        _init_device_runtime()
    return _lib


def _init_device_runtime():
    """The library shares the process's HIP runtime with PyTorch when PyTorch is
    loaded; initialise it through PyTorch first so both see the same device state."""
    import sys
    torch = sys.modules.get("torch")
    if torch is None:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except Exception:
        pass


class MicmHipError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


def check(rc):
    if rc != OK:
        raise MicmHipError(rc, lib().micm_hip_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def rosenbrock_parameters(preset):
    p = RosenbrockParameters()
    check(lib().micm_hip_rosenbrock_parameters_preset(PRESETS[preset], ctypes.byref(p)))
    return p


def to_grouped(host, n_values=None):
    """[n_cells, n_values] cell-major -> flat grouped layout (numpy)."""
    host = np.asarray(host, dtype=np.float64)
    n, v = host.shape if n_values is None else (host.shape[0], n_values)
    g = (n + GROUP - 1) // GROUP
    buf = np.zeros((g * GROUP, v))
    buf[:n] = host.reshape(n, v)
    return np.ascontiguousarray(buf.reshape(g, GROUP, v).transpose(0, 2, 1)).reshape(-1)


def from_grouped(flat, n_cells, n_values):
    g = (n_cells + GROUP - 1) // GROUP
    a = np.asarray(flat, dtype=np.float64).reshape(g, n_values, GROUP).transpose(0, 2, 1).reshape(g * GROUP, n_values)
    return np.ascontiguousarray(a[:n_cells])


class Mechanism:
    """A compiled mechanism resident on the current device (micm_hip_mechanism)."""

    def __init__(self, desc):
        self.desc = desc
        h = _vp()
        check(lib().micm_hip_mechanism_create(ctypes.byref(desc.as_c()), ctypes.byref(h)))
        self.h = h
        info = MechanismInfo()
        check(lib().micm_hip_mechanism_get_info(self.h, ctypes.byref(info)))
        self.info = info
        self.n_species = info.n_species
        self.n_reactions = info.n_reactions
        self.n_custom = info.n_custom_parameters
        self.nnz_jacobian = info.nnz_jacobian
        self.nnz_lower = info.nnz_lower
        self.nnz_upper = info.nnz_upper

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.micm_hip_mechanism_destroy(self.h)
            self.h = None

    def species_order(self):
        out = np.zeros(self.n_species, dtype=np.int64)
        check(lib().micm_hip_mechanism_get_species_order(self.h, _i(out)))
        return out

    def reaction_order(self):
        out = np.zeros(max(self.n_reactions, 1), dtype=np.int64)
        check(lib().micm_hip_mechanism_get_reaction_order(self.h, _i(out)))
        return out[:self.n_reactions]

    def jacobian_diagonal(self):
        out = np.zeros(self.n_species, dtype=np.int64)
        check(lib().micm_hip_mechanism_get_jacobian_diagonal(self.h, _i(out)))
        return out

    def patterns(self):
        a = [np.zeros(max(n, 1), dtype=np.int64) for n in (self.nnz_jacobian, self.nnz_jacobian, self.nnz_lower,
                                                           self.nnz_lower, self.nnz_upper, self.nnz_upper)]
        check(lib().micm_hip_mechanism_get_patterns(self.h, *[_i(x) for x in a]))
        return {"jacobian": (a[0], a[1]), "lower": (a[2], a[3]), "upper": (a[4], a[5])}

    # ---- kernel entry points on device pointers (grouped layout); stream = hipStream_t or None
    def calculate_rate_constants(self, n_cells, conditions, custom, k, stream=None):
        check(lib().micm_hip_calculate_rate_constants(self.h, n_cells, conditions, custom, k, stream))

    def add_forcing_terms(self, n_cells, k, y, f, stream=None):
        check(lib().micm_hip_add_forcing_terms(self.h, n_cells, k, y, f, stream))

    def subtract_jacobian_terms(self, n_cells, k, y, jac, stream=None):
        check(lib().micm_hip_subtract_jacobian_terms(self.h, n_cells, k, y, jac, stream))

    def fill_and_add_forcing_terms(self, n_cells, k, y, f, stream=None):
        """f = f(y): Fill(0) and AddForcingTerms in one pass."""
        check(lib().micm_hip_fill_and_add_forcing_terms(self.h, n_cells, k, y, f, stream))

    def fill_and_subtract_jacobian_terms(self, n_cells, k, y, jac, stream=None):
        """jac = -J(y): Fill(0) and SubtractJacobianTerms in one pass."""
        check(lib().micm_hip_fill_and_subtract_jacobian_terms(self.h, n_cells, k, y, jac, stream))

    def alpha_minus_jacobian(self, n_cells, jac, alpha, stream=None):
        check(lib().micm_hip_alpha_minus_jacobian(self.h, n_cells, jac, float(alpha), stream))

    def lu_decompose(self, n_cells, jac, lower, upper, stream=None):
        check(lib().micm_hip_lu_decompose(self.h, n_cells, jac, lower, upper, stream))

    def lu_solve(self, n_cells, lower, upper, x, stream=None):
        check(lib().micm_hip_lu_solve(self.h, n_cells, lower, upper, x, stream))

    def cooperative_supported(self, stages):
        out = ctypes.c_int32(0)
        check(lib().micm_hip_mechanism_cooperative_supported(self.h, int(stages), ctypes.byref(out)))
        return bool(out.value)

    def normalized_error(self, n_cells, y, ynew, yerr, atol, rtol, error, stream=None):
        check(lib().micm_hip_normalized_error(self.h, n_cells, y, ynew, yerr, atol, float(rtol), error, stream))

    def normalized_error_state(self, n_cells, y, ynew, yerr, atol, rtol, vector_length, error, stream=None):
        """One norm for the whole State, summed in the order of VectorMatrix<vector_length>
        (1: Matrix<Real>); `error` is the address of one device double."""
        check(lib().micm_hip_normalized_error_state(self.h, n_cells, y, ynew, yerr, atol, float(rtol),
                                                    int(vector_length), error, stream))


class DeviceState:
    """Device-resident per-cell arrays and integrator scratch (micm_hip_state)."""

    def __init__(self, mechanism, n_cells):
        self.mechanism = mechanism
        self.n_cells = int(n_cells)
        h = _vp()
        check(lib().micm_hip_state_create(mechanism.h, self.n_cells, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.micm_hip_state_destroy(self.h)
            self.h = None

    def pointers(self):
        p = StatePointers()
        check(lib().micm_hip_state_get_pointers(self.h, ctypes.byref(p)))
        return p

    def set_variables(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(self.n_cells, self.mechanism.n_species)
        check(lib().micm_hip_state_set_variables(self.h, _d(y)))

    def get_variables(self):
        y = np.zeros((self.n_cells, self.mechanism.n_species))
        check(lib().micm_hip_state_get_variables(self.h, _d(y)))
        return y

    def set_conditions(self, cond):
        cond = np.ascontiguousarray(cond, dtype=np.float64).reshape(self.n_cells, 3)
        check(lib().micm_hip_state_set_conditions(self.h, _d(cond)))

    def set_custom_parameters(self, custom):
        if self.mechanism.n_custom == 0:
            return
        custom = np.ascontiguousarray(custom, dtype=np.float64).reshape(self.n_cells, self.mechanism.n_custom)
        check(lib().micm_hip_state_set_custom_parameters(self.h, _d(custom)))

    def get_rate_constants(self):
        k = np.zeros((self.n_cells, self.mechanism.n_reactions))
        check(lib().micm_hip_state_get_rate_constants(self.h, _d(k)))
        return k

    def set_rate_constants(self, k):
        k = np.ascontiguousarray(k, dtype=np.float64).reshape(self.n_cells, self.mechanism.n_reactions)
        check(lib().micm_hip_state_set_rate_constants(self.h, _d(k)))

    def set_tolerances(self, rtol, atol=None):
        a = None if atol is None else np.ascontiguousarray(atol, dtype=np.float64)
        check(lib().micm_hip_state_set_tolerances(self.h, float(rtol), _d(a) if a is not None else _dp()))

    def set_stream(self, stream):
        check(lib().micm_hip_state_set_stream(self.h, stream))

    def set_lockstep_order(self, vector_length):
        """Summation order of the lockstep form's error norm: that of VectorMatrix<vector_length>
        (1, the default: Matrix<Real>)."""
        check(lib().micm_hip_state_set_lockstep_order(self.h, int(vector_length)))

    def set_handoff(self, steps):
        """Step count at which the persistent form hands a cell to the cooperative form
        (0: never, negative: library default)."""
        check(lib().micm_hip_state_set_handoff(self.h, int(steps)))

    def set_cooperative_shape(self, threads=0, stage=-1):
        """Threads per cell (64 .. 1024, 0 = library choice) and LDS staging / layout
        (micm_hip_state_set_cooperative_shape; -1 = library choice)."""
        check(lib().micm_hip_state_set_cooperative_shape(self.h, int(threads), int(stage)))

    def set_lu_blocking(self, min_segment):
        """Shortest segment of the cooperative form's LU schedule: 1, 2, 4, 8, -1 (whole
        sums) or 0 (the library's default)."""
        check(lib().micm_hip_state_set_lu_blocking(self.h, int(min_segment)))

    def results(self):
        states = np.zeros(self.n_cells, dtype=np.int32)
        stats = np.zeros((self.n_cells, 7), dtype=np.int64)
        ft = np.zeros(self.n_cells)
        check(lib().micm_hip_state_get_results(self.h, states.ctypes.data_as(_i32p), _i(stats), _d(ft)))
        return states, stats, ft

    def update_state_parameters(self):
        check(lib().micm_hip_update_state_parameters(self.mechanism.h, self.h))

    def solve(self, time_step, parameters, form=SOLVE_PLAIN):
        check(lib().micm_hip_solve(self.mechanism.h, self.h, float(time_step), ctypes.byref(parameters), int(form)))

    def solve_noalloc(self, time_step, parameters, update_rates=False):
        """The cooperative Solve without device allocation, enqueued on the state's stream
        (results valid once that stream has finished)."""
        check(lib().micm_hip_solve_noalloc(self.mechanism.h, self.h, float(time_step), ctypes.byref(parameters),
                                           SOLVE_UPDATE_RATES if update_rates else 0))

    # ---- device-resident transfers (grouped device arrays given by address)
    def set_variables_device(self, ptr):
        check(lib().micm_hip_state_set_variables_device(self.h, ptr))

    def get_variables_device(self, ptr):
        check(lib().micm_hip_state_get_variables_device(self.h, ptr))

    def set_rate_constants_device(self, ptr):
        check(lib().micm_hip_state_set_rate_constants_device(self.h, ptr))

    def get_rate_constants_device(self, ptr):
        check(lib().micm_hip_state_get_rate_constants_device(self.h, ptr))

    # ---- discrete adjoint
    def solve_recorded(self, time_step, parameters, tape, update_rates=False):
        """A cooperative-form Solve that appends each cell's checkpoints to `tape`."""
        check(lib().micm_hip_solve_recorded(self.mechanism.h, self.h, float(time_step), ctypes.byref(parameters),
                                            SOLVE_UPDATE_RATES if update_rates else 0, tape.h))

    def adjoint(self, parameters, tape, seed, grad_variables, grad_log_rate_constants):
        """Reverse sweep over `tape`; device addresses of grouped arrays: seed = dL/dy_final,
        outputs dL/dy_initial and k dL/dk per cell."""
        check(lib().micm_hip_adjoint(self.mechanism.h, self.h, ctypes.byref(parameters), tape.h, seed,
                                     grad_variables, grad_log_rate_constants))


class Tape:
    """Per-accepted-step checkpoints of recorded Solves (micm_hip_tape): per cell up to
    max_records records (H, y at the step's start) or (-1, y before the clamp)."""

    CLAMP = -1.0

    def __init__(self, mechanism, n_cells, max_records):
        self.mechanism = mechanism
        self.n_cells = int(n_cells)
        self.max_records = int(max_records)
        h = _vp()
        check(lib().micm_hip_tape_create(mechanism.h, self.n_cells, self.max_records, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.micm_hip_tape_destroy(self.h)
            self.h = None

    def reset(self):
        check(lib().micm_hip_tape_reset(self.h))

    def counts(self):
        c = np.zeros(self.n_cells, dtype=np.int64)
        check(lib().micm_hip_tape_get(self.h, 0, self.n_cells, _dp(), _i(c)))
        return c

    def records(self, cells=None):
        """Records of the given cells (all by default), [cells, max_records, n_species + 1]."""
        idx = np.arange(self.n_cells) if cells is None else np.asarray(cells, dtype=np.int64)
        r = np.zeros((len(idx), self.max_records, self.mechanism.n_species + 1))
        for j, c in enumerate(idx):
            check(lib().micm_hip_tape_get(self.h, int(c), 1, _d(r[j]), _ip()))
        return r

    def set(self, records, counts):
        r = np.ascontiguousarray(records, dtype=np.float64)
        c = np.ascontiguousarray(counts, dtype=np.int64)
        if r.shape != (self.n_cells, self.max_records, self.mechanism.n_species + 1) or c.shape != (self.n_cells,):
            raise ValueError("records / counts of the wrong shape")
        check(lib().micm_hip_tape_set(self.h, _d(r), _i(c)))

    def overflowed(self):
        n = ctypes.c_int64(0)
        check(lib().micm_hip_tape_overflowed(self.h, ctypes.byref(n)))
        return n.value


class PackedSolve:
    """Several DeviceStates (any mechanisms) solved together by one graph-captured
    launch, one state after another (micm_hip_packed_*).  Keeps the states
    alive while it exists."""

    def __init__(self, states, form=SOLVE_COOPERATIVE):
        self.states = list(states)
        arr = (_vp * len(self.states))(*[s.h for s in self.states])
        h = _vp()
        check(lib().micm_hip_packed_create(len(self.states), arr, int(form), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().micm_hip_packed_destroy(self.h)
            self.h = None

    def solve(self, time_step, parameters):
        check(lib().micm_hip_packed_solve(self.h, float(time_step), ctypes.byref(parameters)))

    def set_stream(self, stream):
        """Capture and launch on `stream` (hipStream_t, not the legacy default stream)."""
        check(lib().micm_hip_packed_set_stream(self.h, stream))

    @property
    def replayed(self):
        return bool(lib().micm_hip_packed_replayed(self.h))

