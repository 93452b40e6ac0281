"""torch.autograd binding of the discrete adjoint (micm_hip_solve_recorded, micm_hip_adjoint).

    solve = DifferentiableSolve(desc, conditions, custom, tableau=4, time_step=300.0, n_solves=10)
    y0 = torch.tensor(y_init, device="cuda", requires_grad=True)         # [cells, species], declared order
    log_a = torch.zeros(solve.n_reactions, device="cuda", dtype=torch.float64, requires_grad=True)
    y = solve(y0, log_a)                                                 # [cells, species]
    (y[:, targets] ** 2).sum().backward()                                # y0.grad, log_a.grad

Forward: the rate constants of UpdateStateParameters (computed once, at construction)
scaled by exp(log_a) -- a shared [reactions] vector or one row per cell -- then
`n_solves` Solves of `time_step`, each recorded, each followed by the clamp of
solver.hpp:275-297.  Backward: one reverse sweep over the recorded checkpoints on the GPU.
The derivative is that of the recorded map (accepted steps with their step sizes frozen);
log_a of a reaction is the log of a common factor on all of its A-prefactors, so for
every rate type d/dlog_a = k d/dk.

Species and reactions are in declaration order (the order of the mechanism description);
the solver's internal orders (reordered species, type-sorted reactions) stay inside.
All tensors are FP64 on the GPU.  The object owns one device state and one tape: a
forward / backward pair must finish before the next forward.
"""
import torch

from . import _lib

GROUP = _lib.GROUP


def _to_grouped(x):
    """[cells, values] -> the flat grouped layout of include/micm_hip.h (a new tensor)."""
    c, v = x.shape
    g = (c + GROUP - 1) // GROUP
    pad = torch.zeros((g * GROUP, v), dtype=torch.float64, device=x.device)
    pad[:c] = x
    return pad.view(g, GROUP, v).transpose(1, 2).contiguous().view(-1)


def _from_grouped(flat, c, v):
    g = (c + GROUP - 1) // GROUP
    return flat.view(g, v, GROUP).transpose(1, 2).reshape(g * GROUP, v)[:c]


class DifferentiableSolve:
    """A mechanism, a set of cells (conditions, custom parameters) and a Solve sequence
    whose result is differentiable with respect to the initial concentrations and the
    per-reaction log rate parameters."""

    def __init__(self, desc, conditions, custom, tableau=4, time_step=300.0, n_solves=1, max_records=4096,
                 rtol=1e-6, atol=None):
        import numpy as np
        self.mechanism = _lib.Mechanism(desc)
        self.n_cells = int(np.asarray(conditions).shape[0])
        self.n_species = self.mechanism.n_species
        self.n_reactions = self.mechanism.n_reactions
        self.params = _lib.rosenbrock_parameters(tableau)
        self.time_step = float(time_step)
        self.n_solves = int(n_solves)
        self.state = _lib.DeviceState(self.mechanism, self.n_cells)
        self.state.set_conditions(conditions)
        self.state.set_custom_parameters(custom)
        if atol is not None:
            atol = np.broadcast_to(np.asarray(atol, dtype=np.float64), (self.n_species,))
            # the description's order -> the state's
            atol = atol[self.mechanism.species_order()]
        self.state.set_tolerances(rtol, atol)
        self.state.update_state_parameters()
        self.tape = _lib.Tape(self.mechanism, self.n_cells, max_records)
        dev = torch.device("cuda", torch.cuda.current_device())
        self.species = torch.as_tensor(self.mechanism.species_order(), device=dev)     # state -> declared
        self.reactions = torch.as_tensor(self.mechanism.reaction_order(), device=dev)  # sorted -> declared
        k0 = torch.zeros(((self.n_cells + GROUP - 1) // GROUP) * GROUP * max(self.n_reactions, 1),
                         dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        self.state.get_rate_constants_device(k0.data_ptr())
        self.k0 = _from_grouped(k0, self.n_cells, self.n_reactions).clone()  # [cells, sorted reactions]
        self.last_status = None

    def __call__(self, y0, log_a):
        return _RecordedSolve.apply(y0, log_a, self)

    # the two halves, on tensors in the solver's internal orders
    def _forward(self, y0_state, theta_sorted):
        torch.cuda.synchronize()
        self.state.set_stream(torch.cuda.current_stream().cuda_stream)
        k = self.k0 * torch.exp(theta_sorted)
        kg, yg = _to_grouped(k), _to_grouped(y0_state)
        torch.cuda.synchronize()
        self.state.set_rate_constants_device(kg.data_ptr())
        self.state.set_variables_device(yg.data_ptr())
        self.tape.reset()
        for _ in range(self.n_solves):
            self.state.solve_recorded(self.time_step, self.params, self.tape)
        self.last_status = self.state.results()[0]
        if self.tape.overflowed():
            raise _lib.MicmHipError(-1, "tape overflow: raise max_records")
        out = torch.empty_like(yg)
        self.state.get_variables_device(out.data_ptr())
        return _from_grouped(out, self.n_cells, self.n_species)

    def _backward(self, grad_state):
        seed = _to_grouped(grad_state)
        gy = torch.zeros_like(seed)
        gk = torch.zeros(((self.n_cells + GROUP - 1) // GROUP) * GROUP * max(self.n_reactions, 1),
                         dtype=torch.float64, device=seed.device)
        torch.cuda.synchronize()
        self.state.adjoint(self.params, self.tape, seed.data_ptr(), gy.data_ptr(), gk.data_ptr())
        return _from_grouped(gy, self.n_cells, self.n_species), _from_grouped(gk, self.n_cells, self.n_reactions)


class _RecordedSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y0, log_a, solver):
        if y0.dtype != torch.float64 or log_a.dtype != torch.float64:
            raise TypeError("the solver computes in FP64: pass float64 tensors")
        if y0.shape != (solver.n_cells, solver.n_species):
            raise ValueError(f"y0 must be [{solver.n_cells}, {solver.n_species}]")
        shared = log_a.dim() == 1
        if log_a.shape[-1] != solver.n_reactions or (not shared and log_a.shape[0] != solver.n_cells):
            raise ValueError("log_a must be [n_reactions] or [n_cells, n_reactions]")
        theta = log_a[solver.reactions] if shared else log_a[:, solver.reactions]
        y_state = solver._forward(y0[:, solver.species].contiguous(), theta)
        ctx.solver = solver
        ctx.shared = shared
        y = torch.empty_like(y0)
        y[:, solver.species] = y_state
        return y

    @staticmethod
    def backward(ctx, grad_y):
        solver = ctx.solver
        gy_state, gk_sorted = solver._backward(grad_y[:, solver.species].contiguous())
        grad_y0 = grad_log_a = None
        if ctx.needs_input_grad[0]:
            grad_y0 = torch.empty_like(grad_y)
            grad_y0[:, solver.species] = gy_state
        if ctx.needs_input_grad[1]:
            if ctx.shared:
                grad_log_a = torch.empty(solver.n_reactions, dtype=torch.float64, device=grad_y.device)
                grad_log_a[solver.reactions] = gk_sorted.sum(dim=0)
            else:
                grad_log_a = torch.empty((solver.n_cells, solver.n_reactions), dtype=torch.float64,
                                         device=grad_y.device)
                grad_log_a[:, solver.reactions] = gk_sorted
        return grad_y0, grad_log_a, None
