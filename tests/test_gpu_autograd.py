"""The torch.autograd binding of the discrete adjoint (micm_amd/autograd.py) on the GPU:
its forward is the library's Solve (bit for bit against the oracle), its gradients are
the adjoint's, and directional central differences of its own forward agree."""
import numpy as np
import pytest

from micm_amd import synthetic
from oracle import OraclePortableMath

pytestmark = pytest.mark.gpu

TARGETS = [3, 10, 20, 33, 50]


def _setup(n_cells=128, n_solves=2):
    import torch
    from micm_amd.autograd import DifferentiableSolve
    desc = synthetic.mechanism_60_200()
    cells = synthetic.inputs_60_200(desc, n_cells, seed=31)
    solve = DifferentiableSolve(desc, cells.conditions, cells.custom, tableau=4, time_step=300.0,
                                n_solves=n_solves, max_records=2048, atol=1e-12)
    y0 = torch.tensor(cells.variables, dtype=torch.float64, device="cuda")
    return desc, cells, solve, y0


def _loss(y):
    return (y[:, TARGETS] ** 2).sum()


def test_forward_is_the_solve():
    import torch
    desc, cells, solve, y0 = _setup()
    log_a = torch.zeros(solve.n_reactions, dtype=torch.float64, device="cuda")
    y = solve(y0, log_a).cpu().numpy()
    o = OraclePortableMath(desc, 4)
    order = o.species_order()
    k = o.rate_constants(cells.conditions, cells.custom)
    atol = np.full(o.n_species, 1e-12)
    y_ref, _, _, _ = o.solve_record(k, cells.variables[:, order], 300.0, n_solves=2, max_records=2048, atol=atol)
    y_ref_declared = np.empty_like(y_ref)
    y_ref_declared[:, order] = y_ref
    assert np.array_equal(y.view(np.int64), y_ref_declared.view(np.int64))


def test_gradients_match_directional_differences():
    """Directional central differences (h = 1e-6) of the binding's own forward.  Each
    forward records anew, so these differences include how the adaptive step sizes move
    with the input, which the frozen-step adjoint leaves out (measured ~1e-6 relative here);
    hence 1e-5.  The 1e-6 check against differences of the recorded map itself is
    tests/test_gpu_adjoint.py::test_differentiation_workload."""
    import torch
    desc, cells, solve, y0 = _setup()
    y0 = y0.clone().requires_grad_(True)
    log_a = torch.zeros(solve.n_reactions, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = _loss(solve(y0, log_a))
    loss.backward()
    gy, ga = y0.grad.clone(), log_a.grad.clone()
    assert torch.isfinite(gy).all() and torch.isfinite(ga).all()
    g = torch.Generator(device="cpu").manual_seed(7)
    h = 1e-6
    with torch.no_grad():
        for _ in range(3):
            vy = (torch.randn(y0.shape, generator=g, dtype=torch.float64) .to("cuda")) * y0
            fd = (_loss(solve(y0 + h * vy, log_a)) - _loss(solve(y0 - h * vy, log_a))) / (2 * h)
            ad = (gy * vy).sum()
            print(f"  y0 direction: fd {fd.item():.10e} adjoint {ad.item():.10e}", flush=True)
            assert abs(fd.item() - ad.item()) <= 1e-5 * abs(fd.item())
            va = torch.randn(solve.n_reactions, generator=g, dtype=torch.float64).to("cuda")
            fd = (_loss(solve(y0, log_a + h * va)) - _loss(solve(y0, log_a - h * va))) / (2 * h)
            ad = (ga * va).sum()
            print(f"  log A direction: fd {fd.item():.10e} adjoint {ad.item():.10e}", flush=True)
            assert abs(fd.item() - ad.item()) <= 1e-5 * abs(fd.item())


def test_per_cell_parameters_sum_to_the_shared_gradient():
    import torch
    desc, cells, solve, y0 = _setup(n_cells=70, n_solves=1)
    shared = torch.zeros(solve.n_reactions, dtype=torch.float64, device="cuda", requires_grad=True)
    _loss(solve(y0, shared)).backward()
    per_cell = torch.zeros((solve.n_cells, solve.n_reactions), dtype=torch.float64, device="cuda",
                           requires_grad=True)
    _loss(solve(y0, per_cell)).backward()
    assert torch.equal(per_cell.grad.sum(dim=0), shared.grad) or \
        torch.allclose(per_cell.grad.sum(dim=0), shared.grad, rtol=1e-13, atol=0.0)
