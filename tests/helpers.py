"""Shared helpers: device buffers in the grouped layout and comparison metrics."""
import numpy as np

from micm_amd._lib import from_grouped, to_grouped


def dev(host):
    import torch
    return torch.from_numpy(to_grouped(host)).to("cuda")


def dev_zeros(n_cells, n_values):
    import torch
    g = (n_cells + 63) // 64
    return torch.zeros(g * 64 * max(n_values, 1), dtype=torch.float64, device="cuda")


def host(t, n_cells, n_values):
    import torch
    torch.cuda.synchronize()
    return from_grouped(t.cpu().numpy()[: ((n_cells + 63) // 64) * 64 * n_values], n_cells, n_values)


def stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def max_rel(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    mask = np.abs(b) > floor
    if not mask.any():
        return 0.0
    return float(np.max(np.abs(a[mask] - b[mask]) / np.abs(b[mask])))
