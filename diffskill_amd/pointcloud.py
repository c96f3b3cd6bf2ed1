"""Point-cloud losses on the GPU: nearest neighbours by the HIP kernel
``pc_nearest`` and the Chamfer losses built on it, differentiable in both
clouds through ``torch.autograd``.

- ``plb_chamfer_loss``: compute_chamfer_loss_kernel of the reference's loss
  module (plb/engine/losses/loss.py:183-204): the mean Euclidean distance from
  every particle to its nearest target plus the mean from every target to its
  nearest particle; the targets are the grid nodes whose target mass exceeds
  1e-4, at node positions I * dx (loss.py:71-75, ``grid_target_points``).
  The reference's subsampling stride ``min(n // 1000, 1)`` is 1 for every
  cloud of at least 1000 particles, the only case it runs.
- ``chamfer_distance``: nearest squared (or Euclidean) distances averaged over
  the source, one-way or both ways -- the quantity of the chamferdist loss the
  reference builds in create_chamfer_loss (plb/engine/taichi_env.py:28-32; that
  package is a third-party dependency absent here, its exact reductions are
  not reproduced).

The adjoint of a nearest distance d_i = |x_i - y_k(i)| (k(i) fixed, as the
reference's branch-wise Taichi adjoint and chamferdist's saved indices both
hold it) is (x_i - y_k) / d_i for x_i and its negative for y_k, scattered to
k(i); squared distances give 2 (x_i - y_k).
"""
import ctypes

import numpy as np
import torch

from . import _lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _cloud(p, name):
    if not (isinstance(p, torch.Tensor) and p.is_cuda):
        raise TypeError("%s must be a CUDA (HIP) tensor" % name)
    p = p.detach().to(torch.float32).contiguous()
    if p.dim() != 2 or p.shape[1] != 3:
        raise ValueError("%s must be (n, 3), got %s" % (name, tuple(p.shape)))
    return p


def nearest(q, t):
    """For each row of q (n, 3) the nearest row of t (m, 3): (indices (n,)
    int64, squared distances (n,) float32); ties go to the smaller index."""
    L = _lib.load()
    q = _cloud(q, "q")
    t = _cloud(t, "t")
    n, m = q.shape[0], t.shape[0]
    idx = torch.empty(n, dtype=torch.int32, device=q.device)
    d2 = torch.empty(n, dtype=torch.float32, device=q.device)
    work = torch.empty(max(1, n), dtype=torch.int64, device=q.device)
    rc = L.pc_nearest(ctypes.c_void_p(q.data_ptr()), n, ctypes.c_void_p(t.data_ptr()), m, ctypes.c_void_p(work.data_ptr()),
                      ctypes.c_void_p(idx.data_ptr()), ctypes.c_void_p(d2.data_ptr()), _stream())
    if rc != 0:
        raise RuntimeError("pc_nearest failed: %s" % L.pc_last_error().decode(errors="replace"))
    return idx.long(), d2


class _NearestDistance(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, squared):
        idx, d2 = nearest(x, y)
        d = d2 if squared else torch.sqrt(d2)
        ctx.save_for_backward(x, y, idx, d)
        ctx.squared = squared
        return d.to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        x, y, idx, d = ctx.saved_tensors
        diff = x - y.index_select(0, idx)
        if ctx.squared:
            gx = 2.0 * diff * g[:, None]
        else:
            inv = torch.where(d > 0, 1.0 / d.clamp_min(1e-30), torch.zeros_like(d))
            gx = diff * (g * inv.to(g.dtype))[:, None]
        gy = torch.zeros_like(y).index_add_(0, idx, -gx) if ctx.needs_input_grad[1] else None
        return gx if ctx.needs_input_grad[0] else None, gy, None


def nearest_distance(x, y, squared=True):
    """Distance from every row of x to its nearest row of y, (n,), differentiable
    in x and y."""
    return _NearestDistance.apply(x, y, squared)


def chamfer_distance(x, y, bidirectional=False, squared=True):
    """mean_i min_j d(x_i, y_j) (+ mean_j min_i d(x_i, y_j) when bidirectional),
    d the squared (default) or Euclidean distance."""
    loss = nearest_distance(x, y, squared).mean()
    if bidirectional:
        loss = loss + nearest_distance(y, x, squared).mean()
    return loss


def grid_target_points(grid, dx, threshold=1e-4):
    """Target points of the reference's Chamfer loss: the positions I * dx of
    the nodes with mass above the threshold, in numpy's C order (loss.py:71-72)."""
    g = grid.detach().cpu().numpy() if isinstance(grid, torch.Tensor) else np.asarray(grid)
    return (np.array(np.where(g > threshold)).transpose() * dx).astype(np.float64)


def plb_chamfer_loss(x, target_points):
    """compute_chamfer_loss_kernel (loss.py:183-204): Euclidean Chamfer both
    ways, each direction averaged over its own points."""
    t = torch.as_tensor(target_points, dtype=x.dtype, device=x.device)
    return chamfer_distance(x, t, bidirectional=True, squared=False)
