"""Point-cloud kernels: pc_nearest against scipy's k-d tree (fp64) on ragged
sizes, duplicate points and an empty target set; the Chamfer losses and their
gradients against float64 torch on the CPU; the reference's grid-target
Chamfer loss (loss.py:183-204) against a k-d tree."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _cuda(a):
    return torch.as_tensor(np.asarray(a, np.float32), device="cuda")


@pytest.mark.parametrize("nq,nt", [(1000, 3000), (5000, 2501), (257, 2049), (3, 7), (30000, 30000)])
def test_nearest_matches_kdtree(nq, nt):
    from scipy.spatial import cKDTree
    from diffskill_amd.pointcloud import nearest
    print("case nearest %d x %d" % (nq, nt), flush=True)
    rng = np.random.default_rng(nq + nt)
    q = rng.uniform(0.2, 0.8, (nq, 3)).astype(np.float32).astype(np.float64)
    t = rng.uniform(0.2, 0.8, (nt, 3)).astype(np.float32).astype(np.float64)
    idx, d2 = nearest(_cuda(q), _cuda(t))
    idx = idx.cpu().numpy()
    d2 = d2.cpu().double().numpy()
    dd, ii = cKDTree(t).query(q, k=2)
    np.testing.assert_allclose(d2, dd[:, 0] ** 2, rtol=1e-5, atol=1e-9)
    clear = dd[:, 1] ** 2 - dd[:, 0] ** 2 > 1e-6        # fp32 cannot rank closer pairs
    assert clear.mean() > 0.9
    np.testing.assert_array_equal(idx[clear], ii[clear, 0])


def test_nearest_ties_and_empty_targets():
    from diffskill_amd.pointcloud import nearest
    t = np.array([[0.5, 0.5, 0.5], [0.1, 0.1, 0.1], [0.5, 0.5, 0.5], [0.9, 0.9, 0.9]] * 1500)   # exact duplicates
    q = np.array([[0.5, 0.5, 0.52], [0.12, 0.1, 0.1], [0.9, 0.9, 0.9]])
    idx, d2 = nearest(_cuda(q), _cuda(t))
    assert idx.cpu().tolist() == [0, 1, 3]          # first of the equal-distance targets, across chunks
    assert d2[2].item() == 0.0
    idx, d2 = nearest(_cuda(q), torch.zeros((0, 3), device="cuda"))
    assert idx.cpu().tolist() == [-1, -1, -1] and torch.isinf(d2).all()


@pytest.mark.parametrize("bidirectional,squared", [(False, True), (True, True), (False, False), (True, False)])
def test_chamfer_loss_and_gradients(bidirectional, squared):
    from diffskill_amd.pointcloud import chamfer_distance
    rng = np.random.default_rng(5)
    x = rng.uniform(0.3, 0.7, (3000, 3))
    y = rng.uniform(0.35, 0.65, (2000, 3))
    xg = _cuda(x).requires_grad_(True)
    yg = _cuda(y).requires_grad_(True)
    loss = chamfer_distance(xg, yg, bidirectional=bidirectional, squared=squared)
    loss.backward()
    # float64 reference: full distance matrix, argmin, autograd through the chosen pairs
    xr = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    yr = torch.tensor(y, dtype=torch.float64, requires_grad=True)

    def one_way(a, b):
        j = torch.cdist(a.detach(), b.detach()).argmin(dim=1)
        d2 = ((a - b[j]) ** 2).sum(dim=1)
        return (d2 if squared else torch.sqrt(d2)).mean()

    ref = one_way(xr, yr) + (one_way(yr, xr) if bidirectional else 0)
    ref.backward()
    assert abs(loss.item() - ref.item()) < 1e-5 * ref.item()
    for g, gr in ((xg.grad, xr.grad), (yg.grad, yr.grad)):
        g = g.cpu().double().numpy()
        gr = gr.numpy()
        assert np.abs(g - gr).max() < 1e-4 * np.abs(gr).max(), np.abs(g - gr).max()


def test_plb_chamfer_loss_on_grid_targets():
    """loss.py:183-204 with targets from a mass grid (loss.py:71-75)."""
    from scipy.spatial import cKDTree
    from diffskill_amd.pointcloud import grid_target_points, plb_chamfer_loss
    rng = np.random.default_rng(6)
    grid = np.zeros((32, 32, 32))
    grid[10:20, 2:5, 12:18] = rng.uniform(0, 3e-4, (10, 3, 6))
    tp = grid_target_points(grid, 1 / 32)
    assert tp.shape[0] == int((grid > 1e-4).sum())
    np.testing.assert_allclose(tp[0], np.argwhere(grid > 1e-4)[0] / 32)
    x = rng.uniform([0.3, 0.05, 0.35], [0.6, 0.15, 0.55], (2000, 3))
    loss = plb_chamfer_loss(_cuda(x), tp).item()
    a = cKDTree(tp).query(x)[0].mean()
    b = cKDTree(x).query(tp)[0].mean()
    assert abs(loss - (a + b)) < 1e-5 * (a + b)
