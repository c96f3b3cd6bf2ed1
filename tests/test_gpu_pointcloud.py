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


@pytest.mark.parametrize("n,m,p,blur", [(500, 500, 1, 0.001), (300, 433, 1, 0.001), (200, 150, 2, 0.01), (40, 60, 1, 0.05)])
def test_sinkhorn_divergence_matches_oracle(n, m, p, blur):
    """The HIP soft-min / plan-gradient kernels driving the Sinkhorn loop
    against the fp64 restatement of geomloss's algorithm (oracle/sinkhorn.py;
    parity with geomloss itself is unpinned): the divergence to 1e-4 and its
    gradients in both clouds to 1e-3 of their largest entry.  500 x 500 at
    blur 0.001 is the reference's sampled EMD loss (taichi_env.py:262-270)."""
    import sinkhorn as sk
    from diffskill_amd.pointcloud import sinkhorn_divergence
    print("case sinkhorn %d x %d p=%d blur=%g" % (n, m, p, blur), flush=True)
    rng = np.random.default_rng(n * 7 + m)
    x = rng.uniform(0.3, 0.7, (n, 3)).astype(np.float32).astype(np.float64)
    y = (rng.uniform(0.3, 0.7, (m, 3)) + [0.05, 0.0, -0.03]).astype(np.float32).astype(np.float64)
    xt = _cuda(x).requires_grad_(True)
    yt = _cuda(y).requires_grad_(True)
    s = sinkhorn_divergence(xt, yt, p=p, blur=blur)
    s.backward()
    s_ref = sk.sinkhorn_divergence(x, y, p=p, blur=blur)
    assert abs(float(s.detach()) - s_ref) <= 1e-4 * abs(s_ref), (float(s.detach()), s_ref)
    if p == 1:
        gx_ref = sk.sinkhorn_divergence_grad_x(x, y, blur=blur)
        gy_ref = sk.sinkhorn_divergence_grad_x(y, x, blur=blur)
        gx = xt.grad.double().cpu().numpy()
        gy = yt.grad.double().cpu().numpy()
        assert np.abs(gx - gx_ref).max() <= 1e-3 * np.abs(gx_ref).max()
        assert np.abs(gy - gy_ref).max() <= 1e-3 * np.abs(gy_ref).max()


def test_sinkhorn_divergence_dense_clouds():
    """4000 x 4000 (point spacing 0.025, far above blur) against the fp64
    oracle; 30k x 30k (a whole scene against its target, taichi_env.py:303)
    by properties: the divergence of a cloud with itself is 0, a translation
    by t gives a positive divergence below |t| (the library's annealing with
    scaling 0.5 stops short of the transport cost on clouds this dense) that
    grows with |t|, and the gradient pulls the cloud back along -t."""
    import sinkhorn as sk
    from diffskill_amd.pointcloud import sinkhorn_divergence
    print("case sinkhorn 4000 x 4000", flush=True)
    rng = np.random.default_rng(5)
    x = rng.uniform(0.3, 0.7, (4000, 3)).astype(np.float32).astype(np.float64)
    y = (rng.uniform(0.3, 0.7, (4000, 3)) + [0.03, 0.0, 0.02]).astype(np.float32).astype(np.float64)
    s_ref = sk.sinkhorn_divergence(x, y)
    s = float(sinkhorn_divergence(_cuda(x), _cuda(y)))
    assert abs(s - s_ref) <= 1e-4 * abs(s_ref), (s, s_ref)
    print("case sinkhorn 30000 x 30000", flush=True)
    x = _cuda(rng.uniform(0.3, 0.7, (30000, 3)))
    assert abs(float(sinkhorn_divergence(x, x.clone()))) < 1e-5
    t = torch.tensor([0.04, 0.0, -0.03], device="cuda")
    xt = x.clone().requires_grad_(True)
    s1 = sinkhorn_divergence(xt, x + t)
    s1.backward()
    s2 = float(sinkhorn_divergence(x, x + 0.5 * t))
    tn = float(t.norm())
    assert 0 < s2 < float(s1.detach()) < tn
    g = xt.grad.double().sum(0)
    np.testing.assert_allclose((g / g.norm()).cpu().numpy(), -(t / t.norm()).cpu().numpy(), atol=5e-2)   # within ~3 degrees
