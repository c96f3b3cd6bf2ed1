"""The Sinkhorn-divergence oracle (oracle/sinkhorn.py) pinned by mathematics,
since no golden vector of the third-party geomloss exists in the reference:
S(a, a) = 0, symmetry, the exact 1-Wasserstein distance of small clouds (an
assignment solver), translation, and the gradient against central finite
differences."""
import numpy as np
from scipy.optimize import linear_sum_assignment

import sinkhorn as sk


def cloud(n, seed, lo=0.3, hi=0.7):
    return np.random.default_rng(seed).uniform(lo, hi, (n, 3))


def exact_w1(x, y):
    C = np.sqrt(((x[:, None] - y[None]) ** 2).sum(-1))
    r, c = linear_sum_assignment(C)
    return C[r, c].mean()


def test_self_divergence_is_zero_and_symmetric():
    x, y = cloud(40, 0), cloud(55, 1)
    assert abs(sk.sinkhorn_divergence(x, x)) < 1e-12
    assert abs(sk.sinkhorn_divergence(x, y) - sk.sinkhorn_divergence(y, x)) < 1e-9 * abs(sk.sinkhorn_divergence(x, y))


def test_matches_exact_w1_for_small_blur():
    """For blur far below the point spacing and a slow annealing the
    divergence is the optimal transport cost itself (p = 1: the
    1-Wasserstein distance); the library's default annealing (scaling 0.5,
    one symmetric update per eps) stops short of it, below W1."""
    for seed in range(3):
        x, y = cloud(60, 10 + seed), cloud(60, 20 + seed)
        w1 = exact_w1(x, y)
        s = sk.sinkhorn_divergence(x, y, blur=1e-3, scaling=0.99)
        assert abs(s - w1) <= 6e-3 * w1, (s, w1)
        s_default = sk.sinkhorn_divergence(x, y, blur=1e-3)
        assert 0.85 * w1 < s_default < w1, (s_default, w1)
    # blur above the spacing blurs the plan, so the divergence drops below W1
    x, y = cloud(60, 30), cloud(60, 31)
    assert sk.sinkhorn_divergence(x, y, blur=0.1) < exact_w1(x, y)


def test_translation_costs_its_length():
    x = cloud(50, 3)
    t = np.array([0.05, -0.02, 0.03])
    # W1 of a translation is |t| (the identity matching attains the lower bound
    # |mean(y) - mean(x)|); with the slow annealing the divergence reaches it
    assert abs(sk.sinkhorn_divergence(x, x + t, scaling=0.99) - np.linalg.norm(t)) < 1e-3 * np.linalg.norm(t)
    assert abs(sk.sinkhorn_divergence(x, x + t) - np.linalg.norm(t)) < 1e-2 * np.linalg.norm(t)


def test_schedule_follows_the_annealing_rule():
    x, y = cloud(10, 4), cloud(10, 5)
    s = sk.epsilon_schedule(x, y, 1, 0.001, 0.5)
    d = np.linalg.norm(np.concatenate([x, y]).max(0) - np.concatenate([x, y]).min(0))
    assert s[0] == d and s[-1] == 0.001
    r = np.array(s[1:-1])
    np.testing.assert_allclose(r[1:] / r[:-1], 0.5)
    assert r[0] == d and r[-1] > 0.001 >= r[-1] * 0.5


def test_gradient_matches_finite_differences():
    """The envelope gradient (last step, input potentials fixed) against
    central differences of the whole divergence, with the potentials
    converged (extra iterations at the final eps)."""
    x, y = cloud(12, 6), cloud(15, 7)
    kw = dict(blur=0.02, extra=400)
    g = sk.sinkhorn_divergence_grad_x(x, y, **kw)
    h = 1e-6
    for m, c in [(0, 0), (3, 1), (7, 2), (11, 0), (5, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[m, c] += h
        xm[m, c] -= h
        fd = (sk.sinkhorn_divergence(xp, y, **kw) - sk.sinkhorn_divergence(xm, y, **kw)) / (2 * h)
        assert abs(fd - g[m, c]) <= 1e-5 * np.abs(g).max(), (m, c, fd, g[m, c])


def test_device_schedule_matches_oracle_and_degenerate_clouds():
    """The host schedule of the HIP path (diffskill_amd.pointcloud) equals the
    oracle's; clouds within blur of one point start at blur^p, not at 0."""
    import torch
    from diffskill_amd.pointcloud import epsilon_schedule
    x, y = cloud(30, 8), cloud(20, 9)
    np.testing.assert_allclose(epsilon_schedule(torch.tensor(x), torch.tensor(y), 1, 0.001, 0.5),
                               sk.epsilon_schedule(x, y, 1, 0.001, 0.5), rtol=1e-12)
    one = torch.full((5, 3), 0.5, dtype=torch.float64)
    assert epsilon_schedule(one, one, 1, 0.001, 0.5) == [0.001, 0.001]
