"""fp64 CPU restatement of the Sinkhorn divergence the reference uses as its
EMD loss: ``geomloss.SamplesLoss(loss='sinkhorn', p=1, blur=0.001)``
(plb/engine/taichi_env.py:23-26, called at :269-270 and :303, :317).

TEST INFRASTRUCTURE ONLY: imported by ``tests/`` as the checker of the HIP
point-cloud kernels; the product package never imports it.

geomloss is a third-party dependency that is absent here; the reference pins
only its companion pykeops 1.5 (README.md:9), whose contemporary release is
geomloss 0.2.4.  Its published algorithm (Feydy et al., "Interpolating between
Optimal Transport and MMD using Sinkhorn Divergences", AISTATS 2019; the
geomloss documentation of ``SamplesLoss``) is restated here:

- cost C(x, y) = |x - y|^p / p (p = 1: the Euclidean distance, as in the
  library's online backend; its tensorized backend clamps |x - y|^2 at 1e-8
  before the square root);
- uniform weights alpha = 1/N, beta = 1/M;
- epsilon scaling: eps_s = [d^p] + exp(arange(p log d, p log blur, p log
  scaling)) + [blur^p], d the diagonal of the two clouds' bounding box;
- softmin(eps, C, h)_i = -eps log sum_j exp(h_j - C_ij / eps);
- initial potentials from h = log weights, then per eps symmetric coordinate
  ascent (each new potential averaged with the old one) on the x-y problem and
  the two debiasing problems x-x and y-y;
- a last extrapolation step at the final eps (all four potentials updated
  from the previous ones at once);
- S = <alpha, b_x - a_x> + <beta, a_y - b_y>.

Parity unpinned: no golden vector of geomloss exists in the reference.  The
checks in ``tests/test_sinkhorn_oracle.py`` pin it against the mathematics
instead (the debiased divergence of a cloud with itself is 0; symmetry; the
exact 1-Wasserstein distance of small equal-size clouds from an assignment
solver, which S approaches for blur far below the point spacing; a translation
by t costs |t|; the gradient against central finite differences).
"""
import numpy as np


def _cost(x, y, p):
    d = np.sqrt(((x[:, None, :] - y[None, :, :]) ** 2).sum(-1))
    return d if p == 1 else d ** p / p


def _softmin(eps, C, h):
    z = h[None, :] - C / eps
    m = z.max(1, keepdims=True)
    return -eps * (np.log(np.exp(z - m).sum(1)) + m[:, 0])


def epsilon_schedule(x, y, p, blur, scaling):
    """The annealing schedule of geomloss's ``scaling_parameters``."""
    pts = np.concatenate([x, y])
    diameter = float(np.linalg.norm(pts.max(0) - pts.min(0)))
    return [diameter ** p] + [float(np.exp(e)) for e in np.arange(p * np.log(diameter), p * np.log(blur), p * np.log(scaling))] + \
        [blur ** p]


def sinkhorn_potentials(x, y, p=1, blur=0.001, scaling=0.5, extra=0):
    """(eps, a_x, b_x, a_y, b_y) after the last extrapolation, plus the
    potentials they were extrapolated from (for the gradient).  ``extra``
    repeats the final eps that many more times (0: the library's schedule;
    the tests converge the potentials with it)."""
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    N, M = len(x), len(y)
    la, lb = np.full(N, -np.log(N)), np.full(M, -np.log(M))
    Cxy, Cyx = _cost(x, y, p), _cost(y, x, p)
    Cxx, Cyy = _cost(x, x, p), _cost(y, y, p)
    eps_s = epsilon_schedule(x, y, p, blur, scaling)
    eps_s = eps_s + [eps_s[-1]] * int(extra)
    eps = eps_s[0]
    a_y = _softmin(eps, Cyx, la)
    b_x = _softmin(eps, Cxy, lb)
    a_x = _softmin(eps, Cxx, la)
    b_y = _softmin(eps, Cyy, lb)
    for eps in eps_s:
        at_y = _softmin(eps, Cyx, la + b_x / eps)
        bt_x = _softmin(eps, Cxy, lb + a_y / eps)
        at_x = _softmin(eps, Cxx, la + a_x / eps)
        bt_y = _softmin(eps, Cyy, lb + b_y / eps)
        a_y, b_x = 0.5 * (a_y + at_y), 0.5 * (b_x + bt_x)
        a_x, b_y = 0.5 * (a_x + at_x), 0.5 * (b_y + bt_y)
    prev = (a_x, b_x, a_y, b_y)
    a_y, b_x = _softmin(eps, Cyx, la + prev[1] / eps), _softmin(eps, Cxy, lb + prev[2] / eps)
    a_x = _softmin(eps, Cxx, la + prev[0] / eps)
    b_y = _softmin(eps, Cyy, lb + prev[3] / eps)
    return eps, (a_x, b_x, a_y, b_y), prev


def sinkhorn_divergence(x, y, p=1, blur=0.001, scaling=0.5, extra=0):
    """S_eps(alpha, beta) of the two uniformly weighted clouds."""
    _, (a_x, b_x, a_y, b_y), _ = sinkhorn_potentials(x, y, p, blur, scaling, extra)
    return float(np.mean(b_x - a_x) + np.mean(a_y - b_y))


def _unit(d):
    n = np.sqrt((d ** 2).sum(-1, keepdims=True))
    return np.where(n > 0, d / np.where(n > 0, n, 1.0), 0.0)


def sinkhorn_divergence_grad_x(x, y, p=1, blur=0.001, scaling=0.5, extra=0):
    """dS/dx as the library's autograd gives it, p = 1: its cost matrices
    detach their second argument (C_xy = C(x, y.detach()), C_xx = C(x,
    x.detach())) and the last extrapolation detaches its input potentials, so
    only the first-argument dependence of b_x and a_x is differentiated -- the
    envelope-theorem gradient of the divergence at converged potentials:
    g_m = alpha [sum_j P_mj u(x_m - y_j) - sum_k R_mk u(x_m - x_k)],
    P, R the row-normalised kernels of the softmins b_x (x over y) and a_x
    (x over x), u(d) = d / |d|."""
    assert p == 1
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    N, M = len(x), len(y)
    eps, (a_x, b_x, a_y, b_y), (pa_x, pb_x, pa_y, pb_y) = sinkhorn_potentials(x, y, p, blur, scaling, extra)
    Cxy, Cxx = _cost(x, y, 1), _cost(x, x, 1)
    P = np.exp(-np.log(M) + pa_y[None, :] / eps - Cxy / eps + b_x[:, None] / eps)   # [m, j]
    R = np.exp(-np.log(N) + pa_x[None, :] / eps - Cxx / eps + a_x[:, None] / eps)   # [m, k]
    uxy = _unit(x[:, None, :] - y[None, :, :])
    uxx = _unit(x[:, None, :] - x[None, :, :])
    return (1.0 / N) * (np.einsum("mj,mjc->mc", P, uxy) - np.einsum("mk,mkc->mc", R, uxx))
