"""The oracle's adjoint gradient dL/db (L = 1/2 ||u[obs] - d||^2,
-Lap u - kh^2 b u = 0, b per leaf node) pinned by brute force: central
finite differences (step 1e-6) of the oracle's own exact discrete solve on a
4 x 4-leaf problem, and exact zeros where b does not enter the equations
(face and corner nodes)."""
import numpy as np

from oracle import hps_oracle as O


def _problem():
    n, p, kh = 24, 8, 17.0
    a = 1 / (2 * (n / (p - 2)))
    hps = O.HPSOracle(O.helmholtz_pdo(kh), np.array([[0, 1.0], [0, 1.0]]), a, p)
    part = O.SlabPartition(hps.n0, hps.n1, p, 1.0)
    xx = hps.xx_tot
    b = 1 + 0.1 * np.sin(3 * xx[:, 0]) * np.cos(2 * xx[:, 1])
    uu = lambda x: O.uu_dir_func_greens(x, kh, center=(1.5, 1.5))
    Jc = hps.JJ["Jc"]
    interior = (np.arange(hps.nboxes)[:, None] * p * p + Jc[None, :]).ravel()
    return hps, part, kh, b, uu, interior


def test_oracle_adjoint_gradient_vs_finite_differences():
    hps, part, kh, b, uu, interior = _problem()
    rng = np.random.default_rng(0)
    obs = rng.choice(interior, 40, replace=False)
    d = rng.standard_normal(40) * 0.01
    g, L, u = O.adjoint_gradient(hps, part, kh, b, uu, obs, d)
    assert abs(L - 0.5 * np.sum((u[obs] - d) ** 2)) < 1e-15
    h = 1e-6
    gmax = np.abs(g).max()
    for i in np.concatenate((rng.choice(interior, 12, replace=False), obs[:4])):
        Ls = []
        for s in (1.0, -1.0):
            bp = b.copy()
            bp[i] += s * h
            up, _ = O.helmholtz_nodes_solve(hps, part, kh, bp, uu)
            Ls.append(0.5 * np.sum((up[obs] - d) ** 2))
        fd = (Ls[0] - Ls[1]) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * gmax + 1e-4 * abs(fd), (i, g[i], fd)


def test_oracle_adjoint_zero_off_the_interior():
    """b at face and corner nodes multiplies no equation (those rows are
    flux-continuity rows): zero gradient and no change of the solution."""
    hps, part, kh, b, uu, interior = _problem()
    rng = np.random.default_rng(1)
    obs = rng.choice(interior, 20, replace=False)
    d = np.zeros(20)
    g, L, u = O.adjoint_gradient(hps, part, kh, b, uu, obs, d)
    off = np.setdiff1d(np.arange(g.size), interior)
    assert np.count_nonzero(g[off]) == 0
    bp = b.copy()
    bp[off[::7]] += 0.3
    up, _ = O.helmholtz_nodes_solve(hps, part, kh, bp, uu)
    assert np.array_equal(up, u)
