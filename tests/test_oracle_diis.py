"""The DIIS restatement against the constrained least-squares problem it solves."""
import numpy as np

from oracle.diis import DIIS


def test_diis_coefficients_minimise_the_error_norm():
    rng = np.random.default_rng(4)
    d = DIIS(space=4)
    x = rng.standard_normal(50)
    d.update(x)
    for step in range(7):
        x = x + 0.3 ** step * rng.standard_normal(50)
        d.update(x)
        nd = sum(v is not None for v in d.xs)
        c = d.coefficients(nd)[1:]
        E = np.array([d.es[i] for i in range(nd)]).T       # columns e_i
        # min |E c| s.t. sum c = 1  <=>  c = c0 + N y with N a basis of sum-zero vectors
        c0 = np.zeros(nd)
        c0[0] = 1.0
        N = np.eye(nd)[:, 1:] - np.eye(nd)[:, :1]
        y = np.linalg.lstsq(E @ N, -(E @ c0), rcond=None)[0]
        ref = c0 + N @ y
        assert abs(c.sum() - 1.0) < 1e-12
        assert np.linalg.norm(E @ c) <= np.linalg.norm(E @ ref) * (1 + 1e-9) + 1e-14


def test_diis_is_exact_on_a_linear_fixed_point_problem():
    """For x <- A x + b in dimension <= space, DIIS hits the fixed point."""
    rng = np.random.default_rng(5)
    n = 3
    A = 0.5 * rng.standard_normal((n, n)) / np.sqrt(n)
    b = rng.standard_normal(n)
    exact = np.linalg.solve(np.eye(n) - A, b)
    d = DIIS(space=6)
    x = np.zeros(n)
    x = d.update(x)
    for _ in range(8):
        x = d.update(A @ x + b)
    assert np.abs(x - exact).max() < 1e-10
