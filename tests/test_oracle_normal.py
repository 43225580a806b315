"""Pins the oracle's Gauss-Newton normal equations
(gauss_newton_reconstruct, src/Tomography_inverse.jl:489-494; laplace_posterior
:922-924) against what the mathematics fixes:

* a 3 x 3 Jacobian worked by hand (every product and sum written out);
* dense numpy products Ja' diag(W) Ja and Ja' (W r) on random sparse
  Jacobians with an active subset (1e-13 relative: only the summation order
  differs), so a transposed operand, a dropped weight or a wrong column map
  fails;
* symmetry of the normal matrix, positive semi-definiteness;
* zero and non-finite entries are skipped (assemble_forward_and_jacobian,
  src/Tomography.jl:806-818, never passes them on);
* laplace_posterior's every-column form equals the all-active mask.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import _oracle as O


def _rand_csr(n_rows, n_cols, per_row, seed):
    rng = np.random.default_rng(seed)
    rows, cols, vals = [], [], []
    for r in range(n_rows):
        k = rng.integers(0, per_row + 1)
        c = np.sort(rng.choice(n_cols, size=min(k, n_cols), replace=False))
        rows += [r] * len(c)
        cols += list(c)
        vals += list(rng.normal(0.0, 1.0, len(c)) * 10.0 ** rng.uniform(-3, 1, len(c)))
    return sp.csr_matrix((vals, (rows, cols)), shape=(n_rows, n_cols))


def test_hand_example():
    # J = [[1, 2, 0], [0, 3, 4], [5, 0, 6]], W = (1, 2, 3), r = (1, -1, 2)
    J = sp.csr_matrix(np.array([[1.0, 2.0, 0.0], [0.0, 3.0, 4.0], [5.0, 0.0, 6.0]]))
    W = np.array([1.0, 2.0, 3.0])
    r = np.array([1.0, -1.0, 2.0])
    N, g = O.normal_equations(J, r, W)
    want = np.array([[1 * 1 * 1 + 3 * 5 * 5, 1 * 1 * 2, 3 * 5 * 6],
                     [2 * 1 * 1, 1 * 2 * 2 + 2 * 3 * 3, 2 * 3 * 4],
                     [3 * 6 * 5, 2 * 4 * 3, 2 * 4 * 4 + 3 * 6 * 6]], dtype=float)
    assert np.array_equal(N, want)
    assert np.array_equal(g, [1 * 1 * 1 + 5 * 3 * 2, 2 * 1 * 1 + 3 * 2 * -1,
                              4 * 2 * -1 + 6 * 3 * 2])
    # active = columns 0 and 2 only; unweighted
    N, g = O.normal_equations(J, r, None, active=[0, 2])
    assert np.array_equal(N, [[1 + 25, 30], [30, 16 + 36]])
    assert np.array_equal(g, [1 + 10, -4 + 12])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_dense_products(seed):
    J = _rand_csr(300, 120, 12, seed)
    rng = np.random.default_rng(100 + seed)
    W = rng.uniform(0.1, 5.0, J.shape[0])
    r = rng.normal(0.0, 1.0, J.shape[0])
    active = np.flatnonzero(rng.random(J.shape[1]) < 0.6)
    N, g = O.normal_equations(J, r, W, active)
    Ja = J.toarray()[:, active]
    Nd = Ja.T @ (W[:, None] * Ja)
    gd = Ja.T @ (W * r)
    scale = np.abs(Nd).max()
    np.testing.assert_allclose(N, Nd, rtol=1e-13, atol=1e-15 * scale)
    np.testing.assert_allclose(g, gd, rtol=1e-13, atol=1e-15 * np.abs(gd).max())
    np.testing.assert_allclose(N, N.T, rtol=1e-14, atol=1e-15 * scale)
    assert np.linalg.eigvalsh(0.5 * (N + N.T)).min() > -1e-10 * scale
    # unweighted
    N1, g1 = O.normal_equations(J, r, None, active)
    np.testing.assert_allclose(N1, Ja.T @ Ja, rtol=1e-13, atol=1e-15 * np.abs(Ja.T @ Ja).max())
    np.testing.assert_allclose(g1, Ja.T @ r, rtol=1e-13, atol=1e-15 * np.abs(Ja.T @ r).max())


def test_every_column_equals_full_mask():
    J = _rand_csr(80, 40, 8, 7)
    W = np.linspace(0.5, 2.0, 80)
    r = np.sin(np.arange(80.0))
    N0, g0 = O.normal_equations(J, r, W, None)
    N1, g1 = O.normal_equations(J, r, W, np.arange(40))
    assert np.array_equal(N0, N1) and np.array_equal(g0, g1)


def test_zero_and_nonfinite_entries_skipped():
    J = _rand_csr(50, 30, 6, 9).tocoo()
    data = J.data.copy()
    data[::7] = np.nan
    data[3::11] = np.inf
    data[5::13] = 0.0
    Jbad = sp.csr_matrix((data, (J.row, J.col)), shape=J.shape)
    keep = np.isfinite(data) & (data != 0.0)
    Jgood = sp.csr_matrix((data[keep], (J.row[keep], J.col[keep])), shape=J.shape)
    r = np.ones(50)
    Nb, gb = O.normal_equations(Jbad, r)
    Ng, gg = O.normal_equations(Jgood, r)
    assert np.array_equal(Nb, Ng) and np.array_equal(gb, gg)
    assert np.isfinite(Nb).all()


def test_gradient_only_and_empty_rows():
    J = sp.csr_matrix((5, 4))
    N, g = O.normal_equations(J, np.arange(5.0), None, [1, 3])
    assert np.array_equal(N, np.zeros((2, 2))) and np.array_equal(g, [0.0, 0.0])
    N, g = O.normal_equations(_rand_csr(20, 10, 4, 4), np.ones(20), normal=False)
    assert N is None and g.shape == (10,)
