"""Pins of the oracle's ladder solves (CPU only): v1 (inverse) == v3 (solve) == the dense
inverse of a matrix built element by element, residuals, and the beta scaling of the
reference's test_four_point.py:466-532."""
import numpy as np
import pytest

from oracle import dga_oracle as O


def case(nb, niv, nw, nq, seed):
    rng = np.random.default_rng(seed)
    nv = 2 * niv
    chi0_inv = rng.standard_normal((nq,) + (nb,) * 4 + (nw, nv)) + 1j * rng.standard_normal((nq,) + (nb,) * 4 + (nw, nv))
    chi0_inv += 5 * np.einsum("ad,bc->abcd", np.eye(nb), np.eye(nb))[None, ..., None, None]
    gamma = rng.standard_normal((nb,) * 4 + (nw, nv, nv)) + 1j * rng.standard_normal((nb,) * 4 + (nw, nv, nv))
    u = rng.standard_normal((nb,) * 4)
    v = rng.standard_normal((nq,) + (nb,) * 4)
    return chi0_inv, gamma, u, v


@pytest.mark.parametrize("nb,niv", [(1, 3), (2, 2), (3, 1)])
def test_solve_equals_inverse_equals_elementwise_build(nb, niv):
    chi0_inv, gamma, u, v, = case(nb, niv, 2, 2, nb)
    beta = 3.0
    a = O.ladder_sum_inv(chi0_inv, gamma, u, v, beta)
    b = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)
    assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()
    # element-by-element matrix and explicit sum over v'
    nv = 2 * niv
    n = nb * nb * nv
    for q in range(2):
        for w in range(2):
            A = np.zeros((n, n), dtype=complex)
            for o1 in range(nb):
                for o2 in range(nb):
                    for o3 in range(nb):
                        for o4 in range(nb):
                            for p in range(nv):
                                for pp in range(nv):
                                    i = (o1 * nb + o2) * nv + p
                                    j = (o4 * nb + o3) * nv + pp
                                    A[i, j] = (gamma[o1, o2, o3, o4, w, p, pp] - u[o1, o2, o3, o4]
                                               - v[q, o1, o2, o3, o4]) / beta**2
                                    if p == pp:
                                        A[i, j] += chi0_inv[q, o1, o2, o3, o4, w, p]
            Ainv = np.linalg.inv(A)
            for o1, o2, o3, o4, p in [(0, 0, 0, 0, 0), (nb - 1, 0, nb - 1, 0, nv - 1)]:
                ref = sum(Ainv[(o1 * nb + o2) * nv + p, (o4 * nb + o3) * nv + pp] for pp in range(nv)) / beta
                assert abs(b[q, o1, o2, o3, o4, w, p] - ref) < 1e-12 * max(1, abs(ref))


def test_residual_of_solution():
    chi0_inv, gamma, u, v = case(2, 3, 1, 1, 7)
    beta = 2.0
    x = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta) * beta
    A = O.ladder_matrix(chi0_inv[0], gamma, u, v[0], beta)[0]
    nb, nv = 2, 6
    X = x[0, ..., 0, :].transpose(0, 1, 4, 3, 2).reshape(nb * nb * nv, nb * nb)  # rows (o1,o2,v), cols (o4,o3)
    B = O.ladder_rhs(nb, nv)
    assert np.abs(A @ X - B).max() < 1e-12


def test_beta_scaling_of_sum():
    """With A held fixed, the sum scales as 1/beta (test_four_point.py:494-504)."""
    chi0_inv, gamma, u, v = case(1, 4, 2, 2, 9)
    # pre-multiply so that A(beta) is the same matrix for both betas
    b1, b2 = 1.0, 5.0
    g2, u2, v2 = gamma * b2**2, u * b2**2, v * b2**2
    s1 = O.ladder_sum_solve(chi0_inv, gamma * b1**2, u * b1**2, v * b1**2, b1)
    s2 = O.ladder_sum_solve(chi0_inv, g2, u2, v2, b2)
    assert np.allclose(s2, s1 * b1 / b2, rtol=1e-12, atol=1e-14)
