"""The device Krylov-Schur eigensolver of the pairing operator (eliashberg_solver.
solve_eliashberg_krylov_schur) on the CPU, with a dense non-Hermitian test operator in torch:
its leading eigenvalues (largest real part) equal numpy's dense eigenvalues, the returned vectors
are eigenvectors, and restarts are exercised (a small basis against a slowly separating spectrum)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from dgamore_amd.eliashberg_solver import solve_eliashberg_krylov_schur  # noqa: E402


class DenseOp:
    def __init__(self, a, shape):
        self.a = torch.as_tensor(a)
        self.shape = shape
        self.dim = int(np.prod(shape))
        self.g_hat = self.a  # the solver takes its device from here

    def apply(self, x):
        return (self.a @ x.reshape(self.dim)).reshape(x.shape)


def _matrix(n, seed):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(2 * n)
    return a + np.diag(np.linspace(0.0, 1.0, n))  # a spread of real parts, non-Hermitian


@pytest.mark.parametrize("n,ncv", [(240, None), (600, 16)])
def test_leading_eigenpairs_match_dense(n, ncv):
    a = _matrix(n, seed=n)
    op = DenseOp(a, (n // 4, 4))
    lams, vecs, res = solve_eliashberg_krylov_schur(op, 4, tol=1e-12, ncv=ncv)
    ev = np.linalg.eigvals(a)
    ref = ev[np.argsort(-ev.real)][:4]
    assert np.abs(lams - ref).max() / np.abs(ref).max() < 1e-10
    assert np.all(res < 1e-9), res
    for lam, v in zip(lams, vecs):
        v = v.ravel()
        assert np.linalg.norm(a @ v - lam * v) / abs(lam) < 1e-9


def test_start_vector_and_sorting():
    n = 300
    a = _matrix(n, seed=3)
    op = DenseOp(a, (n,))
    lams, _, _ = solve_eliashberg_krylov_schur(op, 3, tol=1e-12, v0=np.ones(n))
    assert np.all(np.diff(lams.real) <= 0)
