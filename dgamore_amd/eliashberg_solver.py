"""Linearised Eliashberg equation with the pairing matvec on the GPU (eliashberg_solver.py:376-471).

solve_eliashberg_lanczos keeps the reference's driver -- scipy's ARPACK eigsh
on a LinearOperator, largest algebraic eigenvalues -- and replaces its matvec
(the einsum / FFT chain of eliashberg_solver.py:420-428) by
dgami_pairing_matvec. The pairing vertex is Fourier transformed and laid out
once (dgami_pairing_setup) and stays resident in HBM; each ARPACK step copies
one gap vector to the device and back.
"""

import numpy as np
import scipy.sparse.linalg as spla

from dgamore_amd import config
from dgamore_amd import kernels as K
from dgamore_amd.four_point import SpinChannel


class PairingOperator:
    """Device-resident linearised gap operator of one channel."""

    def __init__(self, gamma_pp_irr, chi0pp, q_grid, beta, channel):
        import torch

        self.nk = tuple(q_grid.nk)
        self.channel = channel
        self.beta = beta
        gam = np.asarray(gamma_pp_irr)
        self.nb, self.np2 = gam.shape[1], gam.shape[-1]
        self.g_hat, self.g_flip = K.pairing_setup(K.to_dev(gam), K.to_dev(np.asarray(q_grid.irrk_inv), torch.int32),
                                                  self.nk, channel)
        n_k = int(np.prod(self.nk))
        self.chi0pp = K.to_dev(np.asarray(chi0pp).reshape(n_k, self.nb**4, self.np2))
        self.shape = self.nk + (self.nb, self.nb, self.np2)
        self.dim = int(np.prod(self.shape))

    def apply(self, gaps):
        """gaps: device tensor [nvec, nx, ny, nz, nb, nb, np2] -> same."""
        return K.pairing_matvec(self.g_hat, self.g_flip, self.chi0pp, gaps, self.nk, self.channel, self.beta)

    def matvec(self, x):
        d = K.to_dev(np.asarray(x).reshape(self.shape))
        return self.apply(d).cpu().numpy().ravel()


def solve_eliashberg_lanczos(gamma_r_pp, gchi0_q0_pp, ranks=(0, 0), v0=None):
    """Leading eigenvalues and gap functions (eliashberg_solver.py:376-471).

    gamma_r_pp: FourPoint-like with .mat [nq_irr, nb, nb, nb, nb, 2np, 2np] and .channel (SING/TRIP);
    gchi0_q0_pp: FourPoint-like with .mat [nk, nb, nb, nb, nb, 2np] (chi0pp, no -beta factor).
    Returns (lambdas sorted descending, [gap arrays [nx,ny,nz,nb,nb,2np]]).
    """
    ch = gamma_r_pp.channel.value if isinstance(gamma_r_pp.channel, SpinChannel) else str(gamma_r_pp.channel)
    op = PairingOperator(gamma_r_pp.mat, gchi0_q0_pp.mat, config.lattice.q_grid, config.sys.beta, ch)
    lin = spla.LinearOperator(shape=(op.dim, op.dim), matvec=op.matvec, dtype=np.complex128)
    n_eig = config.eliashberg.n_eig
    lambdas, gaps = spla.eigsh(lin, k=n_eig, tol=config.eliashberg.epsilon, v0=v0, which="LA", maxiter=10000)
    order = lambdas.argsort()[::-1]
    lambdas, gaps = lambdas[order], gaps[:, order]
    return lambdas, [gaps[:, i].reshape(op.shape) for i in range(n_eig)]
