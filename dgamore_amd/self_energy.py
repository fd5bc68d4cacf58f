"""Self-energy helpers of the self-consistency loop (self_energy.py:51-75, 153-168).

The momentum average of SelfEnergy.fit_smom runs in libdgami (dgami_kmean); the moment fit
itself reads 2 nb^2 n_fit numbers and stays on the host, as in the reference.
"""

import numpy as np
import torch

from dgamore_amd import kernels as K


def fit_smom(sigma_mat, beta):
    """(mom0, mom1) [nb,nb]: mean of Re Sigma and of Im Sigma * v over the top 20% (at least 4) positive
    frequencies of the k-averaged self-energy (self_energy.py:51-75).

    sigma_mat: [..., nb, nb, 2niv] with any number of leading momentum axes (none: local).
    """
    s = sigma_mat if isinstance(sigma_mat, torch.Tensor) else K.to_dev(np.asarray(sigma_mat))
    nb, nv = s.shape[-2], s.shape[-1]
    loc = K.kmean(s.reshape(-1, nb, nb, nv).contiguous()).cpu().numpy()
    niv = nv // 2
    n_fit = max(int(0.2 * niv), 4)
    fit = loc[..., nv - n_fit:]
    v = np.pi / beta * (2 * np.arange(niv - n_fit, niv) + 1)
    return fit.real.mean(axis=-1), (fit.imag * v[None, None, :]).mean(axis=-1)
