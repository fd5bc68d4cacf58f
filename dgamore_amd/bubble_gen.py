"""BubbleGenerator with the reference's signatures, computed by libdgami (bubble_gen.py:18-279).

The *_cpu / *_gpu / *_auto variants of the reference all map to the one GPU
path here; the MPI distributor and logger arguments are accepted for
signature compatibility and not used (one device, no scatter).
"""

import numpy as np

from dgamore_amd import kernels as K
from dgamore_amd.four_point import FourPoint, FrequencyNotation, SpinChannel


def _giwk_mat(giwk):
    mat = giwk.mat if hasattr(giwk, "mat") else giwk
    if mat.ndim != 6:
        raise ValueError(f"giwk must be [nx,ny,nz,nb,nb,2niv], got shape {mat.shape}")
    return mat


class BubbleGenerator:
    @staticmethod
    def create_generalized_chi0_q_fft_auto(mpi_dist_irrk, giwk, niw, niv, k_grid, beta, logger=None):
        """chi0^{qwv}_{abcd} = -beta/N sum_k G_ad(k,v) G_cb(k-q,v-w) on the irreducible q (bubble_gen.py:33-69)."""
        import torch

        mat = _giwk_mat(giwk)
        nk = tuple(k_grid.nk)
        if mat.shape[:3] != nk:
            raise ValueError(f"giwk grid {mat.shape[:3]} does not match k_grid {nk}")
        ind = K.to_dev(np.asarray(k_grid.irrk_ind), torch.int32)
        chi0 = K.bubble_irrq(K.to_dev(mat), nk, niw, niv, beta, ind)
        nb = mat.shape[3]
        return FourPoint(K.qf_to_ref4(chi0, nb).cpu().numpy(), SpinChannel.NONE, nk, 1, 1, full_niw_range=False,
                         has_compressed_q_dimension=True)

    create_generalized_chi0_q_fft_cpu = staticmethod(
        lambda mpi_dist_irrk, giwk, niw, niv, k_grid, beta: BubbleGenerator.create_generalized_chi0_q_fft_auto(
            mpi_dist_irrk, giwk, niw, niv, k_grid, beta))
    create_generalized_chi0_q_fft_gpu = create_generalized_chi0_q_fft_cpu

    @staticmethod
    def create_generalized_chi0_q_auto(mpi_distributor, giwk, niw, niv, q_list, q_grid, beta, logger=None):
        """Bubble on an explicit q list (bubble_gen.py:147-253); same values via the FFT route."""
        import torch

        mat = _giwk_mat(giwk)
        nk = tuple(q_grid.nk)
        q_list = np.asarray(q_list)
        flat = (q_list[:, 0] * nk[1] + q_list[:, 1]) * nk[2] + q_list[:, 2]
        chi0 = K.bubble_irrq(K.to_dev(mat), nk, niw, niv, beta, K.to_dev(flat, torch.int32))
        return FourPoint(K.qf_to_ref4(chi0, mat.shape[3]).cpu().numpy(), SpinChannel.NONE, nk, 1, 1,
                         full_niw_range=False, has_compressed_q_dimension=True)

    @staticmethod
    def create_generalized_chi0_q_cpu(giwk, niw, niv, q_list, q_grid, beta):
        return BubbleGenerator.create_generalized_chi0_q_auto(None, giwk, niw, niv, q_list, q_grid, beta)

    create_generalized_chi0_q_gpu = create_generalized_chi0_q_cpu

    @staticmethod
    def create_generalized_chi0_q_pp_w0(giwk, niv_pp, q_grid):
        """chi0pp_{abcd}(k,v) = G_ad(k,v) conj(G_cb(k,v)) (bubble_gen.py:268-279), the elementwise input of the
        pairing matvec, computed by dgami_chi0_pp. Returns a FourPoint [nk, nb, nb, nb, nb, 2 niv_pp]."""
        mat = _giwk_mat(giwk)
        nb = mat.shape[3]
        pp = K.chi0_pp(K.to_dev(mat), niv_pp).reshape((-1,) + (nb,) * 4 + (2 * niv_pp,))
        return FourPoint(pp.cpu().numpy(), SpinChannel.NONE, tuple(q_grid.nk), 0, 1, True, True, True,
                         FrequencyNotation.PP)
