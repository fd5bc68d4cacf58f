"""The FP64 lambda-correction module against the reference's known answers (CPU only)."""
import numpy as np

from dgamore_amd import config, lambda_correction as lc
from dgamore_amd.brillouin_zone import KGrid
from dgamore_amd.four_point import FourPoint, SpinChannel
from oracle import dga_oracle as O
from tests.conftest import golden


def test_spch_magn_known_answer():
    config.lattice.q_grid = KGrid((4, 4, 1))
    config.sys.beta = 12.5
    chi = FourPoint(golden("lc_spch_chi_phys_q_magn_before.npy"), SpinChannel.MAGN, (4, 4, 1), 1, 0, False,
                    has_compressed_q_dimension=True)
    loc = O.to_full_niw(golden("lc_spch_chi_magn_loc.npy"), 0).sum() / 12.5
    corrected, lam = lc.perform_single_lambda_correction(chi, loc)
    assert np.isclose(lam, 4.328781)  # test_lambda_correction.py:71
    ref = golden("lc_spch_chi_phys_q_magn.npy")
    assert np.abs(corrected.mat - ref).max() / np.abs(ref).max() < 1e-5


def test_sp_known_answer():
    config.lattice.q_grid = KGrid((4, 4, 1))
    config.sys.beta = 12.5
    chi_d = O.to_full_niw(golden("lc_sp_chi_phys_q_dens_before.npy"), 0)
    s = O.to_full_niw(golden("lc_sp_chi_dens_loc.npy"), 0).sum() + O.to_full_niw(golden("lc_sp_chi_magn_loc.npy"), 0).sum() \
        - 1 / 16 * (config.lattice.q_grid.irrk_count[:, None, None, None, None, None] * chi_d).sum()
    chi = FourPoint(golden("lc_sp_chi_phys_q_magn_before.npy"), SpinChannel.MAGN, (4, 4, 1), 1, 0, False,
                    has_compressed_q_dimension=True)
    _, lam = lc.perform_single_lambda_correction(chi, s / 12.5)
    assert np.isclose(lam, 4.281153)  # test_lambda_correction.py:102


def test_lambda_corrected_step_is_not_capturable():
    """The lambda correction reads chi on the host: capture() refuses it before touching the GPU."""
    import pytest

    from dgamore_amd.pipeline import SDEPipeline

    pipe = SDEPipeline.__new__(SDEPipeline)
    pipe.precision, pipe.lambda_correction = 0, "sp"
    with pytest.raises(ValueError):
        pipe.capture()
