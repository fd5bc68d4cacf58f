"""The reference's iteration body run through the mirror modules on the GPU.

test_reference_iteration_through_the_mirrors restates the call sequence of
calculate_self_energy_q (nonlocal_sde.py:811-928) line by line -- the local vertices loaded from
files as the reference loads them, the FourPoint / SelfEnergy algebra of the reference (negated
double-counting kernel, kernel accumulation, frees), the R-space Sigma that the caller ifft's and
extends, Hartree-Fock, concatenation with the DMFT self-energy, mixing and the iteration file --
and checks the result against the device-resident SDE step of the same model, which the kernel
tests check against the oracle. The linear-mixing branch of apply_mixing_strategy (on the device)
is checked with the reference's own cases (tests/test_mixing.py:110-181).
"""
import os
import sys
from copy import deepcopy

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture
def cfg():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import config

    names = ("sys", "box", "lattice", "self_consistency", "output", "lambda_correction")
    saved = {n: dict(vars(getattr(config, n))) for n in names}
    yield config
    for n, d in saved.items():
        getattr(config, n).__dict__.clear()
        getattr(config, n).__dict__.update(d)


def test_reference_iteration_through_the_mirrors(cfg, tmp_path):
    from dgamore_amd import nonlocal_sde as sde
    from dgamore_amd import synthetic as S
    from dgamore_amd.brillouin_zone import KGrid
    from dgamore_amd.bubble_gen import BubbleGenerator
    from dgamore_amd.four_point import LocalFourPoint, SpinChannel
    from dgamore_amd.pipeline import SDEPipeline
    from dgamore_amd.self_energy import SelfEnergy

    say("reference call sequence through the mirrors: 2 orbitals, 4x4, niw=3, niv=4, shell 2, mixing 0.3")
    m = S.make_model(4, nk=(4, 4, 1), niw=3, niv=4, niv_shell=2)
    nb, beta = m.nb, m.beta
    pipe = SDEPipeline(m, mixing=0.3)
    expected = pipe.step().cpu().numpy()  # the device-resident step from sigma_dmft at mu
    torch.cuda.synchronize()

    out = str(tmp_path)
    grid = KGrid(m.nk)
    cfg.lattice.q_grid = cfg.lattice.k_grid = grid
    cfg.lattice.nk = m.nk
    cfg.sys.beta = beta
    cfg.box.niw_core, cfg.box.niv_core, cfg.box.niv_full = m.niw, m.niv, m.niv_full
    cfg.sys.occ = pipe.occ.cpu().numpy()
    cfg.sys.occ_k = pipe.occ_k.cpu().numpy().reshape(m.nk + (nb, nb))
    cfg.self_consistency.mixing, cfg.self_consistency.mixing_strategy = 0.3, "linear"
    cfg.output.output_path = out
    cfg.lambda_correction.perform_lambda_correction = False
    # the local inputs as files, as the local part of the reference leaves them
    LocalFourPoint(m.f_magn, SpinChannel.MAGN, full_niw_range=False).save(out, "f_magn_loc")
    for ch in ("dens", "magn"):
        LocalFourPoint(m.gamma[ch], SpinChannel(ch), full_niw_range=False).save(out, f"gamma_{ch}_loc")
    sigma_dmft = SelfEnergy(m.sigma_loc[None], (1, 1, 1), True, True)
    sigma_old = sigma_dmft
    u_loc, v_nonloc = m.u, None
    giwk_full = pipe.g_full(pipe.sigma_dmft, m.mu, m.niv_g).cpu().numpy().reshape(m.nk + (nb, nb, 2 * m.niv_g))
    current_iter = 1

    # ---- nonlocal_sde.py:811-893, call for call ------------------------------------------------
    hartree, fock = sde.get_hartree_fock(u_loc, v_nonloc)
    gchi0_q = BubbleGenerator.create_generalized_chi0_q_fft_auto(None, giwk_full, cfg.box.niw_core,
                                                                 cfg.box.niv_full, cfg.lattice.k_grid, beta, None)
    f_dc_loc = 2 * LocalFourPoint.load(os.path.join(out, "f_magn_loc.npy")).permute_orbitals("abcd->cbad")
    kernel = -sde.calculate_sigma_dc_kernel(f_dc_loc, gchi0_q, u_loc)
    f_dc_loc.free()
    gchi0_q_full_sum = 1.0 / beta * gchi0_q.sum_over_all_vn(beta)
    gchi0_q_core = gchi0_q.cut_niv(cfg.box.niv_core)
    gchi0_q.free()
    gchi0_q_core_inv = deepcopy(gchi0_q_core).invert(False)
    gchi0_q_core_sum = 1.0 / beta * gchi0_q_core.sum_over_all_vn(beta)
    gchi0_q_core.free()
    gamma_dens = LocalFourPoint.load(os.path.join(out, "gamma_dens_loc.npy"), SpinChannel.DENS)
    kernel += sde.calculate_sigma_kernel_r_q(gamma_dens, gchi0_q_core_inv, gchi0_q_full_sum, gchi0_q_core_sum,
                                             u_loc, v_nonloc, None)
    gamma_dens.free()
    gamma_magn = LocalFourPoint.load(os.path.join(out, "gamma_magn_loc.npy"), SpinChannel.MAGN)
    kernel += 3 * sde.calculate_sigma_kernel_r_q(gamma_magn, gchi0_q_core_inv, gchi0_q_full_sum, gchi0_q_core_sum,
                                                 u_loc, v_nonloc, None)
    gchi0_q_core_inv.free()
    gchi0_q_full_sum.free()
    gchi0_q_core_sum.free()
    gamma_magn.free()
    sigma_new = sde.calculate_sigma_from_kernel_fft_auto(None, kernel, giwk_full)
    kernel.free()
    assert not sigma_new.full_niv_range and sigma_new.mat.shape == (16, nb, nb, m.niv)
    sigma_new = sigma_new.ifft().to_full_niv_range()
    sigma_new = sigma_new + hartree + fock
    sigma_new += np.zeros((nb, nb, 2 * m.niv))  # delta_sigma: sigma_dmft - sigma_local, zero for this model
    sigma_new = sigma_new.concatenate_self_energies(sigma_dmft)
    sigma_new = sde.apply_mixing_strategy(sigma_new, sigma_old, sigma_dmft, current_iter)
    sigma_new.save(name=f"sigma_dga_iteration_{current_iter}", output_dir=out)
    # ---------------------------------------------------------------------------------------------

    err = np.abs(sigma_new.mat - expected).max() / np.abs(expected).max()
    assert err < 1e-12, err
    back, it = sde.get_starting_sigma(out, sigma_dmft)
    assert it == 1 and np.array_equal(back.mat, sigma_new.mat)
    for ch in ("dens", "magn"):
        chi = np.load(os.path.join(out, f"chi_phys_q_{ch}.npy"))
        assert chi.shape == (grid.nk_irr, nb, nb, nb, nb, m.niw + 1)


def _se(value, niv=8):
    from dgamore_amd.self_energy import SelfEnergy

    return SelfEnergy(np.full((1, 1, 1, 2 * niv), value, dtype=complex), (1, 1, 1), True, True)


@pytest.mark.parametrize("alpha,new,old,expect", [(0.5, 2.0, 0.0, 1.0), (0.0, 5.0, 1.0, 1.0), (1.0, 5.0, 1.0, 5.0),
                                                  (0.5, 2.0 + 2.0j, 0.0, 1.0 + 1.0j)])
def test_linear_mixing_on_the_device(cfg, alpha, new, old, expect):
    from dgamore_amd import nonlocal_sde as sde
    from dgamore_amd.self_energy import SelfEnergy

    say(f"linear mixing alpha={alpha}")
    cfg.self_consistency.mixing_strategy, cfg.self_consistency.mixing = "linear", alpha
    out = sde.apply_mixing_strategy(_se(new), _se(old), _se(0.0), current_iter=1)
    assert isinstance(out, SelfEnergy)
    np.testing.assert_allclose(out.mat, expect, atol=1e-14)


@pytest.mark.parametrize("strategy", ["pulay", "anderson"])
def test_history_mixing_falls_back_to_linear(cfg, strategy):
    """Before mixing_history_length iterations, and with a history without usable directions,
    the reference mixes linearly (test_mixing.py:170-180, 288-299)."""
    from unittest.mock import patch

    from dgamore_amd import nonlocal_sde as sde

    say(f"{strategy} fallbacks to linear mixing")
    cfg.box.niv_core = 4
    cfg.self_consistency.mixing_strategy, cfg.self_consistency.mixing = strategy, 0.5
    cfg.self_consistency.mixing_history_length = 5
    np.testing.assert_allclose(sde.apply_mixing_strategy(_se(2.0), _se(0.0), _se(0.0), current_iter=3).mat, 1.0)
    cfg.self_consistency.mixing_history_length = 3
    v = 3.0 + 1.0j
    files = [np.full((1, 1, 1, 16), v) for _ in range(3)]
    with patch("dgamore_amd.nonlocal_sde.read_last_n_sigmas_from_files", return_value=files):
        out = sde.apply_mixing_strategy(_se(v), _se(v), _se(v), current_iter=10)
    np.testing.assert_allclose(out.mat, v, atol=1e-12)
