"""lambda-correction of the physical susceptibility, FP64 on the host (lambda_correction.py:13-73).

The root find is a short scalar loop over nq*(2niw+1) numbers; it stays in
double precision on the host as in the reference (its arrays there are complex64).
"""

import numpy as np

from dgamore_amd import config


def get_lambda_start(chi_r):
    """-min(1/Re chi(w=0)) (lambda_correction.py:13-18)."""
    w0 = chi_r.shape[-1] // 2
    return -np.min(1.0 / chi_r[..., w0].real)


def apply_lambda(chi_r, lambda_):
    """1/(1/chi + lambda) (lambda_correction.py:21-25)."""
    return 1.0 / (1.0 / chi_r + lambda_)


def find_lambda_weighted(chi_r_mat, chi_r_loc_sum, counts, beta, nk_tot, delta=0.1, eps=1e-7, maxiter=1000):
    """Newton-like iteration with step halving (lambda_correction.py:28-60): the lambda with
    sum_q counts[q] chi/(1 + lambda chi) / (beta nk_tot) = chi_r_loc_sum; chi_r_mat [q_irr, 2niw+1]."""
    chi_r_mat = np.asarray(chi_r_mat, dtype=np.complex128)
    counts = np.asarray(counts)[:, None]
    lambda_start = get_lambda_start(chi_r_mat)
    lambda_ = lambda_start + delta
    factor = 1.0 / beta / nk_tot
    for _ in range(maxiter):
        chi_lam = apply_lambda(chi_r_mat, lambda_)
        f_lam = (counts * chi_lam).sum() * factor - chi_r_loc_sum
        fp_lam = -(counts * chi_lam**2).sum() * factor
        lambda_new = lambda_ - (f_lam / fp_lam).real
        if abs(f_lam.real) < eps:
            return lambda_new
        if lambda_new < lambda_:
            delta /= 2
            lambda_ = lambda_start + delta
        else:
            lambda_ = lambda_new
    return lambda_


def find_lambda(chi_r_mat, chi_r_loc_sum, delta=0.1, eps=1e-7, maxiter=1000):
    """find_lambda_weighted with the irreducible weights, beta and grid size of the module config."""
    return find_lambda_weighted(chi_r_mat, chi_r_loc_sum, config.lattice.q_grid.irrk_count, config.sys.beta,
                                config.lattice.q_grid.nk_tot, delta, eps, maxiter)


def perform_single_lambda_correction(chi_r, chi_r_loc_sum):
    """(corrected chi in the half bosonic range, lambda) for one channel (lambda_correction.py:63-73)."""
    chi_r = chi_r.to_full_niw_range()
    chi_mat = chi_r.compress_q_dimension().mat.squeeze()
    lam = find_lambda(chi_mat, chi_r_loc_sum)
    chi_r.mat = apply_lambda(chi_mat, lam)[:, None, None, None, None, :]
    return chi_r.to_half_niw_range(), lam


def perform_lambda_correction(chi_phys_q_r):
    """lambda correction of one channel's physical susceptibility with the local sums read from the
    files the reference reads (nonlocal_sde.py:339-401): chi_{dens,magn}_loc.npy and, for 'sp',
    chi_phys_q_dens.npy from config.output.output_path; lambda is appended to lambda_{type}.txt.
    'spch' corrects both channels, 'sp' only the magnetic one."""
    import os

    from dgamore_amd.four_point import FourPoint, LocalFourPoint, SpinChannel

    kind = config.lambda_correction.type.lower()
    if kind not in ("spch", "sp"):
        raise ValueError("Lambda correction type must be either 'spch' or 'sp'.")
    out = config.output.output_path
    ch = chi_phys_q_r.channel

    def loc(channel):
        return LocalFourPoint.load(os.path.join(out, f"chi_{channel.value}_loc.npy"), channel,
                                   num_vn_dimensions=0).to_full_niw_range()

    if kind == "spch":
        target = loc(ch).mat.sum() / config.sys.beta
    elif ch != SpinChannel.MAGN:
        return chi_phys_q_r
    else:
        dens = FourPoint.load(os.path.join(out, "chi_phys_q_dens.npy"), SpinChannel.DENS,
                              num_vn_dimensions=0).to_full_niw_range()
        counts = np.asarray(config.lattice.q_grid.irrk_count)[:, None, None, None, None, None]
        total = (loc(SpinChannel.DENS).mat + loc(SpinChannel.MAGN).mat).sum()
        target = (total - (counts * dens.mat).sum() / config.lattice.q_grid.nk_tot) / config.sys.beta
    chi_phys_q_r, lam = perform_single_lambda_correction(chi_phys_q_r, target)
    with open(os.path.join(out, f"lambda_{config.lambda_correction.type}.txt"), "a") as f:
        f.write(f"lambda_{ch.value}: {lam}\n")
    return chi_phys_q_r
