"""Hot-path functions of the reference's nonlocal_sde module, computed by libdgami.

Same names and argument meaning as dgamore/nonlocal_sde.py; they read the
module-level configuration (dgamore_amd.config: sys.beta, box.niv_core,
lattice.k_grid/q_grid) exactly where the reference reads its config. Objects
are FourPoint containers (dgamore_amd.four_point) with numpy storage; each
call uploads its operands, runs the kernels and downloads the result. For a
device-resident iteration use dgamore_amd.pipeline.SDEPipeline.

Interactions: ``u_loc`` is a [nb]*4 array or an object with ``.mat``;
``v_nonloc`` is a [nq, nb, nb, nb, nb] array (irreducible q), an object with
``.mat``, or None for V = 0.
"""

import glob
import os
import re

import numpy as np

from dgamore_amd import config
from dgamore_amd import kernels as K
from dgamore_amd.four_point import FourPoint, SpinChannel
from dgamore_amd.pipeline import u_as_channel


def _mat(x):
    return None if x is None else (x.mat if hasattr(x, "mat") else np.asarray(x))


def _ch(channel):
    return channel.value if isinstance(channel, SpinChannel) else str(channel)


def _v_as_channel(v, channel):
    """V^q in a channel (interaction.py:160-182): dens 2V, magn 0, sing/trip V."""
    if v is None:
        return None
    return {"dens": 2.0, "magn": 0.0, "sing": 1.0, "trip": 1.0}[_ch(channel)] * v


def _dev_m(x):
    return None if x is None else K.orb_to_m(K.to_dev(x))


def _ladder_operands(gamma_r, gchi0_q_inv, u_loc, v_nonloc):
    """Device operands of one channel's ladder: (chi0_inv QF, Gamma CMP, U_r M order, V_r QO or None, inv)."""
    ch = _ch(gamma_r.channel)
    inv = gchi0_q_inv.to_half_niw_range().compress_q_dimension()
    gam = gamma_r.mat if gamma_r.mat.ndim == 7 else None
    if gam is None:
        raise ValueError("gamma_r must be a local vertex [o1,o2,o3,o4,w,v,v']")
    if gam.shape[4] != inv.mat.shape[5] or gam.shape[5] != inv.mat.shape[6]:
        raise ValueError(f"gamma_r frequency box {gam.shape[4:]} does not match chi0^-1 {inv.mat.shape[5:]}")
    ur = u_as_channel(_mat(u_loc), ch)
    vr = _v_as_channel(_mat(v_nonloc), ch)
    return (K.ref4_to_qf(K.to_dev(inv.mat)), K.vertex_to_cmp(K.to_dev(gam)), _dev_m(ur),
            None if vr is None else _dev_m(vr), inv)


def create_auxiliary_chi_r_q(gamma_r, gchi0_q_inv, u_loc, v_nonloc):
    """chi*_r(q) = ((chi0^-1 + Gamma_r/beta^2) - (V_r + U_r)/beta^2)^-1, the full inverse of every
    (q, w) ladder matrix (nonlocal_sde.py:72-82 with FourPoint.invert, four_point.py:590-600).

    Returns a FourPoint [q, o1, o2, o3, o4, w(>=0), v, v'] (two fermionic axes, half bosonic range).
    """
    chi0_inv, gam, ur, vr, inv = _ladder_operands(gamma_r, gchi0_q_inv, u_loc, v_nonloc)
    x, _ = K.ladder_invert(chi0_inv, gam, ur, vr, config.sys.beta)
    return FourPoint(K.qcmp_to_ref8(x, inv.n_bands).cpu().numpy(), gamma_r.channel, inv.nq, 1, 2, False,
                     has_compressed_q_dimension=True)


def create_auxiliary_chi_r_q_sum_v1(gamma_r, gchi0_q_inv, u_loc, v_nonloc):
    """sum_v' chi*_r = sum_v' ((chi0^-1 + Gamma_r/beta^2) - (V_r + U_r)/beta^2)^-1 (nonlocal_sde.py:85-96)."""
    import torch

    beta = config.sys.beta
    ch = _ch(gamma_r.channel)
    inv = gchi0_q_inv.to_half_niw_range().compress_q_dimension()
    nb = inv.n_bands
    gam = gamma_r.mat if gamma_r.mat.ndim == 7 else None
    if gam is None:
        raise ValueError("gamma_r must be a local vertex [o1,o2,o3,o4,w,v,v']")
    if gam.shape[4] != inv.mat.shape[5] or gam.shape[5] != inv.mat.shape[6]:
        raise ValueError(f"gamma_r frequency box {gam.shape[4:]} does not match chi0^-1 {inv.mat.shape[5:]}")
    ur = u_as_channel(_mat(u_loc), ch)
    vr = _v_as_channel(_mat(v_nonloc), ch)
    aux, _ = K.ladder_solve_sum(K.ref4_to_qf(K.to_dev(inv.mat)), K.vertex_to_cmp(K.to_dev(gam))[None],
                                _dev_m(ur)[None], None if vr is None else _dev_m(vr)[None], beta,
                                int(config.precision.ladder))
    return FourPoint(K.qf_to_ref4(aux[0], nb).cpu().numpy(), gamma_r.channel, inv.nq, 1, 1, False,
                     has_compressed_q_dimension=True)


def create_auxiliary_chi_r_q_sum_v3(gamma_r, gchi0_q_inv, u_loc, v_nonloc, mpi_dist_irrq=None):
    """Solve-based variant of the reference (nonlocal_sde.py:129-156); identical result and kernel here."""
    return create_auxiliary_chi_r_q_sum_v1(gamma_r, gchi0_q_inv, u_loc, v_nonloc)


create_auxiliary_chi_r_q_sum_v2 = create_auxiliary_chi_r_q_sum_v3


def create_vrg_r_q(gchi_aux_q_r_sum, gchi0_q_inv):
    """gamma_r = beta chi0^-1 . sum_v' chi*_r per (q,w,v) (nonlocal_sde.py:159-164)."""
    import torch

    beta = config.sys.beta
    nb = gchi_aux_q_r_sum.n_bands
    aux = K.ref4_to_qf(K.to_dev(gchi_aux_q_r_sum.mat))
    inv = K.ref4_to_qf(K.to_dev(gchi0_q_inv.mat))
    nw, _, nv, nq = aux.shape
    vrg = torch.empty_like(aux)
    from dgamore_amd import _lib

    _lib.call("dgami_vrg_chi_phys", nq, nb, nw, nv // 2, float(beta), K._p(inv), K._p(aux), None, None, None, None,
              K._p(vrg), None, None, K._stream())
    return FourPoint(K.qf_to_ref4(vrg, nb).cpu().numpy(), gchi_aux_q_r_sum.channel, gchi_aux_q_r_sum.nq, 1, 1, False,
                     has_compressed_q_dimension=True)


def create_generalized_chi_q_with_shell_correction(gchi_aux_q_sum, gchi0_q_full_sum, gchi0_q_core_sum, u_loc,
                                                   v_nonloc):
    """chi_r = ((chi*_sum + chi0_full - chi0_core)^-1 + U_r + V_r)^-1 (nonlocal_sde.py:167-182).

    gchi_aux_q_sum here is the frequency-summed object (num_vn = 0), as in the reference.
    """
    import torch

    from dgamore_amd import _lib

    ch = _ch(gchi_aux_q_sum.channel)
    x = gchi_aux_q_sum.mat + _mat(gchi0_q_full_sum) - _mat(gchi0_q_core_sum)
    nq, nb, nw = x.shape[0], x.shape[1], x.shape[5]
    ur = u_as_channel(_mat(u_loc), ch)
    vr = _v_as_channel(_mat(v_nonloc), ch)
    # the chi_phys kernel forms sum_v aux / beta: feed the summed object as two frequencies (x, 0) with beta = 1
    aux = K.ref4_to_qf(K.to_dev(np.stack([x, np.zeros_like(x)], axis=-1)))
    zero = torch.zeros((nw, nb**4, nq), dtype=K.C128, device=aux.device)
    chi = torch.empty((nw, nb**4, nq), dtype=K.C128, device=aux.device)
    info = torch.zeros((nw, nq), dtype=torch.int32, device=aux.device)
    _lib.call("dgami_vrg_chi_phys", nq, nb, nw, 1, 1.0, None, K._p(aux), K._p(zero), K._p(zero), K._p(_dev_m(ur)),
              K._p(_dev_m(vr)) if vr is not None else None, None, K._p(chi), K._p(info, torch.int32), K._stream())
    if int(torch.count_nonzero(info)):
        raise np.linalg.LinAlgError("Singular matrix")
    return FourPoint(K.qw_to_ref4w(chi, nb).cpu().numpy(), gchi_aux_q_sum.channel, gchi_aux_q_sum.nq, 1, 0, False,
                     has_compressed_q_dimension=True)


def calculate_sigma_dc_kernel(f_dc_loc, gchi0_q, u_loc):
    """K_dc with f_dc = 2 F_magn,loc[abcd->cbad] already applied by the caller (nonlocal_sde.py:185-198).

    Returns +K_dc (the reference's caller negates it at :820).
    """
    beta = config.sys.beta
    niv_core = config.box.niv_core
    chi0 = gchi0_q.to_half_niw_range().compress_q_dimension()
    nb = chi0.n_bands
    # libdgami takes F_magn,loc and applies the factor 2 and the abcd->cbad permutation itself
    f_magn = 0.5 * np.einsum("cbad...->abcd...", _mat(f_dc_loc))
    out = K.dc_kernel(K.ref4_to_qf(K.to_dev(chi0.mat)), K.to_dev(_mat(u_loc)), K.to_dev(f_magn), nb, niv_core, beta)
    return FourPoint(-K.qf_to_ref4(out, nb).cpu().numpy(), SpinChannel.NONE, chi0.nq, 1, 1, False,
                     has_compressed_q_dimension=True)


def calculate_kernel_r_q(vrg_q_r, gchi_aux_q_r_sum, v_nonloc, u_loc):
    """K_r = u_r . (gamma_r - gamma_r . u_r . chi_r - [magn] 2/3 1) (nonlocal_sde.py:201-215)."""
    ch = _ch(vrg_q_r.channel)
    nb = vrg_q_r.n_bands
    ur = u_as_channel(_mat(u_loc), ch)
    vr = _v_as_channel(_mat(v_nonloc), ch)
    vrg = K.ref4_to_qf(K.to_dev(vrg_q_r.mat))
    chi = K.ref4w_to_qw(K.to_dev(gchi_aux_q_r_sum.mat))
    ker = K.kernel_r(vrg, chi, _dev_m(ur), _dev_m(vr), nb, ch, 1.0)
    return FourPoint(K.qf_to_ref4(ker, nb).cpu().numpy(), vrg_q_r.channel, vrg_q_r.nq, 1, 1, False,
                     has_compressed_q_dimension=True)


def calculate_sigma_from_kernel_fft_auto(mpi_distributor, kernel, giwk):
    """Sigma from the irreducible kernel, returned as the reference returns it: a SelfEnergy in real
    space R, nu >= 0, compressed momentum axis (nonlocal_sde.py:544-594, 650-674), which the caller
    gathers, ifft's and extends to the full fermionic range (nonlocal_sde.py:874-879).

    The irr->FBZ map (mpi_utils.py:123-131) is fused into the Sigma kernel, which produces
    Sigma(k, nu); its nu >= 0 half is brought to R by the device FFT, so that SelfEnergy.ifft()
    returns Sigma(k, nu) exactly.
    """
    import torch

    from dgamore_amd.self_energy import SelfEnergy

    beta = config.sys.beta
    grid = config.lattice.q_grid
    ker = kernel.to_half_niw_range().compress_q_dimension()
    if ker.mat.shape[0] != len(grid.irrk_ind):
        raise ValueError("kernel must be given on the irreducible q-points of config.lattice.q_grid")
    g = _mat(giwk)
    niw = ker.mat.shape[5] - 1
    nk = tuple(grid.nk)
    sig = K.sigma_from_kernel(K.ref4_to_qf(K.to_dev(ker.mat)), K.to_dev(np.asarray(grid.irrk_inv), torch.int32),
                              K.to_dev(g), nk, niw, beta)  # [nx,ny,nz,nb,nb,2niv] in k
    nb, nv = sig.shape[3], sig.shape[-1]
    half = sig[..., nv // 2:].reshape(-1, nb * nb * (nv // 2)).T.contiguous()  # [rest][N]
    K.fft_grid(half, nk, -1, 1.0)  # k -> R, exp(-i k.R), unnormalised (scipy.fft.fftn)
    mat = half.T.contiguous().cpu().numpy().reshape(-1, nb, nb, nv // 2)
    return SelfEnergy(mat, config.lattice.nk, full_niv_range=False, has_compressed_q_dimension=True, calc_smom=False)


def calculate_sigma_from_kernel_auto(mpi_distributor, kernel, giwk, my_full_q_list=None):
    """Direct q-sum form of the reference (nonlocal_sde.py:404-541); same values by the FFT route.
    The reference's direct form returns Sigma(k, nu >= 0) on the momentum grid already: the same
    object after ifft."""
    return calculate_sigma_from_kernel_fft_auto(mpi_distributor, kernel, giwk).ifft()


def calculate_sigma_kernel_r_q(gamma_r, gchi0_q_inv, gchi0_q_full_sum, gchi0_q_core_sum, u_loc, v_nonloc,
                               mpi_dist_irrq=None):
    """The kernel of one spin channel (nonlocal_sde.py:282-336): ladder sums, three-leg vertex,
    shell-corrected physical susceptibility (lambda-corrected if config.lambda_correction asks for
    it; single band), then K_r. With config.output.output_path set, chi_phys_q_{channel} is saved
    in the half bosonic range as the reference saves it."""
    aux = create_auxiliary_chi_r_q_sum_v1(gamma_r, gchi0_q_inv, u_loc, v_nonloc)
    vrg = create_vrg_r_q(aux, gchi0_q_inv)
    chi = create_generalized_chi_q_with_shell_correction(aux.sum_over_all_vn(config.sys.beta), gchi0_q_full_sum,
                                                         gchi0_q_core_sum, u_loc, v_nonloc)
    aux.free()
    if getattr(config.lambda_correction, "perform_lambda_correction", False):
        from dgamore_amd.lambda_correction import perform_lambda_correction

        chi = perform_lambda_correction(chi)
    out = getattr(getattr(config, "output", None), "output_path", None)
    if out:
        chi.save(name=f"chi_phys_q_{chi.channel.value}", output_dir=out)
    return calculate_kernel_r_q(vrg, chi, v_nonloc, u_loc)


def get_hartree_fock(u_loc, v_nonloc, q_list=None):
    """Hartree [1,nb,nb,1] and Fock [nk,nb,nb,1] terms (nonlocal_sde.py:31-69):

        Sigma_H = 2 (U + V^{q=0})_abcd n_dc,   Sigma_F(k) = -(1/N_q) sum_q (U + V^q)_abcd n_dc(k+q),

    with n = config.sys.occ and n(k) = config.sys.occ_k on config.lattice.nk. v_nonloc covers the full
    momentum grid ([nx,ny,nz,nb,nb,nb,nb] or flattened, q = 0 first) or is None. The sum over every q
    of the grid is one FFT cross-correlation (q_list is accepted for signature compatibility).
    """
    nk = tuple(config.lattice.nk)
    u = np.asarray(_mat(u_loc), dtype=complex)
    nb = u.shape[0]
    v = _mat(v_nonloc)
    dv = None if v is None else K.to_dev(np.asarray(v, dtype=complex).reshape(-1, nb, nb, nb, nb))
    occ_k = np.asarray(config.sys.occ_k, dtype=complex).reshape(-1, nb, nb)
    hart, fock = K.hartree_fock(K.to_dev(u), dv, K.to_dev(np.asarray(config.sys.occ, dtype=complex)),
                                K.to_dev(occ_k), nk)
    return hart.cpu().numpy()[None, ..., None], fock.cpu().numpy()[..., None]


def assemble_sigma(sigma_nl, hartree, fock, delta_sigma, sigma_dmft, niv_out, sigma_old=None, alpha=1.0):
    """New self-energy of one iteration, [nk,nb,nb,2 niv_out] (nonlocal_sde.py:884-893):
    sigma + hartree + fock + delta_sigma on the core window, concatenated with the DMFT
    self-energy outside it (self_energy.py:153-168), then alpha new + (1 - alpha) old when sigma_old
    is given (the linear branch of apply_mixing_strategy, nonlocal_sde.py:1145-1151)."""
    snl = np.asarray(sigma_nl)
    nb = snl.shape[-2]
    nk = int(np.prod(snl.shape[:-3]))
    dev = lambda x, shape: None if x is None else K.to_dev(np.asarray(x, dtype=complex).reshape(shape))  # noqa
    nvc = snl.shape[-1]
    out = K.sigma_assemble(K.to_dev(snl.reshape(nk, nb, nb, nvc)), dev(hartree, (nb, nb)), dev(fock, (nk, nb, nb)),
                           dev(delta_sigma, (nb, nb, nvc)), K.to_dev(np.asarray(sigma_dmft).reshape(nb, nb, -1)),
                           niv_out, dev(sigma_old, (nk, nb, nb, 2 * niv_out)), alpha)
    return out.cpu().numpy()


def get_starting_sigma(output_path, default_sigma):
    """(Sigma, iteration) of the last sigma_dga_iteration_N.npy in output_path, or (default_sigma, 0)
    when there is none (nonlocal_sde.py:677-696)."""
    from dgamore_amd.self_energy import SelfEnergy

    if output_path == "" or output_path is None or not os.path.exists(output_path):
        return default_sigma, 0
    found = _sigma_files([output_path])
    if not found:
        return default_sigma, 0
    it, path = found[-1]
    return SelfEnergy(np.load(path, allow_pickle=False), config.lattice.nk, True, True, False), it


def read_last_n_sigmas_from_files(n, output_path="./", previous_sc_path="./"):
    """The last n saved self-energies, oldest first, from output_path and previous_sc_path together
    (nonlocal_sde.py:699-716)."""
    dirs = [output_path]
    if previous_sc_path != "" and previous_sc_path is not None and os.path.exists(previous_sc_path):
        dirs.append(previous_sc_path)
    return [np.load(path, allow_pickle=False) for _, path in _sigma_files(dirs)[-n:]]


def _sigma_files(dirs):
    """[(iteration, path)] of every sigma_dga_iteration_N.npy in dirs, by iteration (a directory
    given twice is listed twice, as in the reference)."""
    out = []
    for d in dirs:
        for f in glob.glob(os.path.join(d, "sigma_dga_iteration_*.npy")):
            m = re.search(r"sigma_dga_iteration_(\d+)\.npy$", f)
            if m:
                out.append((int(m.group(1)), f))
    return sorted(out, key=lambda x: x[0])


def apply_mixing_strategy(sigma_new, sigma_old, sigma_dmft, current_iter, history=None):
    """Mixing of the self-energy (nonlocal_sde.py:979-1151). sigma_*: SelfEnergy objects (or arrays
    [nk,nb,nb,2niv]; sigma_dmft also local).

    linear: alpha sigma_new + (1 - alpha) sigma_old on the device, a new object. pulay / anderson
    (after mixing_history_length iterations): the reference's truncated-SVD least squares on the
    core window in FP64 on the host, written into sigma_new, which is returned (the frequencies
    outside the core window stay sigma_new's). The previous self-energies are read with
    read_last_n_sigmas_from_files from config.output.output_path and
    config.self_consistency.previous_sc_path, as in the reference, unless `history` (oldest
    first) is given. A singular history falls back to linear mixing.
    """
    from dgamore_amd.self_energy import SelfEnergy

    n_hist = config.self_consistency.mixing_history_length
    alpha = config.self_consistency.mixing
    strategy = config.self_consistency.mixing_strategy.lower()
    is_obj = isinstance(sigma_new, SelfEnergy)
    new = sigma_new.mat if is_obj else np.asarray(sigma_new)
    old = np.asarray(_mat(sigma_old))
    if strategy in ("pulay", "anderson") and current_iter > n_hist:
        if history is None:
            history = read_last_n_sigmas_from_files(n_hist, config.output.output_path,
                                                    config.self_consistency.previous_sc_path)
        core = _history_mixing(strategy, new, np.asarray(_mat(sigma_dmft)), list(history)[-n_hist:], n_hist, alpha)
        if core is not None:
            niv, nc = new.shape[-1] // 2, config.box.niv_core
            out = new if is_obj else new.copy()
            out[..., niv - nc:niv + nc] = core
            return sigma_new if is_obj else out
    nb, nv = new.shape[-2], new.shape[-1]
    flat = new.reshape(-1, nb, nb, nv)
    old = np.broadcast_to(old.reshape(-1, nb, nb, nv), flat.shape)  # a local sigma_old applies to every k
    mixed = K.sigma_assemble(K.to_dev(flat), None, None, None, K.to_dev(flat[0]), nv // 2, K.to_dev(old), alpha)
    mixed = mixed.cpu().numpy().reshape(new.shape)
    if is_obj:
        return SelfEnergy(mixed, sigma_new.nq, sigma_new.full_niv_range, sigma_new.has_compressed_q_dimension)
    return mixed


def _history_mixing(strategy, new, sigma_dmft, hist, n_hist, alpha):
    """Pulay / Anderson step on the core window (nonlocal_sde.py:990-1143), FP64 host least squares.
    Returns the new core window [..., 2 niv_core], or None when every singular value of the history
    is below the cutoff (the reference then mixes linearly)."""
    niv = new.shape[-1] // 2
    nc = config.box.niv_core
    sl = slice(niv - nc, niv + nc)
    dmft = np.broadcast_to(sigma_dmft.reshape((1,) + sigma_dmft.shape[-3:]), new.shape)
    proposals = [np.asarray(dmft)] + [np.asarray(h).reshape(new.shape) for h in hist]
    results = [np.asarray(h).reshape(new.shape) for h in hist] + [new]
    shape = new[..., sl].shape
    proposals = [p[..., sl].astype(np.complex128).reshape(-1) for p in proposals]
    results = [r[..., sl].astype(np.complex128).reshape(-1) for r in results]
    ri = lambda z: np.concatenate([z.real, z.imag])  # noqa: E731
    f_i = ri(results[-1] - proposals[-1])
    norm_f = np.linalg.norm(f_i)
    if strategy == "pulay":
        r_m = np.column_stack([ri(proposals[-1 - i] - proposals[-2 - i]) for i in range(n_hist)])
        f_m = np.column_stack([ri(results[-1 - i] - results[-2 - i]) for i in range(n_hist)]) - r_m
        u, s, vh = np.linalg.svd(f_m, full_matrices=False)
        mask = s > 1e-5 * (s[0] if s.size else 1.0)
        if not np.any(mask):
            return None
        coeffs = vh[mask].T @ ((u[:, mask].T @ f_i) / s[mask])
        update = alpha * f_i - (r_m + alpha * f_m) @ coeffs
        norm_u = np.linalg.norm(update)
        if norm_f > 0 and norm_u > 10.0 * norm_f:
            update *= 10.0 * norm_f / norm_u
        n = update.size // 2
        return (proposals[-1] + update[:n] + 1j * update[n:]).reshape(shape)
    dx = np.column_stack([ri(proposals[-1 - i] - proposals[-2 - i]) for i in range(n_hist)])
    df = np.column_stack([ri((results[-1 - i] - proposals[-1 - i]) - (results[-2 - i] - proposals[-2 - i]))
                          for i in range(n_hist)])
    u, s, vh = np.linalg.svd(df, full_matrices=False)
    cutoff = 1e-5 * (s[0] if s.size else 1.0)
    mask = s > cutoff
    if not np.any(mask):
        return None
    coeffs = vh[mask].T @ ((s[mask] / (s[mask] ** 2 + cutoff**2)) * (u[:, mask].T @ f_i))
    x_n = proposals[-1]
    xa = ri(x_n) + f_i - (dx + df) @ coeffs
    n = xa.size // 2
    cand = (1 - alpha) * x_n + alpha * (xa[:n] + 1j * xa[n:])
    upd = cand - x_n
    norm_u = np.linalg.norm(upd)
    if norm_f > 0 and norm_u > 3.0 * norm_f:
        cand = x_n + upd * (3.0 * norm_f / norm_u)
    return cand.reshape(shape)
