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
    """Sigma(k, v) in the full fermionic range from the irreducible kernel.

    Fuses calculate_sigma_from_kernel_fft_* (nonlocal_sde.py:544-674), the irr->FBZ map
    (mpi_utils.py:123-131) and the caller's ifft + to_full_niv_range (nonlocal_sde.py:874-879):
    the return value is the k-space self-energy array [nx,ny,nz,nb,nb,2niv_core].
    """
    import torch

    beta = config.sys.beta
    grid = config.lattice.q_grid
    ker = kernel.to_half_niw_range().compress_q_dimension()
    if ker.mat.shape[0] != len(grid.irrk_ind):
        raise ValueError("kernel must be given on the irreducible q-points of config.lattice.q_grid")
    g = _mat(giwk)
    niw = ker.mat.shape[5] - 1
    sig = K.sigma_from_kernel(K.ref4_to_qf(K.to_dev(ker.mat)), K.to_dev(np.asarray(grid.irrk_inv), torch.int32),
                              K.to_dev(g), tuple(grid.nk), niw, beta)
    return sig.cpu().numpy()


def calculate_sigma_from_kernel_auto(mpi_distributor, kernel, giwk, my_full_q_list=None):
    """Direct q-sum form of the reference (nonlocal_sde.py:404-541); same values by the FFT route."""
    return calculate_sigma_from_kernel_fft_auto(mpi_distributor, kernel, giwk)
