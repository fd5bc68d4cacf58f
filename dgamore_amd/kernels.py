"""Device-tensor wrappers over the libdgami C ABI.

All arrays are torch tensors on the current CUDA (HIP) device; complex data is
torch.complex128 (the memory layout of numpy.complex128). Layout names (REF2,
QF, QW, QO, CMP, M order) are those of include/dgami.h. torch is used for
device memory, streams and layout permutations only; every arithmetic step of
the hot path runs in libdgami.
"""

import ctypes

import numpy as np
import torch

from dgamore_amd import _lib

C128 = torch.complex128


def _p(t, dtype=C128, name="array"):
    if t is None:
        return None
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a torch tensor on the GPU")
    if t.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("dgamore_amd needs a ROCm GPU (gfx950); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def to_dev(x, dtype=C128):
    """numpy -> contiguous device tensor."""
    return torch.as_tensor(np.ascontiguousarray(x)).to(device=device(), dtype=dtype).contiguous()


# ---------------------------------------------------------------------------
# layout conversions (pure permutations)
# ---------------------------------------------------------------------------
def ref4_to_qf(x):
    """[q,o1,o2,o3,o4,w,v] -> QF [w][(o1,o2),(o4,o3)][v][q]."""
    nq, nb, nw, nv = x.shape[0], x.shape[1], x.shape[5], x.shape[6]
    return x.permute(5, 1, 2, 4, 3, 6, 0).reshape(nw, nb**4, nv, nq).contiguous()


def qf_to_ref4(x, nb):
    """QF [w][nb^4][v][q] -> [q,o1,o2,o3,o4,w,v]."""
    nw, _, nv, nq = x.shape
    return x.reshape(nw, nb, nb, nb, nb, nv, nq).permute(6, 1, 2, 4, 3, 0, 5).contiguous()


def ref4w_to_qw(x):
    """[q,o1,o2,o3,o4,w] -> QW [w][nb^4 M order][q]."""
    nq, nb, nw = x.shape[0], x.shape[1], x.shape[5]
    return x.permute(5, 1, 2, 4, 3, 0).reshape(nw, nb**4, nq).contiguous()


def qw_to_ref4w(x, nb):
    nw, _, nq = x.shape
    return x.reshape(nw, nb, nb, nb, nb, nq).permute(5, 1, 2, 4, 3, 0).contiguous()


def orb_to_m(u):
    """[..., o1,o2,o3,o4] (reference order) -> [..., nb^4] M order ((o1,o2),(o4,o3))."""
    nb = u.shape[-1]
    return u.transpose(-1, -2).reshape(*u.shape[:-4], nb**4).contiguous()


def vertex_to_cmp(g):
    """Local vertex [o1,o2,o3,o4,w,v,v'] -> CMP [w][n][n] (local_four_point.py:206-229)."""
    nb, nw, nv = g.shape[0], g.shape[4], g.shape[5]
    n = nb * nb * nv
    return g.permute(4, 0, 1, 5, 3, 2, 6).reshape(nw, n, n).contiguous()


def qcmp_to_ref8(x, nb):
    """[nw][nq][n][n] per-(q, w) compound matrices -> [q,o1,o2,o3,o4,w,v,v'] (four_point.py:251-293)."""
    nw, nq, n, _ = x.shape
    nv = n // (nb * nb)
    return x.reshape(nw, nq, nb, nb, nv, nb, nb, nv).permute(1, 2, 3, 6, 5, 0, 4, 7).contiguous()


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------
def fft_grid(x, nk, sign, scale=1.0, out=None):
    """Batched FFT over the trailing nx*ny*nz axis of x ([batch][nk])."""
    nx, ny, nz = nk
    out = x if out is None else out
    batch = x.numel() // (nx * ny * nz)
    _lib.call("dgami_fft_grid", _p(x), _p(out), batch, nx, ny, nz, sign, float(scale), _stream())
    return out


def g_full(ek, sigma, mu, beta, niv):
    """G(k, v) [N,nb,nb,2 niv] from e_k [N,nb,nb] and Sigma [1 or N][nb][nb][2 niv_sigma] (centred cut).
    mu: a number, or a float64 device tensor of one element (read by the kernel)."""
    N, nb = ek.shape[0], ek.shape[-1]
    g = torch.empty((N, nb, nb, 2 * niv), dtype=C128, device=ek.device)
    if isinstance(mu, torch.Tensor):
        _lib.call("dgami_g_full_dmu", N, nb, int(niv), float(beta), _p(mu, torch.float64, "mu"), _p(ek), _p(sigma),
                  sigma.shape[0], sigma.shape[-1] // 2, _p(g), _stream())
    else:
        _lib.call("dgami_g_full", N, nb, int(niv), float(beta), float(mu), _p(ek), _p(sigma), sigma.shape[0],
                  sigma.shape[-1] // 2, _p(g), _stream())
    return g


def fit_smom(sigma, beta):
    """Local moments of sigma [N,nb,nb,2 niv] on the device: (mom0, mom1) [nb,nb] complex tensors."""
    N, nb, _, nv = sigma.shape
    m0 = torch.empty((nb, nb), dtype=C128, device=sigma.device)
    m1 = torch.empty_like(m0)
    _lib.call("dgami_fit_smom", N, nb, nv // 2, float(beta), _p(sigma), _p(m0), _p(m1), _stream())
    return m0, m1


def update_mu(mu, target, ek, sigma, hloc, smom0, beta, mixing=1.0, info=None):
    """Secant update of the device chemical potential mu (float64 tensor of one element, in place)
    to the target filling, then mu = mixing mu_new + (1 - mixing) mu_old; info (float64 [3] or
    None): mu_new (nan on failure), status, smom0 asymmetry."""
    N, nb = ek.shape[0], ek.shape[-1]
    _lib.call("dgami_update_mu", N, nb, sigma.shape[-1] // 2, float(beta), float(target), float(mixing), _p(ek),
              _p(sigma), sigma.shape[0], _p(hloc), _p(smom0), _p(mu, torch.float64, "mu"),
              _p(info, torch.float64, "info"), _stream())
    return mu


def occupations(ek, sigma, smom0, mu, beta, occ_k, occ):
    """occ_k [N,nb,nb] and occ [nb,nb] (complex, written in place) of the k-dependent model at the device mu."""
    N, nb = ek.shape[0], ek.shape[-1]
    _lib.call("dgami_occupations", N, nb, sigma.shape[-1] // 2, float(beta), _p(ek), _p(sigma), sigma.shape[0],
              _p(smom0), _p(mu, torch.float64, "mu"), _p(occ_k), _p(occ), _stream())
    return occ_k, occ


def bubble_irrq(giwk, nk, niw, niv, beta, irrk_ind):
    """giwk: REF2 [nx,ny,nz,nb,nb,2niv_g]; irrk_ind: int32 [nq]. Returns chi0 QF [niw+1][nb^4][2niv][nq]."""
    nx, ny, nz = nk
    nb = giwk.shape[3]
    niv_g = giwk.shape[-1] // 2
    nq = irrk_ind.shape[0]
    chi0 = torch.empty((niw + 1, nb**4, 2 * niv, nq), dtype=C128, device=giwk.device)
    _lib.call("dgami_bubble_irrq", nx, ny, nz, nb, niv_g, niw, niv, float(beta), _p(giwk, name="giwk"),
              _p(irrk_ind, torch.int32, "irrk_ind"), nq, _p(chi0), _stream())
    return chi0


def chi0_prepare(chi0, nb, niv_core, beta):
    """chi0 QF (niv_full) -> (chi0_inv QF core, full_sum QW, core_sum QW)."""
    nw, _, nvf, nq = chi0.shape
    inv = torch.empty((nw, nb**4, 2 * niv_core, nq), dtype=C128, device=chi0.device)
    full = torch.empty((nw, nb**4, nq), dtype=C128, device=chi0.device)
    core = torch.empty_like(full)
    _lib.call("dgami_chi0_prepare", nq, nb, nw, nvf // 2, niv_core, float(beta), _p(chi0), _p(inv), _p(full),
              _p(core), _stream())
    return inv, full, core


def capturing():
    """True while the current stream records a HIP graph (no host reads are allowed then)."""
    return torch.cuda.is_current_stream_capturing()


def ladder_solve_sum(chi0_inv, gamma_cmp, u_r_m, v_r_m, beta, precision=_lib.PLAIN_FP64, check_info=True):
    """Batched ladder solves for nchan channels.

    chi0_inv: QF [nw][nb^4][2niv][nq]; gamma_cmp: [nchan][nw][n][n]; u_r_m: [nchan][nb^4];
    v_r_m: [nchan][nq][nb^4] or None. Returns (chi_aux_sum [nchan] x QF, info int32 [nchan][nw][nq]).

    With check_info (outside graph capture), the mixed forms' items whose refinement stalled or
    turned non-finite (negative info) are solved again with the plain FP64 LU, and a singular
    matrix raises numpy.linalg.LinAlgError as in the reference (four_point.py:636-638). Under
    capture, or with check_info=False, info is returned as the kernels left it.
    """
    nw, nb4, nv, nq = chi0_inv.shape
    nb = int(round(nb4 ** 0.25))
    nchan = gamma_cmp.shape[0]
    n = nb * nb * nv
    if gamma_cmp.shape[1:] != (nw, n, n):
        raise ValueError(f"gamma must be [nchan,{nw},{n},{n}], got {tuple(gamma_cmp.shape)}")
    out = torch.empty((nchan, nw, nb4, nv, nq), dtype=C128, device=chi0_inv.device)
    info = torch.zeros((nchan, nw, nq), dtype=torch.int32, device=chi0_inv.device)
    _lib.call("dgami_ladder_solve_sum", nchan, nq, nb, nw, nv // 2, float(beta), _p(chi0_inv), _p(gamma_cmp),
              _p(u_r_m), _p(v_r_m), int(precision), _p(out), _p(info, torch.int32), _stream())
    if check_info and not capturing():
        if precision != _lib.PLAIN_FP64:
            _resolve_fp64(out, info, chi0_inv, gamma_cmp, u_r_m, v_r_m, beta)
        bad = torch.nonzero(info).cpu().numpy()
        if len(bad):
            raise np.linalg.LinAlgError(f"singular ladder matrix at (chan, w, q) = {bad[:8].tolist()}")
    return out, info


def _resolve_fp64(out, info, chi0_inv, gamma_cmp, u_r_m, v_r_m, beta):
    """Solve the items whose mixed-precision refinement failed (info < 0) again with the plain
    FP64 LU, one call per (channel, w) with the failed q-points, and put the results and their
    info in place."""
    bad = torch.nonzero(info < 0).cpu().numpy()
    if not len(bad):
        return
    for c, w in sorted({(int(c), int(w)) for c, w, _ in bad}):
        qs = sorted(int(q) for cc, ww, q in bad if cc == c and ww == w)
        qi = torch.tensor(qs, dtype=torch.long, device=out.device)
        ci = chi0_inv[w:w + 1].index_select(3, qi).contiguous()
        v = None if v_r_m is None else v_r_m[c:c + 1].index_select(1, qi).contiguous()
        o, inf = ladder_solve_sum(ci, gamma_cmp[c:c + 1, w:w + 1].contiguous(), u_r_m[c:c + 1].contiguous(), v, beta,
                                  _lib.PLAIN_FP64, check_info=False)
        out[c, w][..., qi] = o[0, 0]
        info[c, w, qi] = inf[0, 0]


def ladder_solve_sum_multi(models, check_info=True, precision=_lib.PLAIN_FP64):
    """Ladder solves of several models in one call (dgami_ladder_solve_sum_multi; with a mixed
    precision dgami_ladder_solve_sum_multi_prec, whose blocked-LU models take that form and whose
    register-resident models stay FP64).

    models: list of (chi0_inv QF, gamma_cmp [nchan][nw][n][n], u_r_m [nchan][nb^4], v_r_m or None, beta).
    Returns a list of (chi_aux_sum [nchan] x QF, info [nchan][nw][nq]) per model. check_info as in
    ladder_solve_sum (stalled or non-finite mixed items solved again in FP64, singular ones raise).
    """
    outs, table = [], (_lib.LadderModel * max(1, len(models)))()
    for m, (chi0_inv, gamma_cmp, u_r_m, v_r_m, beta) in enumerate(models):
        nw, nb4, nv, nq = chi0_inv.shape
        nb = int(round(nb4 ** 0.25))
        nchan = gamma_cmp.shape[0]
        n = nb * nb * nv
        if tuple(gamma_cmp.shape[1:]) != (nw, n, n):
            raise ValueError(f"model {m}: gamma must be [nchan,{nw},{n},{n}], got {tuple(gamma_cmp.shape)}")
        out = torch.empty((nchan, nw, nb4, nv, nq), dtype=C128, device=chi0_inv.device)
        info = torch.zeros((nchan, nw, nq), dtype=torch.int32, device=chi0_inv.device)
        table[m] = _lib.LadderModel(nchan, nq, nb, nw, nv // 2, float(beta), _p(chi0_inv), _p(gamma_cmp), _p(u_r_m),
                                    _p(v_r_m), _p(out), _p(info, torch.int32))
        outs.append((out, info))
    if precision == _lib.PLAIN_FP64:
        _lib.call("dgami_ladder_solve_sum_multi", len(models), ctypes.addressof(table), _stream())
    else:
        _lib.call("dgami_ladder_solve_sum_multi_prec", len(models), ctypes.addressof(table), int(precision), _stream())
    if check_info and not capturing():
        for (out, info), (chi0_inv, gamma_cmp, u_r_m, v_r_m, beta) in zip(outs, models):
            if precision != _lib.PLAIN_FP64:
                _resolve_fp64(out, info, chi0_inv, gamma_cmp, u_r_m, v_r_m, beta)
            bad = torch.nonzero(info).cpu().numpy()
            if len(bad):
                raise np.linalg.LinAlgError(f"singular ladder matrix at (chan, w, q) = {bad[:8].tolist()}")
    return outs


def _raise_singular(info, what):
    bad = torch.nonzero(info).cpu().numpy()
    if len(bad):
        raise np.linalg.LinAlgError(f"singular {what} at (w, q) = {bad[:8].tolist()}")


def ladder_invert(chi0_inv, gamma_cmp, u_r_m, v_r_m, beta, check_info=True):
    """Full inverse X = A^-1 of one channel's ladder matrices.

    chi0_inv: QF [nw][nb^4][2niv][nq]; gamma_cmp: CMP [nw][n][n]; u_r_m: [nb^4]; v_r_m: [nq][nb^4] or None.
    Returns (X [nw][nq][n][n], info int32 [nw][nq]).
    """
    nw, nb4, nv, nq = chi0_inv.shape
    nb = int(round(nb4 ** 0.25))
    n = nb * nb * nv
    if tuple(gamma_cmp.shape) != (nw, n, n):
        raise ValueError(f"gamma must be [{nw},{n},{n}], got {tuple(gamma_cmp.shape)}")
    out = torch.empty((nw, nq, n, n), dtype=C128, device=chi0_inv.device)
    info = torch.zeros((nw, nq), dtype=torch.int32, device=chi0_inv.device)
    _lib.call("dgami_ladder_invert", nq, nb, nw, nv // 2, float(beta), _p(chi0_inv), _p(gamma_cmp), _p(u_r_m),
              _p(v_r_m), _p(out), _p(info, torch.int32), _stream())
    if check_info:
        _raise_singular(info, "ladder matrix")
    return out, info


def invert_dense(a, check_info=True):
    """Inverses of dense complex128 matrices a [..., n, n] (blocked LU, dgami_invert_dense)."""
    n = a.shape[-1]
    flat = a.reshape(-1, n, n).contiguous()
    out = torch.empty_like(flat)
    info = torch.zeros(flat.shape[0], dtype=torch.int32, device=a.device)
    _lib.call("dgami_invert_dense", flat.shape[0], n, _p(flat), _p(out), _p(info, torch.int32), _stream())
    if check_info:
        _raise_singular(info, "matrix")
    return out.reshape(a.shape)


def full_vertex(chi0_inv, gamma_cmp, u_r_m, v_r_m, vrg, chi_phys, beta, check_info=True, precision=0, window=None):
    """Full vertex F_r [nw][nq][n][n] of one channel (vrg QF, chi_phys QW; both None: first term only).

    precision 0: complex128 LU; 1 / 2: complex64 LU with FP32 / split-BF16 matrix-core updates and
    no refinement (relative error of F about 1e-6 of its largest element). window (v_lo, v_hi): only
    the rows and columns with fermionic index in [v_lo, v_hi) are formed; the other entries of the
    returned tensor are undefined."""
    nw, nb4, nv, nq = chi0_inv.shape
    nb = int(round(nb4 ** 0.25))
    n = nb * nb * nv
    if tuple(gamma_cmp.shape) != (nw, n, n):
        raise ValueError(f"gamma must be [{nw},{n},{n}], got {tuple(gamma_cmp.shape)}")
    if vrg is not None and tuple(vrg.shape) != tuple(chi0_inv.shape):
        raise ValueError("vrg must have the shape of chi0_inv")
    if chi_phys is not None and tuple(chi_phys.shape) != (nw, nb4, nq):
        raise ValueError(f"chi_phys must be [{nw},{nb4},{nq}]")
    out = torch.empty((nw, nq, n, n), dtype=C128, device=chi0_inv.device)
    info = torch.zeros((nw, nq), dtype=torch.int32, device=chi0_inv.device)
    v_lo, v_hi = (0, 0) if window is None else (int(window[0]), int(window[1]))
    if precision == 0 and window is None:
        _lib.call("dgami_full_vertex", nq, nb, nw, nv // 2, float(beta), _p(chi0_inv), _p(gamma_cmp), _p(u_r_m),
                  _p(v_r_m), _p(vrg), _p(chi_phys), _p(out), _p(info, torch.int32), _stream())
    else:
        _lib.call("dgami_full_vertex_prec", nq, nb, nw, nv // 2, float(beta), int(precision), v_lo, v_hi,
                  _p(chi0_inv), _p(gamma_cmp), _p(u_r_m), _p(v_r_m), _p(vrg), _p(chi_phys), _p(out),
                  _p(info, torch.int32), _stream())
    if check_info:
        _raise_singular(info, "ladder matrix")
    return out, info


def pairing_vertex_pp(f_dens, f_magn, nb, niv_pp, want_trip=True):
    """Singlet (and triplet) pairing vertices [nq,nb,nb,nb,nb,2np,2np] from F_dens, F_magn [nw][nq][n][n].

    f_magn None: (F_pp of f_dens, None), the transform of one channel.
    """
    nw, nq, n, _ = f_dens.shape
    if f_magn is None:
        want_trip = False
    elif tuple(f_magn.shape) != tuple(f_dens.shape):
        raise ValueError("f_dens and f_magn must have the same shape")
    niv = n // (2 * nb * nb)
    shape = (nq, nb, nb, nb, nb, 2 * niv_pp, 2 * niv_pp)
    gs = torch.empty(shape, dtype=C128, device=f_dens.device)
    gt = torch.empty(shape, dtype=C128, device=f_dens.device) if want_trip else None
    _lib.call("dgami_pairing_vertex_pp", nq, nb, nw, niv, int(niv_pp), _p(f_dens), _p(f_magn), _p(gs), _p(gt),
              _stream())
    return gs, gt


def pairing_vertex_pp_range(f_dens, f_magn, nb, niv_pp, w0, gs, gt=None):
    """Write into gs (and gt) the pairing-vertex entries whose |w| lies in [w0, w0 + nw) from the full
    vertices F_dens, F_magn [nw][nq][n][n] of those bosonic frequencies; other entries are untouched."""
    nw, nq, n, _ = f_dens.shape
    if f_magn is not None and tuple(f_magn.shape) != tuple(f_dens.shape):
        raise ValueError("f_dens and f_magn must have the same shape")
    niv = n // (2 * nb * nb)
    _lib.call("dgami_pairing_vertex_pp_range", nq, nb, nw, int(w0), niv, int(niv_pp), _p(f_dens), _p(f_magn), _p(gs),
              _p(gt), _stream())
    return gs, gt


def generalized_chi(g2_cmp, g, nb, beta, dens):
    """Local chi_r CMP [nw][n][n] from g2 CMP (w >= 0); dens subtracts the disconnected part with
    G [nb][nb][2 niv_g]."""
    nw, n, _ = g2_cmp.shape
    niv = n // (2 * nb * nb)
    out = torch.empty_like(g2_cmp)
    _lib.call("dgami_generalized_chi", nb, nw, niv, 0 if g is None else g.shape[-1] // 2, float(beta),
              1 if dens else 0, _p(g2_cmp), _p(g), _p(out), _stream())
    return out


def local_gamma(chi_inv, chit_inv, u_r_m, nb, beta, niv_full=None):
    """Gamma = beta^2 (chi_inv - chit_inv) + U_r on the core box (CMP [nw][nc][nc]); with niv_full
    also the vertex padded with U_r to the full box. chit_inv None: chi_inv is Gamma itself and
    only the padded vertex is formed. Returns (gamma or None, gamma_pad or None)."""
    nw, nc, _ = chi_inv.shape
    if chit_inv is not None and tuple(chit_inv.shape) != tuple(chi_inv.shape):
        raise ValueError("chi_inv and chit_inv must have the same shape")
    if chit_inv is None and niv_full is None:
        raise ValueError("padding needs niv_full")
    niv_core = nc // (2 * nb * nb)
    gamma = torch.empty_like(chi_inv) if chit_inv is not None else None
    pad = None
    if niv_full is not None:
        nf = nb * nb * 2 * niv_full
        pad = torch.empty((nw, nf, nf), dtype=C128, device=chi_inv.device)
    _lib.call("dgami_local_gamma", nb, nw, niv_core, int(niv_full or niv_core), float(beta), _p(chi_inv),
              _p(chit_inv), _p(u_r_m), _p(gamma), _p(pad), _stream())
    return gamma, pad


def local_sigma(vrg_d, vrg_m, chi_d, chi_m, ud_m, um_m, u_ref, g, beta, occ=None):
    """Local Schwinger-Dyson self-energy [nb][nb][2 niv] (vrg QF and chi QW with nq = 1; the
    Hartree-Fock constant is added when occ [nb][nb] is given)."""
    nw, nb4, nv, nq = vrg_d.shape
    nb = int(round(nb4 ** 0.25))
    if nq != 1 or tuple(chi_d.shape) != (nw, nb4, 1):
        raise ValueError("local_sigma takes local (nq = 1) vertices")
    sigma = torch.empty((nb, nb, nv), dtype=C128, device=vrg_d.device)
    _lib.call("dgami_local_sigma", nb, nw - 1, nv // 2, g.shape[-1] // 2, float(beta), _p(vrg_d), _p(vrg_m),
              _p(chi_d), _p(chi_m), _p(ud_m), _p(um_m), _p(u_ref), _p(g), _p(occ), _p(sigma), _stream())
    return sigma


def vrg_chi_phys(chi0_inv, aux, full, core, u_r_m, v_r_m, nb, beta, want_vrg=True):
    """-> (vrg QF or None, chi_phys QW)."""
    nw, _, nv, nq = aux.shape
    vrg = torch.empty_like(aux) if want_vrg else None
    chi = torch.empty((nw, nb**4, nq), dtype=C128, device=aux.device)
    info = torch.zeros((nw, nq), dtype=torch.int32, device=aux.device)
    _lib.call("dgami_vrg_chi_phys", nq, nb, nw, nv // 2, float(beta), _p(chi0_inv), _p(aux), _p(full), _p(core),
              _p(u_r_m), _p(v_r_m), _p(vrg), _p(chi), _p(info, torch.int32), _stream())
    return vrg, chi


def kernel_r(vrg, chi_phys, u_r_m, v_r_m, nb, channel, weight, kernel=None, accumulate=False):
    """kernel (QF) = [kernel +] weight * K_r; channel 'dens' or 'magn'."""
    nw, _, nv, nq = vrg.shape
    if kernel is None:
        kernel = torch.empty_like(vrg)
        accumulate = False
    ch = {"dens": 0, "magn": 1}[channel]
    _lib.call("dgami_kernel_r", nq, nb, nw, nv // 2, ch, float(weight), _p(vrg), _p(chi_phys), _p(u_r_m),
              _p(v_r_m), int(bool(accumulate)), _p(kernel), _stream())
    return kernel


def dc_kernel(chi0, u_ref, f_magn_loc, nb, niv_core, beta):
    """-K_dc in QF core layout. chi0 QF (niv_full); u_ref [nb]*4; f_magn_loc [nb]*4+[nw,2nvf,2nvf]."""
    nw, _, nvf, nq = chi0.shape
    out = torch.empty((nw, nb**4, 2 * niv_core, nq), dtype=C128, device=chi0.device)
    _lib.call("dgami_dc_kernel", nq, nb, nw, nvf // 2, niv_core, float(beta), _p(chi0), _p(u_ref),
              _p(f_magn_loc), _p(out), _stream())
    return out


def sigma_from_kernel(kernel_irr, irrk_inv, giwk, nk, niw, beta):
    """Sigma(k, v) REF2 [nx,ny,nz,nb,nb,2niv_core] from the irreducible kernel (QF)."""
    nx, ny, nz = nk
    nb = giwk.shape[3]
    niv_g = giwk.shape[-1] // 2
    nv = kernel_irr.shape[2]
    nq = kernel_irr.shape[3]
    sigma = torch.empty((nx, ny, nz, nb, nb, nv), dtype=C128, device=giwk.device)
    _lib.call("dgami_sigma_from_kernel", nx, ny, nz, nb, niw, nv // 2, niv_g, float(beta), _p(kernel_irr),
              _p(irrk_inv, torch.int32, "irrk_inv"), nq, _p(giwk), _p(sigma), _stream())
    return sigma


def orbital_expand_tables(irrk_inv, us, conjs=None):
    """Host tables of the transformed irreducible-to-full-zone map: the distinct (irreducible q,
    U_k, conj_k) triples of the grid become kernel entries j; returns (src [next], u [next,nb,nb],
    conj [next] int32, irrk_inv_ext [N]) as numpy arrays, the untransformed entries first."""
    irrk_inv = np.asarray(irrk_inv)
    us = np.asarray(us, dtype=np.complex128)
    conjs = np.zeros(len(irrk_inv), dtype=bool) if conjs is None else np.asarray(conjs, dtype=bool)
    keys, src, ul, cl = {}, [], [], []
    inv_ext = np.empty(len(irrk_inv), dtype=np.int32)
    order = np.argsort(np.abs(us - np.eye(us.shape[-1])).reshape(len(us), -1).sum(1) + conjs, kind="stable")
    for k in order:  # identity transforms first
        key = (int(irrk_inv[k]), us[k].tobytes(), bool(conjs[k]))
        j = keys.get(key)
        if j is None:
            j = keys[key] = len(src)
            src.append(int(irrk_inv[k]))
            ul.append(us[k])
            cl.append(int(conjs[k]))
        inv_ext[k] = j
    return (np.array(src, dtype=np.int32), np.array(ul), np.array(cl, dtype=np.int32), inv_ext)


def kernel_orbital_expand(kernel_irr, src, u, conj=None):
    """QF kernel on the irreducible q -> QF kernel on the transformed entries (dgami_kernel_orbital_expand).
    src: int32 [next], u: complex128 [next, nb, nb], conj: int32 [next] or None (device tensors)."""
    nw, nb4, nv, nq = kernel_irr.shape
    nb = int(round(nb4 ** 0.25))
    nxt = int(src.shape[0])
    if tuple(u.shape) != (nxt, nb, nb):
        raise ValueError(f"u must be [{nxt},{nb},{nb}], got {tuple(u.shape)}")
    out = torch.empty((nw, nb4, nv, nxt), dtype=C128, device=kernel_irr.device)
    _lib.call("dgami_kernel_orbital_expand", nq, nxt, nb, nw, nv, _p(kernel_irr), _p(src, torch.int32), _p(u),
              _p(conj, torch.int32), _p(out), _stream())
    return out


def hartree_fock(u_ref, v_full, occ, occ_k, nk):
    """(hartree [nb,nb], fock [N,nb,nb]); u_ref [nb]*4, v_full [N,nb,nb,nb,nb] or None (reference order),
    occ [nb,nb], occ_k [N,nb,nb] (complex)."""
    nx, ny, nz = nk
    nb = u_ref.shape[-1]
    N = nx * ny * nz
    hart = torch.empty((nb, nb), dtype=C128, device=u_ref.device)
    fock = torch.empty((N, nb, nb), dtype=C128, device=u_ref.device)
    _lib.call("dgami_hartree_fock", nx, ny, nz, nb, _p(u_ref), _p(v_full), _p(occ), _p(occ_k), _p(hart), _p(fock),
              _stream())
    return hart, fock


def sigma_assemble(sigma_nl, hartree, fock, delta_sigma, sigma_dmft, niv_out, sigma_old=None, alpha=1.0):
    """New self-energy [N,nb,nb,2 niv_out]: core window sigma_nl + H + F + delta_sigma, DMFT outside, then
    alpha new + (1 - alpha) old. sigma_nl [N,nb,nb,2 niv_core]; sigma_dmft [nb,nb,2 niv_dmft]."""
    N, nb, _, nvc = sigma_nl.shape
    out = torch.empty((N, nb, nb, 2 * niv_out), dtype=C128, device=sigma_nl.device)
    _lib.call("dgami_sigma_assemble", N, nb, nvc // 2, int(niv_out), _p(sigma_nl), _p(hartree), _p(fock),
              _p(delta_sigma), _p(sigma_dmft), sigma_dmft.shape[-1] // 2, _p(sigma_old), float(alpha), _p(out),
              _stream())
    return out


def fill_sums(ek, sigma, mu, beta, smom0, hloc=None, want_mean=True):
    """(occ_k [N,nb,nb] float64, occ_mean [nb,nb] or None): (1/beta) sum_v Re[G - G_model]."""
    N, nb = ek.shape[0], ek.shape[-1]
    occ_k = torch.empty((N, nb, nb), dtype=torch.float64, device=ek.device)
    occ_mean = torch.empty((nb, nb), dtype=torch.float64, device=ek.device) if want_mean else None
    _lib.call("dgami_fill_sums", N, nb, sigma.shape[-1] // 2, float(beta), float(mu), _p(ek), _p(sigma),
              sigma.shape[0], _p(smom0), _p(hloc), _p(occ_k, torch.float64), _p(occ_mean, torch.float64), _stream())
    return occ_k, occ_mean


def kmean(x):
    """Mean over the leading (momentum) axis of a complex tensor."""
    out = torch.empty(x.shape[1:], dtype=C128, device=x.device)
    _lib.call("dgami_kmean", x.shape[0], int(np.prod(x.shape[1:])), _p(x), _p(out), _stream())
    return out


def chi0_pp(giwk, niv_pp):
    """chi0pp [N][nb^4 (abcd)][2 niv_pp] from G(k) [..., nb, nb, 2 niv_g] (REF2 or [N,nb,nb,2niv_g])."""
    nb, nvg = giwk.shape[-2], giwk.shape[-1]
    g = giwk.reshape(-1, nb, nb, nvg).contiguous()
    out = torch.empty((g.shape[0], nb**4, 2 * niv_pp), dtype=C128, device=giwk.device)
    _lib.call("dgami_chi0_pp", g.shape[0], nb, nvg // 2, int(niv_pp), _p(g), _p(out), _stream())
    return out


def pairing_setup(gamma_pp_irr, irrk_inv, nk, channel, storage=C128, orbital_transforms=None):
    """gamma_pp_irr [nq,nb,nb,nb,nb,np2,np2] -> (g_hat, g_hat_flip) [N][n][n] in complex128 or complex64.
    orbital_transforms: per-k (us [N,nb,nb], conjs [N]) of the irreducible-zone map (the reference's
    auto symmetry mode) or None for the plain gather; with them the vertex is first expanded to
    its distinct transformed entries (dgami_kernel_orbital_expand)."""
    if orbital_transforms is not None:
        nq0, nb0, np0 = gamma_pp_irr.shape[0], gamma_pp_irr.shape[1], gamma_pp_irr.shape[-1]
        src, u, cj, inv_ext = orbital_expand_tables(np.asarray(irrk_inv.cpu() if hasattr(irrk_inv, "cpu")
                                                               else irrk_inv), *orbital_transforms)
        qf = ref4_to_qf(gamma_pp_irr.reshape((nq0,) + (nb0,) * 4 + (1, np0 * np0)))
        ext = kernel_orbital_expand(qf, to_dev(src, torch.int32), to_dev(u), to_dev(cj, torch.int32))
        gamma_pp_irr = qf_to_ref4(ext, nb0).reshape((len(src),) + (nb0,) * 4 + (np0, np0)).contiguous()
        irrk_inv = to_dev(inv_ext, torch.int32)
    nx, ny, nz = nk
    nq, nb, np2 = gamma_pp_irr.shape[0], gamma_pp_irr.shape[1], gamma_pp_irr.shape[-1]
    n = nb * nb * np2
    N = nx * ny * nz
    g_hat = torch.empty((N, n, n), dtype=storage, device=gamma_pp_irr.device)
    g_flip = torch.empty_like(g_hat)
    ch = {"sing": 0, "trip": 1}[channel]
    name = "dgami_pairing_setup" if storage == C128 else "dgami_pairing_setup_c64"
    _lib.call(name, nx, ny, nz, nb, np2, ch, _p(gamma_pp_irr), _p(irrk_inv, torch.int32), nq,
              _p(g_hat, storage), _p(g_flip, storage), _stream())
    return g_hat, g_flip


def pairing_matvec(g_hat, g_flip, chi0pp, gap, nk, channel, beta):
    """gap [nvec][nx,ny,nz,nb,nb,np2] (or without the vec axis) -> same shape."""
    nx, ny, nz = nk
    squeeze = gap.dim() == 6
    g = gap.unsqueeze(0) if squeeze else gap
    nvec, nb, np2 = g.shape[0], g.shape[4], g.shape[-1]
    g = g.contiguous()
    out = torch.empty_like(g)
    ch = {"sing": 0, "trip": 1}[channel]
    name = "dgami_pairing_matvec" if g_hat.dtype == C128 else "dgami_pairing_matvec_c64"
    _lib.call(name, nx, ny, nz, nb, np2, ch, float(beta), _p(g_hat, g_hat.dtype), _p(g_flip, g_hat.dtype),
              _p(chi0pp), nvec, _p(g), _p(out), _stream())
    return out[0] if squeeze else out
