"""NumPy complex128 restatement of the reference ladder-DGA hot path.

TEST INFRASTRUCTURE ONLY. Imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``; never by the product package.

The reference (Julpe/DGAmore) stores every n-point object as complex64
(``dgamore/n_point_base.py:64-74``). This restatement keeps the reference's
arithmetic structure with that cast removed, so it computes in complex128.
Array layouts are the reference's C-order layouts (``FourPoint.mat``):

* two-point objects   ``[nx, ny, nz, o1, o2, v]``
* 4-point, irr. q     ``[q, o1, o2, o3, o4, w(>=0), v]`` (one fermionic axis)
* local vertex        ``[o1, o2, o3, o4, w(>=0), v, v']``
* compound index      rows ``(o1, o2, v)``, cols ``(o4, o3, v')``
  (``dgamore/four_point.py:184-221``)

Parity pins (see ``PARITY``): each function lists the check in ``tests/``
that pins it against reference-held vectors, known answers, or mathematics.
"""

import numpy as np
import scipy.fft as spfft
import scipy.linalg as spla

# -----------------------------------------------------------------------------
# Parity status of each oracle function: name -> how it is pinned (tests/...).
# -----------------------------------------------------------------------------
PARITY = {
    # helpers (layouts, grids, channel algebra): pinned through the functions that use them
    "vn": "closed form pi/beta (2n+1); with g_full/bubble_fft reproduces the reference-held gchi0_q_rank_0.npy",
    "vn_int": "index helper of vn (same pins)",
    "wn_int": "index helper of the bosonic axis (pinned with bubble_fft)",
    "read_hk": "reads the reference's wannier.hk; with bubble_fft reproduces gchi0_q_rank_0.npy (test_oracle_golden.py)",
    "filter_small": "the reference's 1e-12 filter (n_point_base.py:205-211); part of the bubble_fft pin",
    "cut_niv_g": "frequency window helper of the bubble (pinned with bubble_fft)",
    "u_as_channel": "the reference's expectations of tests/test_interaction.py:93-108 (test_oracle_interaction.py); "
                    "with vrg/chi_phys_shell it reproduces the reference chain vrg_q_* -> gchi_aux_q_*_sum",
    "v_as_channel": "the reference's expectations of tests/test_interaction.py:166-200 (test_oracle_interaction.py)",
    "sum_over_vn": "(1/beta) sum over nu; part of the vrg -> gchi_aux_q_*_sum reference chain",
    "to_full_niw": "F(-w) = conj F(w); reproduces the reference's lambda-correction files (test_oracle_golden.py)",
    "local_uncompound": "layout helper, inverse of compound_local_vertex; pinned through gamma_shell, which returns "
                        "the planted core vertex, and full_vertex_from_gamma, which reproduces the reference's SrVO3 "
                        "f_{dens,magn}_loc.npy (test_oracle_local_sde.py)",
    "apply_lambda": "closed form; with find_lambda reproduces the reference's corrected chi files to 1e-5 "
                    "(the complex64 limit, test_oracle_golden.py)",
    "map_irr_to_fbz": "K_FBZ[k] = K_irr[irrk_inv[k]]; sigma_fft through it equals the direct q-sum form",
    "ladder_matrix": "ladder_sum_inv/_solve pins (v1 = v3 = element-by-element inverse, residual, beta scaling)",
    "ladder_rhs": "the nb^2 indicator columns of four_point.py:661-668; v1 = v3 pin",
    "block_invert_vn0": "textbook nb^2 x nb^2 inverse; part of the chi_phys_shell reference chain",
    "identity_vn1": "helper (block identity), pinned with block_invert_vn1",
    "orb_matrix": "layout helper of auxiliary_chi_q / full_vertex_q (pinned with them)",
    "chi0_inv_compound": "layout helper of the compound matrix (pinned with auxiliary_chi_q)",
    "compound_local_vertex": "layout helper (rows (o1,o2,v), cols (o4,o3,v')), pinned with ladder_matrix",
    "compound_to_ref8": "layout helper of full_vertex_q (pinned with it)",
    "chi0_pp_w0": "one band: |G(k,nu)|^2, real and non-negative on the central 2 niv_pp frequencies; two orbitals: "
                  "an orbital-diagonal complex G puts g_a conj(g_b) exactly at (a, b, b, a) and zero elsewhere, and a "
                  "unitary orbital rotation of a general G transforms the result as W W conj(W) conj(W) "
                  "(test_oracle_eliashberg.py)",
    "pairing_setup": "with pairing_matvec equals the explicit k-space double sum of the linearised gap equation "
                     "(q = k'-k and the flipped q = -k'-k term) on a 4x4, 2-orbital, 2np=4 case without symmetries; "
                     "a missing roll, a flip of the wrong frequency axis, a wrong orbital permutation or the wrong "
                     "sign of the flipped term fails it (test_oracle_eliashberg.py)",
    "dense_solve_reference": "numpy.linalg.solve (the reference's own call), used as a reference inside tests",
    "orbital_expand_fbz": "the kernel of every full-zone q computed directly equals the irreducible one expanded "
                          "with the orbital transforms on an orbital-symmetric 2- and 3-orbital model; the plain "
                          "gather does not (test_oracle_sigma.py)",
    "kgrid_irr_maps": "known irreducible counts 6/45/91/153/325/561 + brute-force orbit check (test_oracle_grid.py)",
    "g_full": "closed form: (iv+mu-ek-sigma) @ G == 1 (test_oracle_golden.py)",
    "bubble_fft": "reference-held gchi0_q_rank_0.npy (test_oracle_golden.py) + brute-force k-sum bubble_direct",
    "bubble_direct": "reference-held gchi0_q_rank_0.npy (test_oracle_golden.py)",
    "block_invert_vn1": "reference-held gchi0_q_inv_rank_0.npy (test_oracle_golden.py)",
    "ladder_sum_inv": "v1 == v3 == dense inverse of the explicitly built matrix, FP64 residual, beta scaling (test_oracle_ladder.py)",
    "ladder_sum_solve": "same as ladder_sum_inv",
    "vrg": "reference-held chain vrg_q -> gchi_aux_q_*_sum (test_oracle_golden.py)",
    "chi_phys_shell": "reference-held gchi_aux_q_{dens,magn}_sum_rank_0.npy (test_oracle_golden.py)",
    "kernel_r": "closed form for nb=1: K = u*(g - g*u*chi - [magn]2/3); nb = 2, 3: the same as products of "
                "nb^2 x nb^2 compound matrices in the (a,b) x (d,c) order of the reference's matmuls "
                "(test_oracle_sigma.py)",
    "dc_kernel": "brute-force loops on a tiny case (test_oracle_sigma.py)",
    "sigma_fft": "equals the direct q-sum sigma_direct (nonlocal_sde.py:404-427) on small grids (test_oracle_sigma.py)",
    "sigma_direct": "brute-force loop over (k,q,w,a,d) on a tiny grid (test_oracle_sigma.py)",
    "find_lambda": "reference known answers lambda=-37.450340, 4.328781, 4.281153 + fixtures (test_oracle_golden.py)",
    "pairing_matvec": "equals the explicit k-space double sum (see pairing_setup; test_oracle_eliashberg.py)",
    "hartree_fock": "closed forms: local U gives k-independent Fock -U n and Hartree 2 U n; cos(q) V on a "
                    "cos(k) occupation gives -(V/2) b cos(k) (test_oracle_selfcons.py)",
    "fit_smom": "closed form on Sigma = m0 + m1/(i v) (test_oracle_selfcons.py)",
    "total_fill": "reference-held mu_history.npy: update_mu on sigma_dga_iteration_1.npy reproduces mu_1 "
                  "(test_oracle_selfcons.py); Fermi-Dirac limit at Sigma = 0",
    "fill_nonlocal": "Sigma = 0 gives the Fermi function of e_k exactly; target filling of the golden run",
    "update_mu": "reference-held mu_history.npy (test_oracle_selfcons.py)",
    "sigma_assemble": "reference-held sigma_dga_iteration_1.npy equals sigma_dmft outside the core window "
                      "(test_oracle_selfcons.py); mixing closed form",
    "auxiliary_chi_q": "A X == 1, row sums equal ladder_sum_solve (test_oracle_vertex.py)",
    "full_vertex_q": "BSE form beta^2 (K - K X K), Woodbury form beta^2 D X K, K = 0 gives 0, second term against "
                     "the reference's einsum contractions (test_oracle_vertex.py)",
    "vertex_pp_w0": "Matsubara-frequency form F_pp(v,v') = -F(w=v-v'; v, -v'), both bosonic signs, on an "
                    "analytic vertex (test_oracle_vertex.py)",
    "pairing_vertices": "reference-held gamma_irrq_{sing,trip}_pp.npy = combination of f_irrq_{dens,magn}_pp.npy "
                        "(test_oracle_vertex.py)",
    "local_bubble": "reference-held gchi0_loc.npy from g_dmft.npy (test_oracle_local_sde.py)",
    "generalized_chi": "closed form beta g2 - 2 beta delta_{w0} G G (dens) on an analytic g2 (test_oracle_local_sde.py)",
    "gamma_shell": "no shell: beta^2 (chi^-1 - chi0^-1); with a shell: an RPA chi built from a constant U_r on the "
                   "full box returns Gamma = U_r exactly, which the core-only chi0 fails (test_oracle_local_sde.py)",
    "full_vertex_from_gamma": "one band, Gamma = U: F = U / (1 + U sum_v chi0(v) / beta^2) in every (v, v'); the "
                              "BSE form beta^2 (chi0^-1 - chi0^-1 chi chi0^-1) of the same Gamma (test_oracle_local_sde.py)",
    "local_sigma": "reference-held siw_dga_local.npy from vrg_{dens,magn}_loc.npy, chi_{dens,magn}_loc.npy and "
                   "g_dmft.npy, up to its Hartree-Fock constant (test_oracle_local_sde.py)",
    "local_hartree": "closed form sum_cd (2 U_abcd - U_adcb) n_dc; one band: U n (test_oracle_local_sde.py)",
}


# -----------------------------------------------------------------------------
# Matsubara grids  (dgamore/matsubara_frequencies.py:21-62)
# -----------------------------------------------------------------------------
def vn_int(niv):
    """Fermionic integers in [-niv, niv) (matsubara_frequencies.py:42-52)."""
    return np.arange(-niv, niv)


def wn_int(niw, positive=False):
    """Bosonic integers in [-niw, niw] or [0, niw] (matsubara_frequencies.py:21-30)."""
    return np.arange(0, niw + 1) if positive else np.arange(-niw, niw + 1)


def vn(niv, beta):
    """nu_n = (2n+1) pi / beta (matsubara_frequencies.py:54-62)."""
    return np.pi / beta * (2 * vn_int(niv) + 1)


# -----------------------------------------------------------------------------
# k-grid symmetry reduction  (dgamore/brillouin_zone.py:128-218, 349-363)
# -----------------------------------------------------------------------------
def _inv_sym(mat, axis):
    """Inversion along one axis of a [0, 2pi) grid (brillouin_zone.py:128-142)."""
    n = mat.shape[axis]
    half, odd = n // 2, n % 2
    src = [slice(None)] * mat.ndim
    dst = [slice(None)] * mat.ndim
    src[axis] = slice(1, half + odd)
    dst[axis] = slice(half + 1, None)
    mat[tuple(dst)] = np.flip(mat[tuple(src)], axis=axis)


def _pair_sym(mat, a, b):
    """Exchange symmetry of two axes (brillouin_zone.py:166-176)."""
    if mat.shape[a] != mat.shape[b]:
        return
    merged = np.minimum(mat, mat.swapaxes(a, b))
    mat[...] = np.minimum(merged, merged.swapaxes(a, b))


SYMMETRY_SETS = {
    # brillouin_zone.py:78-118
    "two_dimensional_square": ("x_inv", "y_inv", "x_y_sym"),
    "three_dimensional_cubic": ("x_inv", "y_inv", "z_inv", "x_y_sym", "x_z_sym", "y_z_sym"),
    "two_dimensional_nematic": ("x_inv", "y_inv"),
    "quasi_two_dimensional_square": ("x_inv", "y_inv", "z_inv", "x_y_sym"),
    "none": (),
}


def kgrid_irr_maps(nk, symmetries="two_dimensional_square"):
    """Irreducible-k maps of ``KGrid`` (brillouin_zone.py:349-363).

    Returns a dict with irrk_ind, irrk_inv, irrk_count (np.unique of the
    symmetrised fbz2irrk index grid), the full and irreducible index lists.
    """
    fbz = np.arange(int(np.prod(nk))).reshape(nk)
    for s in SYMMETRY_SETS[symmetries] if isinstance(symmetries, str) else symmetries:
        if s == "x_inv":
            _inv_sym(fbz, 0)
        elif s == "y_inv":
            _inv_sym(fbz, 1)
        elif s == "z_inv":
            _inv_sym(fbz, 2)
        elif s == "x_y_sym":
            _pair_sym(fbz, 0, 1)
        elif s == "x_z_sym":
            _pair_sym(fbz, 0, 2)
        elif s == "y_z_sym":
            _pair_sym(fbz, 1, 2)
        else:
            raise ValueError(s)
    _, ind, inv, cnt = np.unique(fbz, return_index=True, return_inverse=True, return_counts=True)
    mesh = np.array(np.meshgrid(*[np.arange(n) for n in nk], indexing="ij")).reshape(3, -1)
    return {
        "irrk_ind": ind,
        "irrk_inv": inv.ravel(),
        "irrk_count": cnt,
        "q_list": mesh.T.copy(),
        "irrq_list": mesh[:, ind].T.copy(),
    }


# -----------------------------------------------------------------------------
# One-particle objects
# -----------------------------------------------------------------------------
def g_full(sigma, mu, ek, beta):
    """G(k,v) = [(iv + mu) 1 - e_k - Sigma(k,v)]^-1 (greens_function.py:125-142).

    sigma: [nx,ny,nz,nb,nb,2niv] (or broadcastable [1,1,1,...]); ek: [nx,ny,nz,nb,nb].
    """
    nb = ek.shape[-1]
    niv = sigma.shape[-1] // 2
    eye = np.eye(nb)
    iv = 1j * vn(niv, beta)
    mat = (iv[None, None, :] * eye[..., None] + mu * eye[..., None])[None, None, None] - ek[..., None] - sigma
    inv = np.linalg.inv(mat.transpose(0, 1, 2, 5, 3, 4))
    return inv.transpose(0, 1, 2, 4, 5, 3)


def read_hk(path):
    """wannier90-style H(k) text file -> ([nk, nb, nb] complex, kpoints) (hamiltonian.py:234-318)."""
    with open(path, encoding="utf-8") as f:
        header = f.readline().split("#")[0].split()
        nkp, nbands = int(header[0]), int(header[1])
        data = np.array(f.read().split(), dtype=float)
    data = data.reshape(-1, 3 + 2 * nbands * nbands)[:nkp]
    hk = data[:, 3:].reshape(nkp, nbands, nbands, 2)
    hk = hk[..., 0] + 1j * hk[..., 1]
    # the reference transposes the wannier (row, col) order into its G convention
    return hk.transpose(0, 2, 1), data[:, :3]


def filter_small(x, threshold=1e-12):
    """Zero entries whose |Re| and |Im| are both below threshold (n_point_base.py:205-211)."""
    x = x.copy()
    x[(np.abs(x.real) < threshold) & (np.abs(x.imag) < threshold)] = 0.0
    return x


def cut_niv_g(g, niv_cut):
    """Symmetric fermionic cut of a two-point object (local_n_point.py:134-154)."""
    niv = g.shape[-1] // 2
    if niv_cut > niv:
        return g
    return g[..., niv - niv_cut: niv + niv_cut]


# -----------------------------------------------------------------------------
# Bubble  (dgamore/bubble_gen.py)
# -----------------------------------------------------------------------------
def bubble_fft(giwk, niw, niv, irrk_ind, beta, filter_values=True):
    """chi0^{q w v}_{abcd} = -beta/N sum_k G_ad(k,v) G_cb(k-q,v-w), FFT form (bubble_gen.py:33-69).

    giwk: [nx,ny,nz,nb,nb,2niv_g] with niv_g >= niv+niw. Returns [nq_irr,nb,nb,nb,nb,niw+1,2niv].
    """
    nk = giwk.shape[:3]
    ntot = int(np.prod(nk))
    nb = giwk.shape[3]
    g_k = cut_niv_g(giwk, niv + niw)
    g_r = spfft.fftn(g_k, axes=(0, 1, 2))
    # G(-R): flip + roll(1) (n_point_base.py:553-574), then transpose orbitals
    g_r_rev = np.roll(np.flip(g_r, axis=(0, 1, 2)), 1, axis=(0, 1, 2)).swapaxes(3, 4)
    start, end = niw, niw + 2 * niv
    g_v = g_r[..., start:end]
    out = np.zeros((len(irrk_ind), nb, nb, nb, nb, niw + 1, 2 * niv), dtype=np.complex128)
    for iw in range(niw + 1):
        g_vw = g_r_rev[..., start - iw: end - iw]
        buf = g_v[:, :, :, :, None, None, :, :] * g_vw[:, :, :, None, :, :, None, :]
        out[..., iw, :] = np.fft.ifftn(buf, axes=(0, 1, 2)).reshape((ntot, nb, nb, nb, nb, 2 * niv))[irrk_ind]
    out *= -beta / ntot
    return filter_small(out) if filter_values else out


def bubble_direct(giwk, niw, niv, q_list, beta, filter_values=True):
    """Direct k-sum form of the bubble (bubble_gen.py:147-183). Oracle for bubble_fft."""
    nk = giwk.shape[:3]
    nb = giwk.shape[3]
    g = cut_niv_g(giwk, niv + niw)
    gniv = g.shape[-1] // 2
    g_left = g[..., gniv - niv: gniv + niv]
    out = np.zeros((len(q_list), nb, nb, nb, nb, niw + 1, 2 * niv), dtype=np.complex128)
    idx = [np.arange(n) for n in nk]
    for iq, q in enumerate(q_list):
        g_r = g[(idx[0] - q[0]) % nk[0]][:, (idx[1] - q[1]) % nk[1]][:, :, (idx[2] - q[2]) % nk[2]]
        for iw in range(niw + 1):
            s = gniv - niv - iw
            out[iq, ..., iw, :] = np.einsum("xyzadv,xyzcbv->abcdv", g_left, g_r[..., s: s + 2 * niv])
    out *= -beta / int(np.prod(nk))
    return filter_small(out) if filter_values else out


def chi0_pp_w0(giwk, niv_pp):
    """chi0pp_{abcd}(k,v) = G_ad(k,v) conj(G_cb(k,v)), no -beta (bubble_gen.py:268-279).

    Returns [nk_tot, nb, nb, nb, nb, 2niv_pp].
    """
    g = cut_niv_g(giwk, niv_pp)
    g = g.reshape((-1,) + g.shape[3:])
    return filter_small(g[:, :, None, None, :, :] * np.conj(g.swapaxes(1, 2))[:, None, :, :, None, :])


# -----------------------------------------------------------------------------
# Compound-index helpers  (dgamore/four_point.py:184-293)
# -----------------------------------------------------------------------------
def block_invert_vn1(x):
    """Per-(q,w,v) inverse of the nb^2 x nb^2 block [(o1,o2),(o4,o3)] (four_point.py:601-613).

    x: [q,o1,o2,o3,o4,w,v] -> same shape.
    """
    nq, nb = x.shape[0], x.shape[1]
    nw, nv = x.shape[5], x.shape[6]
    m = x.transpose(0, 5, 6, 1, 2, 4, 3).reshape(nq, nw, nv, nb * nb, nb * nb)
    m = np.linalg.inv(m)
    return m.reshape(nq, nw, nv, nb, nb, nb, nb).transpose(0, 3, 4, 6, 5, 1, 2)


def block_invert_vn0(x):
    """Per-(q,w) inverse of the nb^2 x nb^2 block (four_point.py:207-209, 251-260, 615-618).

    x: [q,o1,o2,o3,o4,w] -> same shape.
    """
    nq, nb, nw = x.shape[0], x.shape[1], x.shape[5]
    m = x.transpose(0, 5, 1, 2, 4, 3).reshape(nq, nw, nb * nb, nb * nb)
    m = np.linalg.inv(m)
    return m.reshape(nq, nw, nb, nb, nb, nb).transpose(0, 2, 3, 5, 4, 1)


def compound_local_vertex(gamma):
    """Local vertex [o1,o2,o3,o4,w,v,v'] -> compound [w, n, n] (local_four_point.py:206-229)."""
    nb, nw, nv = gamma.shape[0], gamma.shape[4], gamma.shape[5]
    n = nb * nb * nv
    return gamma.transpose(4, 0, 1, 5, 3, 2, 6).reshape(nw, n, n)


# -----------------------------------------------------------------------------
# Channel algebra  (dgamore/interaction.py:43-65, 160-182)
# -----------------------------------------------------------------------------
def u_as_channel(u, channel):
    """Local U_abcd in a spin channel (interaction.py:43-65)."""
    ut = np.einsum("abcd->adcb", u)
    return {"dens": 2 * u - ut, "magn": -ut, "sing": u + ut, "trip": u - ut}[channel]


def v_as_channel(v, channel):
    """Non-local V^q_abcd in a spin channel (interaction.py:160-182)."""
    return {"dens": 2 * v, "magn": 0 * v, "sing": v, "trip": v}[channel]


# -----------------------------------------------------------------------------
# Ladder  (dgamore/nonlocal_sde.py:85-156, dgamore/four_point.py:620-684)
# -----------------------------------------------------------------------------
def ladder_matrix(chi0_inv_q, gamma, u_r, v_r_q, beta):
    """A(q,w) = chi0^-1(q,w,v) delta_vv' + (Gamma_r(w) - U_r - V_r^q)/beta^2 in compound form.

    chi0_inv_q: [o1,o2,o3,o4,w,v] (one q); gamma: [o1,o2,o3,o4,w,v,v'];
    u_r: [nb]*4; v_r_q: [nb]*4. Returns [w, n, n] (nonlocal_sde.py:93-95,
    four_point.py:332-430 with local_n_point.py:163-175).
    """
    nb, nw, nv = gamma.shape[0], gamma.shape[4], gamma.shape[5]
    full = gamma / beta**2 - ((u_r + v_r_q) / beta**2)[..., None, None, None]
    eye = np.eye(nv)
    full = full + chi0_inv_q[..., None] * eye
    return compound_local_vertex(full)


def ladder_sum_inv(chi0_inv, gamma, u_r, v_r, beta):
    """sum_v' [A^-1] / beta via explicit inverse, v1 path (four_point.py:620-642).

    chi0_inv: [q,o1..o4,w,v]; gamma: [o1..o4,w,v,v']; u_r: [nb]*4; v_r: [q,nb,nb,nb,nb].
    Returns chi*_sum [q,o1,o2,o3,o4,w,v].
    """
    nq, nb, nw, nv = chi0_inv.shape[0], chi0_inv.shape[1], chi0_inv.shape[5], chi0_inv.shape[6]
    out = np.empty_like(chi0_inv)
    for iq in range(nq):
        a = ladder_matrix(chi0_inv[iq], gamma, u_r, v_r[iq], beta)
        inv = np.linalg.inv(a).reshape(nw, nb, nb, nv, nb, nb, nv)
        out[iq] = inv.transpose(1, 2, 5, 4, 0, 3, 6).sum(axis=-1)
    return out / beta


def ladder_rhs(nb, nv):
    """Indicator right-hand sides B[(o4,o3,v'), (o4,o3)] = 1 (four_point.py:661-668)."""
    n = nb * nb * nv
    idx = np.arange(n)
    rhs = np.zeros((n, nb * nb))
    rhs[idx, (idx // (nb * nv)) * nb + (idx // nv) % nb] = 1.0
    return rhs


def ladder_sum_solve(chi0_inv, gamma, u_r, v_r, beta):
    """Same as ladder_sum_inv through LU solves with nb^2 indicator RHS (four_point.py:644-684)."""
    nq, nb, nw, nv = chi0_inv.shape[0], chi0_inv.shape[1], chi0_inv.shape[5], chi0_inv.shape[6]
    rhs = ladder_rhs(nb, nv)
    out = np.empty_like(chi0_inv)
    for iq in range(nq):
        a = ladder_matrix(chi0_inv[iq], gamma, u_r, v_r[iq], beta)
        x = np.linalg.solve(a, np.broadcast_to(rhs, (nw,) + rhs.shape))
        out[iq] = x.reshape(nw, nb, nb, nv, nb, nb).transpose(1, 2, 5, 4, 0, 3)
    return out / beta


def vrg(chi0_inv, chi_aux_sum, beta):
    """gamma_r = beta chi0^-1 . chi*_sum per (q,w,v) (nonlocal_sde.py:159-164, four_point.py:522-557)."""
    return beta * np.einsum("qabcdwv,qdcefwv->qabefwv", chi0_inv, chi_aux_sum)


def sum_over_vn(x, beta):
    """(1/beta) sum over the last fermionic axis (four_point.py:114-144)."""
    return x.sum(axis=-1) / beta


def chi_phys_shell(chi_aux_sum_v, chi0_full_sum, chi0_core_sum, u_r, v_r):
    """chi_r = ((chi*_sum + chi0_full - chi0_core)^-1 + U_r + V_r^q)^-1 (nonlocal_sde.py:167-182).

    All inputs [q,o1..o4,w] except u_r [nb]*4, v_r [q,nb]*4.
    """
    x = block_invert_vn0(chi_aux_sum_v + chi0_full_sum - chi0_core_sum)
    x = x + (u_r[None] + v_r)[..., None]
    return block_invert_vn0(x)


def identity_vn1(nq, nb, nw, nv):
    """Compound identity delta_ad delta_bc (four_point.py:713-742)."""
    e = np.einsum("ad,bc->abcd", np.eye(nb), np.eye(nb))
    return np.broadcast_to(e[None, ..., None, None], (nq, nb, nb, nb, nb, nw, nv)).copy()


def kernel_r(vrg_r, chi_phys, u_r, v_r, channel):
    """K_r = u_r . (gamma_r - gamma_r . u_r . chi_r - [magn] 2/3 1) (nonlocal_sde.py:201-215)."""
    ur = u_r[None] + v_r  # [q,nb,nb,nb,nb]
    gu = np.einsum("qabijwv,qjief->qabefwv", vrg_r, ur)
    kern = vrg_r - np.einsum("qabcdwv,qdcefw->qabefwv", gu, chi_phys)
    if channel == "magn":
        nq, nb, nw, nv = vrg_r.shape[0], vrg_r.shape[1], vrg_r.shape[5], vrg_r.shape[6]
        kern = kern - 2.0 / 3.0 * identity_vn1(nq, nb, nw, nv)
    return np.einsum("qabij,qjiefwv->qabefwv", ur, kern)


def dc_kernel(f_magn_loc, chi0, u, beta, niv_core):
    """Double-counting kernel K_dc (nonlocal_sde.py:185-198, 817-820 prefactor 2 and abcd->cbad).

    f_magn_loc: [o1..o4,w,v,v'] on the chi0 fermionic box; chi0: [q,o1..o4,w,v].
    Returns +K_dc [q,o1..o4,w,v'] cut to niv_core (the caller subtracts it).
    """
    f_dc = 2.0 * np.einsum("abcd...->cbad...", f_magn_loc)
    up = np.einsum("abcd->adcb", u)
    k0 = np.einsum("abij,qjiefwv->qabefwv", up, chi0) / beta**2
    out = np.einsum("qabcdwv,dcefwvp->qabefwp", k0, f_dc)
    niv = out.shape[-1] // 2
    return out[..., niv - niv_core: niv + niv_core]


# -----------------------------------------------------------------------------
# Full vertex and pairing vertices  (dgamore/nonlocal_sde.py:72-82,
# dgamore/eliashberg_solver.py:44-147, 175-209, 620-629)
# -----------------------------------------------------------------------------
def orb_matrix(x):
    """[..., o1, o2, o3, o4] -> [..., nb^2, nb^2] with rows (o1,o2), cols (o4,o3) (four_point.py:184-221)."""
    nb = x.shape[-1]
    return x.swapaxes(-1, -2).reshape(*x.shape[:-4], nb * nb, nb * nb)


def chi0_inv_compound(chi0_inv_q):
    """chi0^-1 of one q ([o1..o4, w, v]) as the block-diagonal compound matrix [w, n, n]
    (local_n_point.py:163-175 extend_vn_to_diagonal)."""
    nv = chi0_inv_q.shape[-1]
    return compound_local_vertex(chi0_inv_q[..., None] * np.eye(nv))


def auxiliary_chi_q(chi0_inv, gamma, u_r, v_r, beta):
    """chi*_r(q,w) = ((chi0^-1 + Gamma_r/beta^2) - (U_r + V_r^q)/beta^2)^-1 as full inverses
    (nonlocal_sde.py:72-82, four_point.py:590-600). Returns compound [q, w, n, n]."""
    nq = chi0_inv.shape[0]
    return np.stack([np.linalg.inv(ladder_matrix(chi0_inv[iq], gamma, u_r, v_r[iq], beta)) for iq in range(nq)])


def full_vertex_q(chi0_inv, gamma, u_r, v_r, vrg_r, chi_phys, beta):
    """Full ladder vertex F_r(q) (eliashberg_solver.py:78-147 / 175-209), compound [q, w, n, n]:

        beta^2 (chi0^-1 - chi0^-1 chi* chi0^-1) + u @ (vrg * vrg) - u @ chi_r @ u @ (vrg * vrg),

    u = U_r + V_r^q, (vrg * vrg)[(a,b,v),(f,e,v')] = sum_dc vrg_abcd(v) vrg_dcef(v') (four_point.py:442-473).
    vrg_r: [q,o1..o4,w,v] or None (then only the first term); chi_phys: [q,o1..o4,w].
    """
    x = auxiliary_chi_q(chi0_inv, gamma, u_r, v_r, beta)
    nq, nw, n = x.shape[0], x.shape[1], x.shape[2]
    out = np.empty_like(x)
    for iq in range(nq):
        d = chi0_inv_compound(chi0_inv[iq])
        f = beta**2 * (d - d @ x[iq] @ d)
        if vrg_r is not None:
            u = orb_matrix(u_r + v_r[iq])
            s = orb_matrix(np.moveaxis(chi_phys[iq], -1, 0))           # [w, T, T]
            y = u - u @ s @ u
            vm = orb_matrix(np.moveaxis(vrg_r[iq], (-2, -1), (0, 1)))   # [w, v, T, T]
            p = y[:, None] @ vm
            f = f + np.einsum("wvab,wpbc->wavcp", p, vm).reshape(nw, n, n)
        out[iq] = f
    return out


def compound_to_ref8(x, nb):
    """[q, w, n, n] compound -> [q, o1, o2, o3, o4, w, v, v'] (four_point.py:251-293)."""
    nq, nw, n = x.shape[0], x.shape[1], x.shape[2]
    nv = n // (nb * nb)
    return x.reshape(nq, nw, nb, nb, nv, nb, nb, nv).transpose(0, 2, 3, 6, 5, 1, 4, 7)


def vertex_pp_w0(f_ref, niv_pp):
    """ph -> pp notation at w_pp = 0 (eliashberg_solver.py:44-58): cut to niv_pp, full bosonic
    range, flip v', then F_pp[..., v, v'] = -F[..., w = v - v', v, v'] on the integer grids.

    f_ref: [q, o1..o4, w(>=0), v, v'] (niv >= niv_pp). Returns [q, o1..o4, 2niv_pp, 2niv_pp].
    """
    niv = f_ref.shape[-1] // 2
    niw = f_ref.shape[-3] - 1
    cut = f_ref[..., niv - niv_pp: niv + niv_pp, niv - niv_pp: niv + niv_pp]
    full = np.flip(to_full_niw(cut, 2), axis=-1)
    v = vn_int(niv_pp)
    omega = v[:, None] - v[None, :]
    out = np.zeros(f_ref.shape[:5] + (2 * niv_pp, 2 * niv_pp), dtype=f_ref.dtype)
    for idx, w in enumerate(wn_int(niw)):
        mask = omega == w
        out[..., mask] = -full[..., idx, :, :][..., mask]
    return out


def pairing_vertices(f_dens_pp, f_magn_pp):
    """Singlet and triplet pairing vertices (eliashberg_solver.py:620-629)."""
    return 0.5 * f_dens_pp - 1.5 * f_magn_pp, 0.5 * f_dens_pp + 0.5 * f_magn_pp


# -----------------------------------------------------------------------------
# Self-energy  (dgamore/nonlocal_sde.py:404-427, 544-594)
# -----------------------------------------------------------------------------
def to_full_niw(x, n_vn_axes):
    """Half -> full bosonic range: F(-w) = conj(F(w)) with all frequency axes flipped (local_n_point.py:191-225)."""
    w_axis = x.ndim - 1 - n_vn_axes
    freq_axes = tuple(range(w_axis, x.ndim))
    pos = x
    neg = np.conj(np.flip(np.take(x, np.arange(1, x.shape[w_axis]), axis=w_axis), axis=freq_axes))
    return np.concatenate([neg, pos], axis=w_axis)


def map_irr_to_fbz(x_irr, irrk_inv):
    """K_FBZ[k] = K_irr[irrk_inv[k]] (n_point_base.py:462-507 without orbital transform)."""
    return x_irr[irrk_inv]


def orbital_expand_fbz(x_irr, irrk_inv, us, conjs=None):
    """Irreducible -> full zone with the per-k orbital transformation of the reference's auto
    mode (n_point_base.py:462-507 with symmetry_reduction.py:791-903, four orbital axes):
    X_abcd(k) = U_aa' U*_bb' U_cc' U*_dd' [X_a'b'c'd'(rep(k))]^(*conj_k); sigma_k^2 = 1.
    x_irr: [nq, nb, nb, nb, nb, ...]; us: [N, nb, nb]; conjs: [N] bool or None."""
    x = map_irr_to_fbz(x_irr, irrk_inv).copy()
    if conjs is not None:
        x[conjs] = np.conj(x[conjs])
    return np.einsum("kap,kbq,kcr,kds,kpqrs...->kabcd...", us, us.conj(), us, us.conj(), x, optimize=True)


def sigma_fft(kernel_irr, giwk, irrk_inv, beta, niv_core, niw):
    """Sigma(k,v) via the R-space FFT form (nonlocal_sde.py:544-594, 874-879).

    kernel_irr: [q_irr,o1..o4,w(>=0),2niv_core]; giwk: [nx,ny,nz,nb,nb,2niv_g].
    Returns Sigma [nx,ny,nz,nb,nb,2niv_core] in the full fermionic range.
    """
    nk = giwk.shape[:3]
    ntot = int(np.prod(nk))
    nb = giwk.shape[3]
    gniv = giwk.shape[-1] // 2
    k_fbz = map_irr_to_fbz(kernel_irr, irrk_inv)
    k_full = to_full_niw(k_fbz, 1)[..., niv_core:]  # [N, nb^4, 2niw+1, niv_core]
    k_full = k_full.reshape(nk + k_full.shape[1:])
    k_hat = np.conj(spfft.fftn(np.conj(k_full), axes=(0, 1, 2))).reshape((ntot,) + k_full.shape[3:])
    g_r = spfft.fftn(giwk, axes=(0, 1, 2)).reshape(ntot, nb, nb, -1)
    mat = np.zeros((ntot, nb, nb, niv_core), dtype=np.complex128)
    for iw, w in enumerate(wn_int(niw)):
        g_slice = g_r[..., gniv - w: gniv + niv_core - w]
        mat += np.einsum("Radv,Raijdv->Rijv", g_slice, k_hat[..., iw, :])
    mat *= -0.5 / beta / ntot
    sig = spfft.ifftn(mat.reshape(nk + mat.shape[1:]), axes=(0, 1, 2))
    return np.concatenate([np.conj(np.flip(sig, axis=-1)), sig], axis=-1)


def sigma_direct(kernel_irr, giwk, irrk_inv, beta, niv_core, niw):
    """Direct q-sum form Sigma = -1/(2 beta N) sum_{q,w} K(q,w,v) G(k+q,v-w) (nonlocal_sde.py:404-427)."""
    nk = giwk.shape[:3]
    ntot = int(np.prod(nk))
    nb = giwk.shape[3]
    gniv = giwk.shape[-1] // 2
    k_full = to_full_niw(map_irr_to_fbz(kernel_irr, irrk_inv), 1)
    q_list = np.array(np.meshgrid(*[np.arange(n) for n in nk], indexing="ij")).reshape(3, -1).T
    mat = np.zeros(nk + (nb, nb, niv_core), dtype=np.complex128)
    for iq, q in enumerate(q_list):
        shifted = np.roll(giwk, [-i for i in q], axis=(0, 1, 2))
        for iw, w in enumerate(wn_int(niw)):
            g_qk = shifted[..., gniv - w: gniv + niv_core - w]
            mat += np.einsum("aijdv,xyzadv->xyzijv", k_full[iq, ..., iw, niv_core:], g_qk)
    mat *= -0.5 / beta / ntot
    return np.concatenate([np.conj(np.flip(mat, axis=-1)), mat], axis=-1)


# -----------------------------------------------------------------------------
# Hartree-Fock, self-energy assembly and mixing, fillings and the chemical potential
# (dgamore/nonlocal_sde.py:31-69, 884-893, 1145-1151; dgamore/self_energy.py:51-75, 153-168;
#  dgamore/greens_function.py:20-80, 212-240)
# -----------------------------------------------------------------------------
def hartree_fock(u, v_full, occ, occ_k):
    """Hartree [nb,nb] and Fock [nx,ny,nz,nb,nb] terms (nonlocal_sde.py:31-69).

    u: [nb]*4; v_full: [nx,ny,nz,nb,nb,nb,nb] or None; occ: [nb,nb]; occ_k: [nx,ny,nz,nb,nb].
    Fock by the reference's loop: for every q, n(k+q) is a roll of occ_k by -q.
    """
    nk = occ_k.shape[:3]
    nb = u.shape[0]
    v = np.zeros(nk + (nb,) * 4, dtype=complex) if v_full is None else v_full
    hartree = 2 * np.einsum("abcd,dc->ab", u + v[0, 0, 0], occ)
    fock = np.zeros(nk + (nb, nb), dtype=complex)
    for q in np.ndindex(*nk):
        rolled = np.roll(occ_k, [-i for i in q], axis=(0, 1, 2))
        fock += np.einsum("abcd,xyzdc->xyzab", u + v[q], rolled)
    return hartree, -fock / np.prod(nk)


def fit_smom(sigma, beta):
    """First two local moments of Sigma from the top 20% (at least 4) positive frequencies
    (self_energy.py:51-75). sigma: [..k axes.., nb, nb, 2niv] with 3 momentum axes or none."""
    s = sigma if sigma.ndim == 3 else sigma.reshape((-1,) + sigma.shape[-3:]).mean(axis=0)
    niv = s.shape[-1] // 2
    half = s[..., niv:]
    iv = vn(niv, beta)[niv:]
    n_fit = max(int(0.2 * niv), 4)
    fit = half[..., niv - n_fit:]
    return fit.real.mean(axis=-1), (fit.imag * iv[niv - n_fit:][None, None, :]).mean(axis=-1)


def _fermi_matrix(m):
    """Fermi function (1 + exp(M))^-1 of a batch of matrices via their eigenvalues (greens_function.py:38-45)."""
    ev, evec = np.linalg.eig(m)
    rho = np.empty_like(ev)
    pos = ev > 0
    rho[pos] = np.exp(-ev[pos]) / (1 + np.exp(-ev[pos]))
    rho[~pos] = 1 / (1 + np.exp(ev[~pos]))
    return evec @ (rho[..., :, None] * np.eye(m.shape[-1])) @ np.linalg.inv(evec)


def total_fill(mu, ek, sigma_k, beta, smom0):
    """2 Tr occ with the local model Green's function (greens_function.py:20-51).

    ek: [nx,ny,nz,nb,nb]; sigma_k: [nk,nb,nb,2niv]; smom0: [nb,nb].
    """
    nb = sigma_k.shape[-2]
    niv = sigma_k.shape[-1] // 2
    eye = np.eye(nb)
    iv = 1j * vn(niv, beta)
    hloc = ek.mean(axis=(0, 1, 2))
    g_model = np.linalg.inv(iv[:, None, None] * eye + mu * eye - hloc - smom0)
    e = ek.reshape(-1, nb, nb)
    mat = iv[None, :, None, None] * eye + mu * eye - e[:, None] - sigma_k.transpose(0, 3, 1, 2)
    g_loc = np.linalg.inv(mat).mean(axis=0)
    rho = _fermi_matrix(beta * (hloc.real + smom0 - mu * eye))
    occ = rho + (g_loc.real - g_model.real).sum(axis=0) / beta
    return 2.0 * np.trace(occ).real


def update_mu(mu0, target, ek, sigma_k, beta, smom0):
    """Newton (secant) root of total_fill - target with tol 1e-6 (greens_function.py:54-80)."""
    import scipy.optimize as opt

    mu = opt.newton(lambda m: total_fill(m, ek, sigma_k, beta, smom0) - target, mu0, tol=1e-6)
    return float(np.real(mu))


def fill_nonlocal(sigma_k, ek, mu, beta, smom0):
    """(n, occ_mean [nb,nb], occ_k [nx,ny,nz,nb,nb]) with the k-dependent model (greens_function.py:212-240).

    sigma_k: [nk or 1, nb, nb, 2niv]; ek: [nx,ny,nz,nb,nb]; smom0: [nb,nb].
    """
    nk = ek.shape[:3]
    nb = ek.shape[-1]
    niv = sigma_k.shape[-1] // 2
    eye = np.eye(nb)
    iv = 1j * vn(niv, beta)
    e = ek.reshape(-1, nb, nb)
    base = iv[None, :, None, None] * eye + mu * eye - e[:, None]
    g = np.linalg.inv(base - sigma_k.transpose(0, 3, 1, 2))
    g_model = np.linalg.inv(base - smom0)
    rho = _fermi_matrix(beta * (e.real + smom0 - mu * eye))
    occ_k = rho + (g.real - g_model.real).sum(axis=1) / beta
    occ_k.real[np.abs(occ_k) < 1e-12] = 0.0
    occ_mean = occ_k.mean(axis=0)
    occ_mean.real[np.abs(occ_mean) < 1e-12] = 0.0
    return 2.0 * np.trace(occ_mean).real, occ_mean, occ_k.reshape(nk + (nb, nb))


def sigma_assemble(sigma_nl, hartree, fock, delta_sigma, sigma_dmft, niv_out, sigma_old=None, alpha=1.0):
    """sigma + hartree + fock + delta_sigma on the core window, the DMFT self-energy outside it, then
    linear mixing (nonlocal_sde.py:884-893, 1145-1151; self_energy.py:153-168).

    sigma_nl: [nk,nb,nb,2niv_core]; hartree [nb,nb]; fock [nk,nb,nb]; delta_sigma [nb,nb,2niv_core];
    sigma_dmft [nb,nb,2niv_dmft]. Returns [nk,nb,nb,2niv_out].
    """
    nk = sigma_nl.shape[0]
    new = sigma_nl + hartree[None, ..., None] + fock[..., None] + delta_sigma[None]
    nd = sigma_dmft.shape[-1] // 2
    dmft = np.tile(sigma_dmft[..., nd - niv_out: nd + niv_out][None], (nk, 1, 1, 1))
    diff = niv_out - new.shape[-1] // 2
    out = np.concatenate((dmft[..., :diff], new, dmft[..., diff + new.shape[-1]:]), axis=-1)
    return out if sigma_old is None else alpha * out + (1 - alpha) * sigma_old


# -----------------------------------------------------------------------------
# Lambda correction, FP64  (dgamore/lambda_correction.py:13-73)
# -----------------------------------------------------------------------------
def find_lambda(chi_r_full, chi_loc_sum, irrk_count, beta, nk_tot, delta=0.1, eps=1e-7, maxiter=1000):
    """Newton-like root find of sum_q w_q chi/(1+lambda chi)/(beta N) = chi_loc_sum (lambda_correction.py:28-60).

    chi_r_full: [q_irr, 2niw+1] (full bosonic range, single band).
    """
    w0 = chi_r_full.shape[-1] // 2
    lam_start = -np.min(1.0 / chi_r_full[..., w0].real)
    lam = lam_start + delta
    factor = 1.0 / beta / nk_tot
    for _ in range(maxiter):
        chi_lam = 1.0 / (1.0 / chi_r_full + lam)
        f = (irrk_count[:, None] * chi_lam).sum() * factor - chi_loc_sum
        fp = -(irrk_count[:, None] * chi_lam**2).sum() * factor
        lam_new = lam - (f / fp).real
        if abs(f.real) < eps:
            return lam_new
        if lam_new < lam:
            delta /= 2
            lam = lam_start + delta
        else:
            lam = lam_new
    return lam


def apply_lambda(chi, lam):
    """chi_lambda = 1/(1/chi + lambda) (lambda_correction.py:21-25)."""
    return 1.0 / (1.0 / chi + lam)


# -----------------------------------------------------------------------------
# Eliashberg pairing matvec  (dgamore/eliashberg_solver.py:376-428)
# -----------------------------------------------------------------------------
def pairing_setup(gamma_pp_irr, irrk_inv, nk, channel):
    """Gamma_hat = s fft_q Gamma_pp and its flipped copy (eliashberg_solver.py:390-398).

    gamma_pp_irr: [q_irr, o1..o4, v, v'] (pp, w=0). Returns (g_hat, g_hat_flip, sign) with
    arrays [nx,ny,nz,o1..o4,v,v'].
    """
    sign = 1 if channel == "sing" else -1
    g = gamma_pp_irr[irrk_inv].reshape(tuple(nk) + gamma_pp_irr.shape[1:])
    g_hat = sign * spfft.fftn(g, axes=(0, 1, 2))
    flip = np.roll(np.flip(g_hat, axis=(0, 1, 2)), 1, axis=(0, 1, 2))
    flip = np.flip(flip, axis=-1)
    flip = np.einsum("...abcdvp->...adcbvp", flip)
    return g_hat, flip, sign


def pairing_matvec(g_hat, g_hat_flip, chi0pp, gap, sign, beta):
    """One application of the linearised gap operator (eliashberg_solver.py:420-428).

    chi0pp: [nx,ny,nz,o1..o4,v]; gap: [nx,ny,nz,o1,o2,v]. Returns the new gap, same shape.
    """
    nk = gap.shape[:3]
    norm = 0.5 / int(np.prod(nk)) / beta
    gg = np.fft.fftn(np.einsum("xyzabcdv,xyzcdv->xyzabv", chi0pp, gap), axes=(0, 1, 2))
    gg_flip = np.roll(np.flip(gg, axis=(0, 1, 2)), 1, axis=(0, 1, 2))
    new = np.einsum("xyzacbdvp,xyzcdp->xyzabv", g_hat, gg) + sign * np.einsum(
        "xyzacbdvp,xyzcdp->xyzabv", g_hat_flip, gg_flip
    )
    return np.fft.ifftn(norm * new, axes=(0, 1, 2))


def dense_solve_reference(a, b):
    """LAPACK reference solve, used by tests as an independent check."""
    return spla.solve(a, b)


# -----------------------------------------------------------------------------
# Local Schwinger-Dyson equation  (dgamore/local_sde.py:17-217; bubble_gen.py:20-30)
# Local objects: [o1,o2,o3,o4,w(>=0),v(,v')], compound rows (o1,o2,v), cols (o4,o3,v').
# -----------------------------------------------------------------------------
def local_bubble(g_loc, niw, niv, beta, filter_values=True):
    """chi0_abcd(w, v) = -beta G_ad(v) G_cb(v - w), w in [0, niw], v in [-niv, niv) (bubble_gen.py:20-30,
    which forms every w and keeps w >= 0 in its objects). g_loc: [nb, nb, 2 niv_g]."""
    nivg = g_loc.shape[-1] // 2
    v = np.arange(-niv, niv)
    w = np.arange(niw + 1)
    gl = g_loc[:, None, None, :, None, nivg + v]                               # G_ad(v)
    gr = g_loc.swapaxes(0, 1)[None, :, :, None][..., nivg + v[None, :] - w[:, None]]  # G_cb(v - w)
    out = -beta * gl * gr
    return filter_small(out) if filter_values else out


def generalized_chi(g2, g_loc, beta, channel, niv_core):
    """chi_r = beta g2 (w >= 0); density: chi[w = 0] -= 2 beta G_ab(v) G_cd(v') (local_sde.py:17-29)."""
    chi = beta * np.asarray(g2, dtype=complex).copy()
    if channel == "dens":
        nivg = g_loc.shape[-1] // 2
        gs = g_loc[..., nivg - niv_core:nivg + niv_core]
        chi[:, :, :, :, 0] -= 2.0 * beta * gs[:, :, None, None, :, None] * gs[None, None, :, :, None, :]
    return chi


def local_uncompound(m, nb):
    """[w, n, n] -> [o1,o2,o3,o4,w,v,v'] (inverse of compound_local_vertex)."""
    nw, n = m.shape[0], m.shape[1]
    nv = n // (nb * nb)
    return m.reshape(nw, nb, nb, nv, nb, nb, nv).transpose(1, 2, 5, 4, 0, 3, 6)


def _u_full(u_r, nw, nv):
    """A local interaction on every (w, v, v') of a [o1..o4, w, v, v'] object."""
    return np.broadcast_to(np.asarray(u_r, dtype=complex)[..., None, None, None], u_r.shape + (nw, nv, nv))


def gamma_shell(chi_r, chi0_full, u_r, beta, niv_core):
    """Gamma_r = beta^2 (chi_r^-1 - chi_tilde_core^-1) + U_r with chi_tilde = (chi0^-1 + U_r / beta^2)^-1 on the
    full box (U_r on every (v, v')), cut to the core (local_sde.py:40-51).

    chi_r: [o1..o4, w, 2nc, 2nc]; chi0_full: [o1..o4, w, 2nf] (diagonal in v); u_r: [nb]*4."""
    nb, nw, nvf = chi0_full.shape[0], chi0_full.shape[4], chi0_full.shape[5]
    nf, nc = nvf // 2, niv_core
    inv0 = block_invert_vn1(chi0_full[None])[0]
    diag = np.einsum("abcdwv,vp->abcdwvp", inv0, np.eye(nvf))
    chi_t = np.linalg.inv(compound_local_vertex(diag + _u_full(u_r, nw, nvf) / beta**2))
    core = local_uncompound(chi_t, nb)[..., nf - nc:nf + nc, nf - nc:nf + nc]
    chi_t_core_inv = np.linalg.inv(compound_local_vertex(core))
    g = beta**2 * (np.linalg.inv(compound_local_vertex(chi_r)) - chi_t_core_inv)
    return local_uncompound(g, nb) + _u_full(u_r, nw, 2 * nc)


def full_vertex_from_gamma(gamma, chi0_full, u_r, beta):
    """F_r = Gamma_u [1 + chi0 Gamma_u / beta^2]^-1 on the full box, Gamma_u = Gamma in the core and U_r outside
    (local_sde.py:69-76 with local_four_point.py:652-668)."""
    nb, nw, nvf = chi0_full.shape[0], chi0_full.shape[4], chi0_full.shape[5]
    nvc = gamma.shape[-1]
    d = (nvf - nvc) // 2
    gu = np.array(_u_full(u_r, nw, nvf))
    gu[..., d:d + nvc, d:d + nvc] = gamma
    gm = compound_local_vertex(gu)
    c0 = compound_local_vertex(np.einsum("abcdwv,vp->abcdwvp", chi0_full, np.eye(nvf)))
    n = gm.shape[-1]
    return local_uncompound(gm @ np.linalg.inv(np.eye(n) + c0 @ gm / beta**2), nb)


def local_sigma(vrg_d, vrg_m, chi_d, chi_m, g_loc, u, beta, niw, niv):
    """The frequency-dependent part of the local Schwinger-Dyson self-energy (local_sde.py:150-168):
    inner_r = vrg_r - vrg_r . U_r . chi_r, inner = (inner_d - inner_m) / 2 on the full bosonic range,
    Sigma_ij(v) = -(1/beta) sum_w U_kjop inner_ilpo(w, v) G_lk(v - w). Returns [nb, nb, 2 niv]."""
    def inner(vr, ur, cr):
        vu = np.einsum("abcdwv,dcef->abefwv", vr, ur)
        return vr - np.einsum("abefwv,feghw->abghwv", vu, cr)

    full = to_full_niw(0.5 * (inner(vrg_d, u_as_channel(u, "dens"), chi_d) -
                              inner(vrg_m, u_as_channel(u, "magn"), chi_m)), 1)
    nivg = g_loc.shape[-1] // 2
    w = np.arange(-niw, niw + 1)
    v = np.arange(-niv, niv)
    g_wv = g_loc[..., nivg + v[None, :] - w[:, None]]  # GreensFunction.get_g_wv: G_ab(v - w)
    return -1.0 / beta * np.einsum("kjop,ilpowv,lkwv->ijv", u, full, g_wv)


def local_hartree(u, occ):
    """Hartree-Fock constant of the local self-energy: sum_cd U_dens,abcd n_dc (local_sde.py:169)."""
    return np.einsum("abcd,dc->ab", u_as_channel(u, "dens"), occ)
