"""Linearised Eliashberg equation with the pairing matvec on the GPU (eliashberg_solver.py:376-471),
and the full ladder vertex and pairing vertices it starts from (eliashberg_solver.py:44-147, 620-629).

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
from dgamore_amd.four_point import FourPoint, FrequencyNotation, SpinChannel
from dgamore_amd.nonlocal_sde import _ladder_operands


def _ph_compound(f_q_r):
    """FourPoint [q, o1..o4, w(>=0), v, v'] -> device [nw][nq][n][n] compound (four_point.py:184-221)."""
    mat = f_q_r.to_half_niw_range().compress_q_dimension().mat
    nq, nb, nw, nv = mat.shape[0], mat.shape[1], mat.shape[5], mat.shape[6]
    n = nb * nb * nv
    d = K.to_dev(mat)
    return d.permute(5, 0, 1, 2, 6, 4, 3, 7).reshape(nw, nq, n, n).contiguous(), nb


def create_full_vertex_q_r(u_loc, v_nonloc, gamma_r, niv_pp, gchi0_q_inv, vrg_q_r, gchi_aux_q_r_sum,
                           save_fq=None, construct_fq_cheap=None):
    """Full ladder vertex of one channel (eliashberg_solver.py:78-147, and the per-q _v2 form :175-209):

        F_r = beta^2 (chi0^-1 - chi0^-1 chi*_r chi0^-1) + u (vrg * vrg) - u chi_r u (vrg * vrg),  u = U_r + V_r^q

    The reference reads chi0^-1, vrg and chi_r (its gchi_aux_q_r_sum file, the shell-corrected
    physical susceptibility) from per-rank files; here they are passed in. With save_fq the
    particle-hole vertex [q, o1..o4, w, v, v'] is returned, otherwise its particle-particle
    transform at w = 0 [q, o1..o4, 2 niv_pp, 2 niv_pp] (transform_vertex_q_frequencies_w0).
    construct_fq_cheap cuts every operand to niv_pp fermionic frequencies first.
    """
    save_fq = config.eliashberg.save_fq if save_fq is None else save_fq
    cheap = config.eliashberg.construct_fq_cheap if construct_fq_cheap is None else construct_fq_cheap
    if cheap:
        gamma_r = gamma_r.cut_niv(niv_pp)
        gchi0_q_inv = gchi0_q_inv.cut_niv(niv_pp)
        vrg_q_r = vrg_q_r.cut_niv(niv_pp)
    chi0_inv, gam, ur, vr, inv = _ladder_operands(gamma_r, gchi0_q_inv, u_loc, v_nonloc)
    nb = inv.n_bands
    vrg = K.ref4_to_qf(K.to_dev(vrg_q_r.to_half_niw_range().compress_q_dimension().mat))
    chi = K.ref4w_to_qw(K.to_dev(gchi_aux_q_r_sum.to_half_niw_range().compress_q_dimension().mat))
    f, _ = K.full_vertex(chi0_inv, gam, ur, vr, vrg, chi, config.sys.beta)
    if save_fq:
        return FourPoint(K.qcmp_to_ref8(f, nb).cpu().numpy(), gamma_r.channel, inv.nq, 1, 2, False,
                         has_compressed_q_dimension=True)
    pp, _ = K.pairing_vertex_pp(f, None, nb, niv_pp)
    return FourPoint(pp.cpu().numpy(), gamma_r.channel, inv.nq, 0, 2, False, True, True, FrequencyNotation.PP)


def transform_vertex_q_frequencies_w0(f_q_r, niv_pp):
    """Particle-hole -> particle-particle notation at w = 0 (eliashberg_solver.py:44-58, 69-75):
    F_pp[..., v, v'] = -F[..., w = v - v', v, -v' (flipped index)], negative w through F(-w) = conj F(w)
    with both fermionic axes flipped."""
    f, nb = _ph_compound(f_q_r)
    pp, _ = K.pairing_vertex_pp(f, None, nb, niv_pp)
    return FourPoint(pp.cpu().numpy(), f_q_r.channel, f_q_r.nq, 0, 2, False, True, True, FrequencyNotation.PP)


def create_pairing_vertices_pp(f_dens, f_magn, niv_pp):
    """Singlet and triplet pairing vertices from the particle-hole full vertices of the density and
    magnetic channels: 0.5 F_dens,pp - 1.5 F_magn,pp and 0.5 F_dens,pp + 0.5 F_magn,pp
    (the combination in solve, eliashberg_solver.py:620-629, fused with the w = 0 transform)."""
    fd, nb = _ph_compound(f_dens)
    fm, _ = _ph_compound(f_magn)
    gs, gt = K.pairing_vertex_pp(fd, fm, nb, niv_pp)
    nq = f_dens.nq
    return (FourPoint(gs.cpu().numpy(), SpinChannel.SING, nq, 0, 2, False, True, True, FrequencyNotation.PP),
            FourPoint(gt.cpu().numpy(), SpinChannel.TRIP, nq, 0, 2, False, True, True, FrequencyNotation.PP))


def pairing_vertices_from_ladder(chi0_inv, gammas, u_rs, v_rs, vrgs, chis, beta, nb, niv_pp, w_chunk=16,
                                 precision=0, window_rhs=True):
    """Singlet and triplet pairing vertices [nq, nb, nb, nb, nb, 2np, 2np] on the device from the
    ladder of both channels (the chain of solve, eliashberg_solver.py:584-629, with
    create_full_vertex_q_r, :78-147, and the w = 0 transform, :44-75).

    chi0_inv: QF [nw][nb^4][2niv][nq] (niv >= niv_pp; the reference's construct_fq_cheap cuts its
    operands to niv_pp first); gammas: [2][nw][n][n] (density, magnetic) compound local vertices;
    u_rs: [2][nb^4]; v_rs: [2][nq][nb^4] or None; vrgs: [2] x QF; chis: [2] x QW. The particle-hole
    full vertices are formed for w_chunk bosonic frequencies at a time and transformed before the
    next range, so F of every w is never held at once; pp entries with |v - v'| >= nw stay zero,
    as in the reference. precision 1 / 2 factors the ladder matrices in complex64 with FP32 /
    split-BF16 matrix-core updates (dgami_full_vertex_prec). With window_rhs, the full vertices of
    each range are formed only at the fermionic frequencies the transform reads (the LU carries
    those right-hand sides only); bosonic frequencies |w| >= 2 np, whose entries stay zero, are
    skipped."""
    import torch

    nw, nb4, nv, nq = chi0_inv.shape
    shape = (nq, nb, nb, nb, nb, 2 * niv_pp, 2 * niv_pp)
    gs = torch.zeros(shape, dtype=K.C128, device=chi0_inv.device)
    gt = torch.zeros(shape, dtype=K.C128, device=chi0_inv.device)
    off = nv // 2 - niv_pp
    for w0 in range(0, nw, w_chunk):
        w1 = min(nw, w0 + w_chunk)
        if w0 >= 2 * niv_pp:  # no pairing entry has |v - v'| this large
            break
        # the transform of |w| in [w0, w1) reads F(w) only at fermionic indices in [w0 + off, 2 np + off)
        window = (w0 + off, 2 * niv_pp + off) if window_rhs else None
        fs = [K.full_vertex(chi0_inv[w0:w1].contiguous(), gammas[c, w0:w1].contiguous(), u_rs[c],
                            None if v_rs is None else v_rs[c], vrgs[c][w0:w1].contiguous(), chis[c][w0:w1].contiguous(),
                            beta, precision=precision, window=window)[0] for c in range(2)]
        K.pairing_vertex_pp_range(fs[0], fs[1], nb, niv_pp, w0, gs, gt)
        del fs
    return gs, gt


class PairingOperator:
    """Device-resident linearised gap operator of one channel."""

    def __init__(self, gamma_pp_irr, chi0pp, q_grid, beta, channel, matrix_precision="fp64", orbital_transforms=None):
        """gamma_pp_irr, chi0pp: arrays or device tensors. matrix_precision "fp32" stores the
        transformed vertex in complex64 (half the bytes per application); the vectors and the
        sums stay complex128. orbital_transforms: per-k (us, conjs) of the irreducible-zone map
        (the reference's auto symmetry mode) or None."""
        import torch

        def dev(x):
            return x.contiguous() if isinstance(x, torch.Tensor) else K.to_dev(np.asarray(x))

        self.nk = tuple(q_grid.nk)
        self.channel = channel
        self.beta = beta
        gam = dev(gamma_pp_irr)
        self.nb, self.np2 = gam.shape[1], gam.shape[-1]
        storage = {"fp64": torch.complex128, "fp32": torch.complex64}[matrix_precision]
        self.g_hat, self.g_flip = K.pairing_setup(gam, K.to_dev(np.asarray(q_grid.irrk_inv), torch.int32),
                                                  self.nk, channel, storage, orbital_transforms)
        n_k = int(np.prod(self.nk))
        self.chi0pp = dev(chi0pp).reshape(n_k, self.nb**4, self.np2).contiguous()
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
    op = PairingOperator(gamma_r_pp.mat, gchi0_q0_pp.mat, config.lattice.q_grid, config.sys.beta, ch,
                         getattr(config.eliashberg, "matvec_precision", "fp64"))
    lin = spla.LinearOperator(shape=(op.dim, op.dim), matvec=op.matvec, dtype=np.complex128)
    n_eig = config.eliashberg.n_eig
    lambdas, gaps = spla.eigsh(lin, k=n_eig, tol=config.eliashberg.epsilon, v0=v0, which="LA", maxiter=10000)
    order = lambdas.argsort()[::-1]
    lambdas, gaps = lambdas[order], gaps[:, order]
    return lambdas, [gaps[:, i].reshape(op.shape) for i in range(n_eig)]


def solve_eliashberg_krylov_schur(op, n_eig=4, tol=1e-12, v0=None, ncv=None, max_restarts=500):
    """Leading eigenpairs (largest real part) of the pairing operator by the Krylov-Schur method
    (G. W. Stewart, SIAM J. Matrix Anal. Appl. 23 (2001) 601-614) with the basis on the device:
    the operator is applied to device vectors and the Arnoldi vectors are orthogonalised there in
    FP64 (classical Gram-Schmidt, applied twice); only the ncv x ncv projected matrix goes to the
    host for its Schur form. No counterpart in the reference (its eigsh runs on the host, with the
    vectors copied for every application, eliashberg_solver.py:430-443). Returns what
    solve_eliashberg_arnoldi returns: (lambdas sorted by real part, descending; gaps; residuals
    |A x - lambda x| / |lambda| of the returned vectors)."""
    import scipy.linalg as sla
    import torch

    n = op.dim
    m = ncv or max(2 * n_eig + 12, 24)
    keep_min = n_eig + max(2, (m - n_eig) // 2)
    dev = op.g_hat.device
    V = torch.zeros((m + 1, n), dtype=torch.complex128, device=dev)
    x0 = torch.as_tensor(np.asarray(v0).ravel(), dtype=torch.complex128, device=dev) if v0 is not None else \
        torch.randn(n, dtype=torch.complex128, device=dev, generator=torch.Generator(device=dev).manual_seed(11))
    V[0] = x0 / torch.linalg.vector_norm(x0)
    H = np.zeros((m + 1, m), dtype=complex)

    def apply(v):
        return op.apply(v.reshape(op.shape)).reshape(n)

    def vdot_rows(Vk, w):
        # V_k^H w in blocks of 12 basis vectors: measured on 1.18M-element vectors, 24 rows in one
        # product take 2.2 ms (the BLAS switches kernels above 12 rows), two blocks of 12 take 0.36 ms
        return torch.cat([Vk[i:i + 12].conj() @ w for i in range(0, Vk.shape[0], 12)])

    j0 = 0
    for _ in range(max_restarts):
        for j in range(j0, m):
            w = apply(V[j])
            h = vdot_rows(V[:j + 1], w)
            w = w - h @ V[:j + 1]
            h2 = vdot_rows(V[:j + 1], w)  # second pass (classical Gram-Schmidt twice)
            w = w - h2 @ V[:j + 1]
            beta = torch.linalg.vector_norm(w)
            H[:j + 1, j] = (h + h2).cpu().numpy()
            H[j + 1, j] = float(beta)
            V[j + 1] = w / beta
        ev = np.linalg.eigvals(H[:m, :m])
        thr = np.sort(ev.real)[::-1][keep_min - 1]
        # Schur form with the eigenvalues of largest real part leading; any leading block of it
        # keeps the Krylov relation A (V Z) = (V Z) T + v_m r with the residual row r = H[m] Z
        T, Z, sdim = sla.schur(H[:m, :m], output="complex", sort=lambda x: x.real >= thr)
        keep = min(sdim, m - 1)
        r = H[m, :m] @ Z
        # Ritz pairs of the kept block: eigenvectors of the upper-triangular T[:keep, :keep]
        lam, S = np.linalg.eig(T[:keep, :keep])
        res = np.abs(r[:keep] @ S) / np.maximum(np.abs(lam), 1e-300)
        want = np.argsort(-lam.real)[:n_eig]
        if np.all(res[want] < tol):
            break
        # restart: keep the leading Schur vectors and the residual spike
        Vk = torch.as_tensor(Z[:, :keep].T.copy(), device=dev) @ V[:m]
        V[:keep] = Vk
        V[keep] = V[m]
        H[:] = 0.0
        H[:keep, :keep] = T[:keep, :keep]
        H[keep, :keep] = r[:keep]
        j0 = keep
    else:
        raise RuntimeError(f"Krylov-Schur did not converge in {max_restarts} restarts (residuals {res[want]})")
    Y = torch.as_tensor((Z[:, :keep] @ S[:, want]).T.copy(), device=dev) @ V[:m]  # Ritz vectors
    lams = lam[want]
    gaps, resid = [], []
    for i in range(n_eig):
        y = Y[i] / torch.linalg.vector_norm(Y[i])
        resid.append(float(torch.linalg.vector_norm(apply(y) - complex(lams[i]) * y)) / abs(lams[i]))
        gaps.append(y.reshape(op.shape).cpu().numpy())
    return lams, gaps, np.array(resid)


def solve_eliashberg_arnoldi(op, n_eig=4, tol=1e-12, v0=None, which="LR"):
    """Leading eigenpairs of the pairing operator by ARPACK's non-Hermitian Arnoldi (eigs) around the
    device matvec, with their residuals |A x - lambda x| / |lambda|. No counterpart in the reference,
    which runs the Hermitian eigsh (eliashberg_solver.py:430-443) on this operator although it is not
    Hermitian; eigs solves the same eigenproblem without that assumption. Returns (lambdas sorted by
    real part, descending; gaps [n_eig][nx,ny,nz,nb,nb,2np]; residuals)."""
    lin = spla.LinearOperator(shape=(op.dim, op.dim), matvec=op.matvec, dtype=np.complex128)
    lambdas, vecs = spla.eigs(lin, k=n_eig, which=which, tol=tol, v0=v0, maxiter=20000)
    order = np.argsort(-lambdas.real)
    lambdas, vecs = lambdas[order], vecs[:, order]
    res = np.array([np.linalg.norm(op.matvec(vecs[:, i]) - lambdas[i] * vecs[:, i]) /
                    (abs(lambdas[i]) * np.linalg.norm(vecs[:, i])) for i in range(n_eig)])
    return lambdas, [vecs[:, i].reshape(op.shape) for i in range(n_eig)], res
