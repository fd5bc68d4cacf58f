"""Pins of the oracle's full-vertex and pairing-vertex functions (CPU only).

The reference's end_2_end data holds the pp full vertices f_irrq_{dens,magn}_pp and the
pairing vertices built from them, but not the local irreducible vertex they come from, so
the ph full vertex itself is pinned by algebra that a dropped term, a wrong sign or a
transposed orbital pair would break:
  * the inverse against the solve-and-sum form of the ladder (itself pinned elsewhere);
  * F1 = beta^2 (D - D X D) against the Bethe-Salpeter form beta^2 (K - K X K) and the
    Woodbury form beta^2 D X K, with D and K assembled here independently;
  * the second term against the reference's own einsum contractions (four_point.py:442-557);
  * the ph -> pp frequency transform against a direct evaluation on Matsubara integers.
"""
import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import golden, ladder_case


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _compound_blocks(chi0_inv_q, nb, nv):
    """Block-diagonal D and the index sets of every v, built with explicit index arrays."""
    nw = chi0_inv_q.shape[4]
    T = nb * nb
    n = T * nv
    m = O.orb_matrix(np.moveaxis(chi0_inv_q, (-2, -1), (0, 1)))  # [w, v, T, T]
    d = np.zeros((nw, n, n), dtype=complex)
    for v in range(nv):
        idx = np.arange(T) * nv + v
        d[:, idx[:, None], idx[None, :]] = m[:, v]
    return d


def _k_matrix(gamma, u, vq, beta, nb, nv):
    """K = (Gamma - U - V^q)/beta^2 with the interaction spread over every (v, v') block."""
    g = gamma.transpose(4, 0, 1, 5, 3, 2, 6).reshape(gamma.shape[4], nb * nb * nv, nb * nb * nv)
    uv = np.kron(O.orb_matrix(u + vq), np.ones((nv, nv)))
    return (g - uv[None]) / beta**2


@pytest.mark.parametrize("nb,niv,niw,nq", [(1, 5, 2, 2), (2, 3, 1, 2), (3, 2, 1, 1)])
def test_auxiliary_chi_rows_sum_to_ladder(nb, niv, niw, nq):
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=11 + nb)
    x = O.auxiliary_chi_q(chi0_inv, gamma, u, v, beta)
    nv = 2 * niv
    for iq in range(nq):
        a = _compound_blocks(chi0_inv[iq], nb, nv) + _k_matrix(gamma, u, v[iq], beta, nb, nv)
        assert np.abs(a @ x[iq] - np.eye(a.shape[-1])).max() < 1e-10
    # sum over v' of X / beta, per column orbital pair, is the summed ladder
    x8 = O.compound_to_ref8(x, nb)
    assert _rel(x8.sum(-1) / beta, O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)) < 1e-12


@pytest.mark.parametrize("nb,niv,niw,nq", [(1, 5, 2, 2), (2, 3, 1, 2), (3, 2, 1, 1)])
def test_full_vertex_first_term_is_bethe_salpeter(nb, niv, niw, nq):
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=21 + nb)
    f = O.full_vertex_q(chi0_inv, gamma, u, v, None, None, beta)
    nv = 2 * niv
    for iq in range(nq):
        d = _compound_blocks(chi0_inv[iq], nb, nv)
        k = _k_matrix(gamma, u, v[iq], beta, nb, nv)
        x = np.linalg.inv(d + k)
        assert _rel(f[iq], beta**2 * (k - k @ x @ k)) < 1e-10
        assert _rel(f[iq], beta**2 * d @ x @ k) < 1e-10


def test_full_vertex_vanishes_without_irreducible_part():
    nb, niv, niw = 2, 3, 1
    chi0_inv, _, u, v, beta = ladder_case(nb, niv, niw, 1, seed=5)
    gamma = np.broadcast_to((u + v[0])[..., None, None, None], (nb,) * 4 + (niw + 1, 2 * niv, 2 * niv)).copy()
    f = O.full_vertex_q(chi0_inv, gamma, u, v, None, None, beta)
    assert np.abs(f).max() < 1e-10 * np.abs(chi0_inv).max() * beta**2


@pytest.mark.parametrize("nb,niv,niw,nq", [(1, 4, 1, 2), (2, 3, 1, 2), (3, 2, 1, 2)])
def test_full_vertex_second_term_matches_reference_contractions(nb, niv, niw, nq):
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=31 + nb)
    rng = np.random.default_rng(4)
    shape = chi0_inv.shape
    vrg = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    chi = rng.standard_normal(shape[:-1]) + 1j * rng.standard_normal(shape[:-1])
    f = O.compound_to_ref8(O.full_vertex_q(chi0_inv, gamma, u, v, vrg, chi, beta), nb)
    f1 = O.compound_to_ref8(O.full_vertex_q(chi0_inv, gamma, u, v, None, None, beta), nb)
    # FourPoint.mul (four_point.py:442-473), Interaction @ FourPoint (:475-518), vn0 @ vn2 (:520-557)
    uq = u[None] + v
    vv = np.einsum("qabcdwv,qdcefwp->qabefwvp", vrg, vrg)
    t1 = np.einsum("qabij,qjiefwvp->qabefwvp", uq, vv)
    us = np.einsum("qabij,qjiefw->qabefw", uq, chi)
    usu = np.einsum("qabijw,qjief->qabefw", us, uq)
    t2 = np.einsum("qabcdw,qdcefwvp->qabefwvp", usu, vv)
    assert _rel(f - f1, t1 - t2) < 1e-12


def _analytic_vertex(nb, niw, niv, seed):
    """A vertex given by a formula of the Matsubara integers (w >= 0), with its full-range values."""
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((nb,) * 4) + 1j * rng.standard_normal((nb,) * 4)

    def g(w, v, vp):
        return c * (1.0 + 0.1j * w + 0.03 * v - 0.05j * vp + 0.002 * v * vp + 0.01j * w * w)

    def g_full(w, v, vp):  # F(-w, v, v') = conj F(w, -v-1, -v'-1)
        return g(w, v, vp) if w >= 0 else np.conj(g(-w, -v - 1, -vp - 1))

    vs = O.vn_int(niv)
    f = np.empty((1,) + (nb,) * 4 + (niw + 1, 2 * niv, 2 * niv), dtype=complex)
    for w in range(niw + 1):
        for i, v in enumerate(vs):
            for j, vp in enumerate(vs):
                f[0, ..., w, i, j] = g(w, v, vp)
    return f, g_full


@pytest.mark.parametrize("nb,niw,niv,niv_pp", [(1, 6, 5, 3), (2, 4, 3, 2), (1, 3, 4, 3)])
def test_vertex_pp_w0_on_matsubara_integers(nb, niw, niv, niv_pp):
    f, g_full = _analytic_vertex(nb, niw, niv, seed=nb)
    got = O.vertex_pp_w0(f, niv_pp)
    vs = O.vn_int(niv_pp)
    for i, v in enumerate(vs):
        for j, vpp in enumerate(vs):
            w = v - vpp
            # the pp partner frequency -v''-1 is the flipped fermionic index
            want = -g_full(w, v, -vpp - 1) if abs(w) <= niw else np.zeros((nb,) * 4)
            assert np.abs(got[0, ..., i, j] - want).max() < 1e-12


def test_pairing_vertices_match_reference_fixtures():
    sing, trip = O.pairing_vertices(golden("e2e_f_irrq_dens_pp.npy"), golden("e2e_f_irrq_magn_pp.npy"))
    assert _rel(sing, golden("e2e_gamma_irrq_sing_pp.npy")) < 1e-6
    assert _rel(trip, golden("e2e_gamma_irrq_trip_pp.npy")) < 1e-6


def test_vertex_pp_w0_conjugation_symmetry_matches_fixture():
    """For w = v - v' != 0 the transform takes F(-w) from F(w) by conjugation with both fermionic
    axes flipped, so F_pp(i, j) = conj F_pp(2np-1-i, 2np-1-j) off the diagonal. The reference's own
    f_irrq_dens_pp.npy satisfies this exactly; the oracle's transform must too, for any F."""
    fix = golden("e2e_f_irrq_dens_pp.npy")
    p2 = fix.shape[-1]
    off = ~np.eye(p2, dtype=bool)
    assert np.abs((fix - np.conj(fix[..., ::-1, ::-1]))[..., off]).max() == 0.0
    rng = np.random.default_rng(9)
    f = rng.standard_normal((2, 2, 2, 2, 2, 7, 8, 8)) + 1j * rng.standard_normal((2, 2, 2, 2, 2, 7, 8, 8))
    pp = O.vertex_pp_w0(f, 3)
    off = ~np.eye(6, dtype=bool)
    assert np.abs((pp - np.conj(pp[..., ::-1, ::-1]))[..., off]).max() == 0.0
