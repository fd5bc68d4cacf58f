"""The oracle's local Schwinger-Dyson functions against the reference's fixtures, closed forms and
exact identities (CPU only).

Fixtures (complex64 files of the reference's test data, so agreement is bounded near 1e-6):
* end_2_end: gchi0_loc.npy is the local bubble of g_loc.npy (beta = 12.5, niw = 20, niv_full = 30);
  siw_dga_local.npy is the local Sigma of vrg_{dens,magn}_loc, chi_{dens,magn}_loc and g_dmft.npy
  (U = 8 on aaaa), whose Hartree-Fock constant uses the DMFT occupation of the reference's
  1p-data.hdf5: checked here as a constant, and exactly, with the occupation read from that file, in
  test_dmft_interface.py.
* srvo3_end2end: f_{dens,magn}_loc.npy is the full vertex of gamma_{dens,magn}_loc.npy with the bubble
  of g_loc.npy, no shell (niv_shell = 0), 3 orbitals. The run's beta is not stored in these files; the
  tail of G, Im G(v_n) -> -1 / v_n, gives beta = 1.0 (checked below), and F matches only with it.
"""
import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import golden


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _u_aaaa(nb, U):
    u = np.zeros((nb,) * 4)
    for a in range(nb):
        u[a, a, a, a] = U
    return u


def test_local_bubble_matches_reference_fixture():
    g = golden("e2e_g_loc.npy").astype(complex)
    ref = golden("e2e_gchi0_loc.npy")
    chi0 = O.local_bubble(g, 20, 30, 12.5)
    assert chi0.shape == ref.shape
    assert rel(chi0, ref) < 1e-6


def test_local_bubble_brute_force():
    """-beta G_ad(v) G_cb(v - w) element by element on a random G."""
    rng = np.random.default_rng(3)
    nb, nivg, beta = 2, 12, 5.0
    g = rng.standard_normal((nb, nb, 2 * nivg)) + 1j * rng.standard_normal((nb, nb, 2 * nivg))
    chi0 = O.local_bubble(g, 3, 4, beta, filter_values=False)
    for (a, b, c, d, w, v) in [(0, 1, 1, 0, 2, 3), (1, 0, 0, 1, 3, 7), (0, 0, 1, 1, 0, 0)]:
        n = v - 4
        assert abs(chi0[a, b, c, d, w, v] - (-beta * g[a, d, nivg + n] * g[c, b, nivg + n - w])) < 1e-13


def test_srvo3_beta_from_the_green_function_tail():
    g = golden("srvo3_g_loc.npy")
    niv = g.shape[-1] // 2
    n = np.arange(niv - 5, niv)
    beta = np.pi * (2 * n + 1) * (-g[0, 0, niv + n].imag)
    assert np.abs(beta - 1.0).max() < 2e-3


@pytest.mark.parametrize("channel", ["dens", "magn"])
def test_full_vertex_from_gamma_matches_srvo3_fixture(channel):
    g = golden("srvo3_g_loc.npy").astype(complex)
    gamma = golden(f"srvo3_gamma_{channel}_loc.npy").astype(complex)
    ref = golden(f"srvo3_f_{channel}_loc_w0_1_10.npy")
    ws = [0, 1, 10]
    chi0 = O.local_bubble(g, 10, 10, 1.0)[..., ws, :]
    f = O.full_vertex_from_gamma(gamma[..., ws, :, :], chi0, np.zeros((3,) * 4), 1.0)
    assert rel(f, ref) < 1e-6
    # the wrong temperature does not match
    f2 = O.full_vertex_from_gamma(gamma[..., ws, :, :], O.local_bubble(g, 10, 10, 1.1)[..., ws, :],
                                  np.zeros((3,) * 4), 1.1)
    assert rel(f2, ref) > 1e-3


def test_local_sigma_matches_reference_fixture_up_to_its_hartree_constant():
    g = golden("e2e_g_dmft.npy").astype(complex)
    vd, vm = golden("e2e_vrg_dens_loc.npy").astype(complex), golden("e2e_vrg_magn_loc.npy").astype(complex)
    cd, cm = golden("e2e_chi_dens_loc.npy").astype(complex), golden("e2e_chi_magn_loc.npy").astype(complex)
    ref = golden("e2e_siw_dga_local.npy")[0, 0, 0].astype(complex)
    u = _u_aaaa(2, 8.0)
    s = O.local_sigma(vd, vm, cd, cm, g, u, 12.5, 20, 20)
    diff = ref - s
    const = diff[..., :1]
    assert np.abs(diff - const).max() < 2e-6  # frequency independent
    const = const[..., 0]
    # the constant is U_dens . n with U_dens,aaaa = 8 and n the DMFT occupation per spin: real, diagonal,
    # and close to the occupation of the Matsubara sum of G
    assert np.abs(const - np.diag(np.diag(const))).max() < 1e-6
    assert np.abs(const.imag).max() < 1e-6
    beta = 12.5
    niv = g.shape[-1] // 2
    n = 0.5 * np.eye(2) + np.sum(g - np.eye(2)[..., None] / (1j * O.vn(niv, beta)), axis=-1) / beta
    assert np.abs(np.diag(const).real / 8.0 - np.diag(n).real).max() < 0.015
    # a dropped magnetic channel or a flipped G argument changes Sigma by far more than the fixture's noise
    s_no_m = O.local_sigma(vd, 0 * vm, cd, 0 * cm, g, u, 12.5, 20, 20)
    assert np.abs((ref - s_no_m) - (ref - s_no_m)[..., :1]).max() > 1e-2


def test_local_hartree_closed_forms():
    n = np.array([[0.4]])
    assert np.allclose(O.local_hartree(np.full((1, 1, 1, 1), 3.0), n), 3.0 * 0.4)
    occ = np.diag([0.3, 0.45])
    assert np.allclose(O.local_hartree(_u_aaaa(2, 8.0), occ), 8.0 * occ)


def test_generalized_chi_of_a_noninteracting_g2_is_the_bubble():
    """Wick's theorem for U = 0: G2_upup = delta_w0 G_ab(v) G_cd(v') - delta_vv' G_ad(v) G_cb(v - w),
    G2_updn = delta_w0 G_ab(v) G_cd(v'); dens = upup + updn and magn = upup - updn. Both channels'
    generalized susceptibilities equal the bubble chi0 on the diagonal v = v' and vanish elsewhere."""
    rng = np.random.default_rng(5)
    nb, nivg, niw, niv, beta = 2, 16, 3, 5, 4.0
    g = rng.standard_normal((nb, nb, 2 * nivg)) + 1j * rng.standard_normal((nb, nb, 2 * nivg))
    chi0 = O.local_bubble(g, niw, niv, beta, filter_values=False)
    exch = -chi0 / beta  # G_ad(v) G_cb(v - w)
    eye = np.eye(2 * niv)
    gs = g[..., nivg - niv:nivg + niv]
    disc = np.zeros((nb,) * 4 + (niw + 1, 2 * niv, 2 * niv), dtype=complex)
    disc[:, :, :, :, 0] = gs[:, :, None, None, :, None] * gs[None, None, :, :, None, :]
    upup = disc - np.einsum("abcdwv,vp->abcdwvp", exch, eye)
    updn = disc
    diag = np.einsum("abcdwv,vp->abcdwvp", chi0, eye)
    assert np.abs(O.generalized_chi(upup + updn, g, beta, "dens", niv) - diag).max() < 1e-12
    assert np.abs(O.generalized_chi(upup - updn, g, beta, "magn", niv) - diag).max() < 1e-12


def _box(seed, nb=2, niw=2, nc=3, ns=2, beta=3.0):
    rng = np.random.default_rng(seed)
    nivg = nc + ns + niw + 8
    v = O.vn(nivg, beta)
    eps = rng.standard_normal((nb, nb)) * 0.3
    eps = eps + eps.T
    g = np.linalg.inv(1j * v[:, None, None] * np.eye(nb) - eps[None] - 0.5j * np.sign(v)[:, None, None] * np.eye(nb))
    g = g.transpose(1, 2, 0)
    chi0 = O.local_bubble(g, niw, nc + ns, beta, filter_values=False)
    return rng, g, chi0, beta


def test_gamma_shell_without_a_shell_is_the_plain_inverse_difference():
    rng, g, chi0, beta = _box(1, ns=0)
    nb, nw, nv = 2, chi0.shape[4], chi0.shape[5]
    chi = rng.standard_normal((nb,) * 4 + (nw, nv, nv)) + 1j * rng.standard_normal((nb,) * 4 + (nw, nv, nv))
    chi = chi * 0.1 + np.einsum("abcdwv,vp->abcdwvp", chi0, np.eye(nv))
    u = rng.standard_normal((nb,) * 4)
    ur = O.u_as_channel(u, "dens")
    gam = O.gamma_shell(chi, chi0, ur, beta, nv // 2)
    inv0 = np.linalg.inv(O.compound_local_vertex(np.einsum("abcdwv,vp->abcdwvp", chi0, np.eye(nv))))
    plain = beta**2 * (np.linalg.inv(O.compound_local_vertex(chi)) - inv0)
    assert np.abs(O.compound_local_vertex(gam) - plain).max() / np.abs(plain).max() < 1e-12


def test_gamma_shell_recovers_a_vertex_that_equals_u_outside_the_core():
    """chi from a Gamma that is arbitrary in the core and U_r in the shell, on the full box:
    chi = (chi0^-1 + Gamma / beta^2)^-1, cut to the core. The shell-corrected Gamma of that chi is the
    planted core Gamma exactly (the two matrices inverted differ only in their core block, so their
    Schur complements differ by it); the core-only formula beta^2 (chi^-1 - chi0_core^-1) is not."""
    rng, g, chi0, beta = _box(2)
    nb, nw, nvf = 2, chi0.shape[4], chi0.shape[5]
    nc = 3
    u = rng.standard_normal((nb,) * 4)
    ur = O.u_as_channel(u, "magn")
    core = 3.0 * (rng.standard_normal((nb,) * 4 + (nw, 2 * nc, 2 * nc)) +
                  1j * rng.standard_normal((nb,) * 4 + (nw, 2 * nc, 2 * nc)))
    gfull = np.array(O._u_full(ur, nw, nvf))
    d = (nvf - 2 * nc) // 2
    gfull[..., d:d + 2 * nc, d:d + 2 * nc] = core
    diag0 = np.einsum("abcdwv,vp->abcdwvp", chi0, np.eye(nvf))
    a = np.linalg.inv(O.compound_local_vertex(diag0)) + O.compound_local_vertex(gfull) / beta**2
    chi = O.local_uncompound(np.linalg.inv(a), nb)[..., d:d + 2 * nc, d:d + 2 * nc]
    gam = O.gamma_shell(chi, chi0, ur, beta, nc)
    assert np.abs(gam - core).max() / np.abs(core).max() < 1e-10
    chi0c = chi0[..., d:d + 2 * nc]
    naive = beta**2 * (np.linalg.inv(O.compound_local_vertex(chi)) -
                       np.linalg.inv(O.compound_local_vertex(np.einsum("abcdwv,vp->abcdwvp", chi0c,
                                                                       np.eye(2 * nc)))))
    assert np.abs(O.local_uncompound(naive, nb) - core).max() / np.abs(core).max() > 1e-3


def test_full_vertex_of_a_constant_one_band_vertex():
    """One band, Gamma = U everywhere: F = U / (1 + U sum_v chi0(v) / beta^2) in every (v, v')."""
    _, g, chi0, beta = _box(4, nb=1)
    nw, nvf = chi0.shape[4], chi0.shape[5]
    U = 2.5
    f = O.full_vertex_from_gamma(np.full((1, 1, 1, 1, nw, 4, 4), U, dtype=complex), chi0,
                                 np.full((1, 1, 1, 1), U), beta)
    expect = U / (1 + U * chi0[0, 0, 0, 0].sum(axis=-1) / beta**2)
    assert np.abs(f[0, 0, 0, 0] - expect[:, None, None]).max() / abs(U) < 1e-12


def test_full_vertex_from_gamma_equals_the_bse_form_without_a_shell():
    """create_full_vertex (local_sde.py:79-84): beta^2 (chi0^-1 - chi0^-1 chi chi0^-1) with
    chi = (chi0^-1 + Gamma / beta^2)^-1 is Gamma [1 + chi0 Gamma / beta^2]^-1."""
    rng, g, chi0, beta = _box(6, ns=0)
    nb, nw, nv = 2, chi0.shape[4], chi0.shape[5]
    gam = rng.standard_normal((nb,) * 4 + (nw, nv, nv)) + 1j * rng.standard_normal((nb,) * 4 + (nw, nv, nv))
    f = O.full_vertex_from_gamma(gam, chi0, np.zeros((nb,) * 4), beta)
    k = np.linalg.inv(O.compound_local_vertex(np.einsum("abcdwv,vp->abcdwvp", chi0, np.eye(nv))))
    chi = np.linalg.inv(k + O.compound_local_vertex(gam) / beta**2)
    bse = beta**2 * (k - k @ chi @ k)
    assert np.abs(O.compound_local_vertex(f) - bse).max() / np.abs(bse).max() < 1e-11
