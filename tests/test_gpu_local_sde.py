"""The local Schwinger-Dyson producer on the GPU against the oracle and the reference's fixtures
(needs an MI355X).

* local bubble of the end_2_end g_loc.npy vs the oracle and gchi0_loc.npy;
* local Sigma from the end_2_end vrg/chi fixtures vs the oracle (with a Hartree-Fock occupation)
  and vs siw_dga_local.npy (with the occupation its constant implies);
* SrVO3 full vertices from gamma_{dens,magn}_loc.npy (beta = 1, no shell) vs the oracle and the
  reference's f_{dens,magn}_loc.npy;
* the whole local SDE (both channels, shell-corrected Gamma, F, vrg, chi(w), Sigma) on g2 built
  from a planted vertex, vs the oracle chain; the planted core vertex is recovered.
"""
import sys

import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import golden

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture
def cfg():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import config

    saved = (dict(vars(config.sys)), dict(vars(config.box)))
    yield config
    vars(config.sys).update(saved[0])
    vars(config.box).update(saved[1])


def _u_aaaa(nb, U):
    u = np.zeros((nb,) * 4, dtype=complex)
    for a in range(nb):
        u[a, a, a, a] = U
    return u


def test_local_bubble_gpu(cfg):
    from dgamore_amd import kernels as K
    from dgamore_amd import local_sde as LS

    say("local bubble, end_2_end g_loc, niw=20, niv=30")
    g = golden("e2e_g_loc.npy").astype(complex)
    chi0 = K.qf_to_ref4(LS.local_bubble(K.to_dev(g), 20, 30, 12.5), 2)[0].cpu().numpy()
    assert rel(chi0, O.local_bubble(g, 20, 30, 12.5)) < 1e-13
    assert rel(chi0, golden("e2e_gchi0_loc.npy")) < 1e-6


def test_local_sigma_gpu(cfg):
    from dgamore_amd import kernels as K
    from dgamore_amd import local_sde as LS
    from dgamore_amd.four_point import LocalFourPoint, SpinChannel
    from dgamore_amd.pipeline import u_as_channel

    say("local Sigma from the end_2_end vrg/chi fixtures")
    g = golden("e2e_g_dmft.npy").astype(complex)
    vd, vm = golden("e2e_vrg_dens_loc.npy").astype(complex), golden("e2e_vrg_magn_loc.npy").astype(complex)
    cd, cm = golden("e2e_chi_dens_loc.npy").astype(complex), golden("e2e_chi_magn_loc.npy").astype(complex)
    u = _u_aaaa(2, 8.0)
    occ = np.array([[0.42, 0.03 - 0.01j], [0.03 + 0.01j, 0.41]])
    qf = lambda x: K.ref4_to_qf(K.to_dev(x[None]))  # noqa: E731
    qw = lambda x: K.ref4w_to_qw(K.to_dev(x[None]))  # noqa: E731
    s = K.local_sigma(qf(vd), qf(vm), qw(cd), qw(cm), K.orb_to_m(K.to_dev(u_as_channel(u, "dens"))),
                      K.orb_to_m(K.to_dev(u_as_channel(u, "magn"))), K.to_dev(u), K.to_dev(g), 12.5,
                      K.to_dev(occ)).cpu().numpy()
    ref = O.local_sigma(vd, vm, cd, cm, g, u, 12.5, 20, 20) + O.local_hartree(u, occ)[..., None]
    assert rel(s, ref) < 1e-12
    # the reference-shaped seam with the occupation the fixture's constant implies
    fix = golden("e2e_siw_dga_local.npy")[0, 0, 0]
    const = (fix - O.local_sigma(vd, vm, cd, cm, g, u, 12.5, 20, 20))[..., 0]
    cfg.sys.beta = 12.5
    cfg.sys.occ_dmft = np.diag(np.diag(const).real / 8.0)
    se = LS.get_loc_self_energy_vrg(LocalFourPoint(vd, SpinChannel.DENS, 1, 1, False),
                                    LocalFourPoint(vm, SpinChannel.MAGN, 1, 1, False),
                                    LocalFourPoint(cd, SpinChannel.DENS, 1, 0, False),
                                    LocalFourPoint(cm, SpinChannel.MAGN, 1, 0, False), g, u)
    assert se.mat.shape == (1, 1, 1, 2, 2, 40)
    assert np.abs(se.mat[0, 0, 0] - fix).max() < 2e-6


@pytest.mark.parametrize("channel", ["dens", "magn"])
def test_full_vertex_from_gamma_gpu_srvo3(cfg, channel):
    from dgamore_amd import kernels as K
    from dgamore_amd import local_sde as LS
    from dgamore_amd.four_point import LocalFourPoint, SpinChannel

    say(f"SrVO3 local F from Gamma ({channel}), 3 orbitals, niw=10, niv=10, beta=1")
    cfg.sys.beta = 1.0
    g = golden("srvo3_g_loc.npy").astype(complex)
    gamma = golden(f"srvo3_gamma_{channel}_loc.npy").astype(complex)
    chi0 = K.qf_to_ref4(LS.local_bubble(K.to_dev(g), 10, 10, 1.0), 3)[0].cpu().numpy()
    f = LS.create_full_vertex_from_gamma(LocalFourPoint(gamma, SpinChannel(channel), 1, 2, False),
                                         LocalFourPoint(chi0, SpinChannel.NONE, 1, 1, False),
                                         np.zeros((3,) * 4)).mat
    ws = [0, 1, 10]
    assert rel(f[..., ws, :, :], golden(f"srvo3_f_{channel}_loc_w0_1_10.npy")) < 1e-6
    assert rel(f, O.full_vertex_from_gamma(gamma, chi0, np.zeros((3,) * 4), 1.0)) < 1e-11


def _planted_g2(g, chi0_full, ur, channel, beta, nc, rng):
    """g2 of one channel whose shell-corrected vertex is a planted core vertex (U_r in the shell)."""
    nb, nw, nvf = chi0_full.shape[0], chi0_full.shape[4], chi0_full.shape[5]
    core = 2.0 * (rng.standard_normal((nb,) * 4 + (nw, 2 * nc, 2 * nc)) +
                  1j * rng.standard_normal((nb,) * 4 + (nw, 2 * nc, 2 * nc)))
    gfull = np.array(O._u_full(ur, nw, nvf))
    d = (nvf - 2 * nc) // 2
    gfull[..., d:d + 2 * nc, d:d + 2 * nc] += core
    diag0 = np.einsum("abcdwv,vp->abcdwvp", chi0_full, np.eye(nvf))
    a = np.linalg.inv(O.compound_local_vertex(diag0)) + O.compound_local_vertex(gfull) / beta**2
    chi = O.local_uncompound(np.linalg.inv(a), nb)[..., d:d + 2 * nc, d:d + 2 * nc]
    g2 = chi / beta
    if channel == "dens":
        nivg = g.shape[-1] // 2
        gs = g[..., nivg - nc:nivg + nc]
        g2[:, :, :, :, 0] += 2.0 * gs[:, :, None, None, :, None] * gs[None, None, :, :, None, :]
    return g2, gfull[..., d:d + 2 * nc, d:d + 2 * nc]


def test_perform_local_schwinger_dyson_gpu(cfg):
    from dgamore_amd import local_sde as LS
    from dgamore_amd.four_point import LocalFourPoint, SpinChannel

    beta, niw, nc, ns = 12.5, 4, 6, 4
    nf = nc + ns
    say(f"local SDE, 2 orbitals, beta={beta}, niw={niw}, niv_core={nc}, niv_shell={ns}, planted vertices")
    g = golden("e2e_g_dmft.npy").astype(complex)
    u = _u_aaaa(2, 8.0) + 0.5 * np.einsum("ac,bd->abcd", np.eye(2), np.eye(2))
    occ = np.diag([0.43, 0.44]).astype(complex)
    rng = np.random.default_rng(11)
    chi0 = O.local_bubble(g, niw, nf, beta, filter_values=False)
    g2, planted = {}, {}
    for ch in ("dens", "magn"):
        g2[ch], planted[ch] = _planted_g2(g, chi0, O.u_as_channel(u, ch), ch, beta, nc, rng)
    cfg.sys.beta = beta
    cfg.sys.occ_dmft = occ
    cfg.box.niw_core, cfg.box.niv_core, cfg.box.niv_shell, cfg.box.niv_full = niw, nc, ns, nf
    out = LS.perform_local_schwinger_dyson(
        g, LocalFourPoint(g2["dens"], SpinChannel.DENS, 1, 2, False),
        LocalFourPoint(g2["magn"], SpinChannel.MAGN, 1, 2, False), u)
    gd, gm, cd, cm, vd, vm, fd, fm, xd, xm, sig = out
    got = {"dens": (gd, cd, vd, fd, xd), "magn": (gm, cm, vm, fm, xm)}

    d = ns
    inv_core = O.block_invert_vn1(chi0[None][..., d:d + 2 * nc])
    full_sum = chi0.sum(-1)[None] / beta**2
    core_sum = chi0[..., d:d + 2 * nc].sum(-1)[None] / beta**2
    ref = {}
    for ch in ("dens", "magn"):
        ur = O.u_as_channel(u, ch)
        chi_r = O.generalized_chi(g2[ch], g, beta, ch, nc)
        gam = O.gamma_shell(chi_r, chi0, ur, beta, nc)
        f = O.full_vertex_from_gamma(gam, chi0, ur, beta)
        aux = O.ladder_sum_solve(inv_core, gam, ur, np.zeros((1,) + (2,) * 4), beta)
        vrg = O.vrg(inv_core, aux, beta)
        chi_p = O.chi_phys_shell(O.sum_over_vn(aux, beta), full_sum, core_sum, ur, np.zeros((1,) + (2,) * 4))
        ref[ch] = (gam, chi_p[0], vrg[0], f, chi_r)
        names = ("gamma", "chi(w)", "vrg", "F", "chi_r")
        for name, a, b in zip(names, got[ch], ref[ch]):
            e = rel(a.mat, b)
            say(f"  {ch} {name}: rel. error {e:.2e}")
            assert a.mat.shape == b.shape, (name, a.mat.shape, b.shape)
            assert e < 1e-10, (ch, name, e)
        assert rel(gam, planted[ch]) < 1e-9
    s_ref = O.local_sigma(ref["dens"][2], ref["magn"][2], ref["dens"][1], ref["magn"][1], g, u, beta, niw, nc)
    s_ref = s_ref + O.local_hartree(u, occ)[..., None]
    e = rel(sig.mat[0, 0, 0], s_ref)
    say(f"  Sigma_loc: rel. error {e:.2e}")
    assert e < 1e-10
