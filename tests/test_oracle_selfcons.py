"""Pins of the oracle's Hartree-Fock, self-energy assembly, filling and chemical-potential
functions (CPU only).

The reference's end_2_end run holds mu before and after its first iteration
(mu_history.npy) and the self-energy it produced (sigma_dga_iteration_1.npy): the target
filling from the DMFT self-energy and the secant update on the new self-energy must give
the second mu back. The rest is pinned by closed forms.
"""
import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import GOLDEN, golden

BETA = 12.5


def _e2e_ek():
    hk, _ = O.read_hk(f"{GOLDEN}/e2e_wannier.hk")
    return hk.reshape(4, 4, 1, 2, 2)


def test_update_mu_reproduces_reference_mu_history():
    ek = _e2e_ek()
    mu = golden("e2e_mu_history.npy").real
    sd = golden("e2e_sigma_dmft.npy")[0, 0, 0]
    s1 = golden("e2e_sigma_dga_iteration_1.npy")
    m0, _ = O.fit_smom(sd, BETA)
    n_target, _, _ = O.fill_nonlocal(sd[None], ek, mu[0], BETA, m0)
    m1, _ = O.fit_smom(s1, BETA)
    assert abs(O.update_mu(mu[0], n_target, ek, s1, BETA, m1) - mu[1]) < 1e-6


def test_reference_iteration_keeps_dmft_tails_outside_core_window():
    """sigma_dga_iteration_1 (mixing 1) is the DMFT self-energy outside |v| < niv_core = 20."""
    sd = golden("e2e_sigma_dmft.npy")[0, 0, 0]
    s1 = golden("e2e_sigma_dga_iteration_1.npy")
    niv, nc = 100, 20
    core = s1[..., niv - nc: niv + nc]
    out = O.sigma_assemble(core, np.zeros((2, 2)), np.zeros((16, 2, 2)), np.zeros((2, 2, 2 * nc)), sd, niv)
    assert np.abs(out - s1).max() == 0.0
    # and the window is the core one: the first frequency inside differs from DMFT
    assert np.abs(s1[..., niv - nc] - sd[None, ..., niv - nc]).max() > 1e-3


def test_linear_mixing():
    rng = np.random.default_rng(0)
    new = rng.standard_normal((3, 2, 2, 8)) + 0j
    old = rng.standard_normal((3, 2, 2, 8)) + 0j
    out = O.sigma_assemble(new, np.zeros((2, 2)), np.zeros((3, 2, 2)), np.zeros((2, 2, 8)), new[0], 4, old, 0.3)
    assert np.abs(out - (0.3 * new + 0.7 * old)).max() < 1e-15


def test_hartree_fock_local_interaction_closed_form():
    rng = np.random.default_rng(1)
    nb, nk = 2, (4, 6, 1)
    u = rng.standard_normal((nb,) * 4)
    occ_k = rng.standard_normal(nk + (nb, nb)) + 1j * rng.standard_normal(nk + (nb, nb))
    occ = occ_k.mean(axis=(0, 1, 2))
    h, f = O.hartree_fock(u, None, occ, occ_k)
    want_f = -np.einsum("abcd,dc->ab", u, occ)
    assert np.abs(h - 2 * np.einsum("abcd,dc->ab", u, occ)).max() < 1e-13
    assert np.abs(f - want_f[None, None, None]).max() < 1e-13


def test_fock_of_cosine_interaction_closed_form():
    """V(q) = V0 (cos qx + cos qy), n(k) = a + b cos kx: (1/N) sum_q V(q) n(k+q) = V0 b cos(kx) / 2."""
    nx, ny = 8, 6
    kx = 2 * np.pi * np.arange(nx) / nx
    ky = 2 * np.pi * np.arange(ny) / ny
    v0, a, b = 0.7, 0.45, 0.2
    vq = v0 * (np.cos(kx)[:, None] + np.cos(ky)[None, :])
    nkk = a + b * np.cos(kx)[:, None] + 0 * ky[None, :]
    u = np.full((1, 1, 1, 1), 2.0)
    _, f = O.hartree_fock(u, vq.reshape(nx, ny, 1, 1, 1, 1, 1), np.array([[a]]),
                          nkk.reshape(nx, ny, 1, 1, 1).astype(complex))
    want = -(2.0 * a + v0 * b * np.cos(kx)[:, None] / 2 + 0 * ky[None, :])
    assert np.abs(f[:, :, 0, 0, 0] - want).max() < 1e-13


def test_fit_smom_closed_form():
    niv, nb = 40, 2
    v = O.vn(niv, BETA)
    m0 = np.array([[0.3, 0.1], [0.1, -0.2]])
    m1 = np.array([[2.0, 0.5], [0.5, 1.0]])
    sigma = m0[..., None] + m1[..., None] / (1j * v)
    mom0, mom1 = O.fit_smom(sigma, BETA)
    assert np.abs(mom0 - m0).max() < 1e-13
    assert np.abs(mom1 + m1).max() < 1e-13  # Im(m1 / (i v)) v = -m1


def test_fillings_fermi_dirac_limit():
    """Sigma = 0, smom0 = 0: the k-dependent model is G itself, so occ_k is the Fermi function of e_k."""
    ek = _e2e_ek()
    mu = 0.4
    nb = 2
    sig = np.zeros((1, nb, nb, 60), dtype=complex)
    n, _, occ_k = O.fill_nonlocal(sig, ek, mu, BETA, np.zeros((nb, nb)))
    w, vecs = np.linalg.eigh(ek.reshape(-1, nb, nb))
    f = 1 / (1 + np.exp(BETA * (w - mu)))
    rho = np.einsum("kab,kb,kcb->kac", vecs, f, vecs.conj())
    assert np.abs(occ_k.reshape(-1, nb, nb) - rho).max() < 1e-12
    assert abs(n - 2 * f.sum() / len(f)) < 1e-12
    # the local model converges to the same filling as the frequency box grows
    sig_big = np.zeros((1, nb, nb, 4000), dtype=complex)
    assert abs(O.total_fill(mu, ek, sig_big, BETA, np.zeros((nb, nb))) - n) < 1e-3


@pytest.mark.parametrize("alpha", [0.5, 1.0])
def test_mixing_of_full_assembly(alpha):
    rng = np.random.default_rng(3)
    nk, nb, nc, nout, nd = 4, 2, 3, 5, 7
    snl = rng.standard_normal((nk, nb, nb, 2 * nc)) + 0j
    h = rng.standard_normal((nb, nb)) + 0j
    f = rng.standard_normal((nk, nb, nb)) + 0j
    ds = rng.standard_normal((nb, nb, 2 * nc)) + 0j
    sd = rng.standard_normal((nb, nb, 2 * nd)) + 0j
    old = rng.standard_normal((nk, nb, nb, 2 * nout)) + 0j
    out = O.sigma_assemble(snl, h, f, ds, sd, nout, old, alpha)
    val = np.tile(sd[None, ..., nd - nout: nd + nout], (nk, 1, 1, 1))
    val[..., nout - nc: nout + nc] = snl + h[None, ..., None] + f[..., None] + ds[None]
    assert np.abs(out - (alpha * val + (1 - alpha) * old)).max() < 1e-14
