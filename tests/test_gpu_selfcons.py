"""Parity of the Hartree-Fock, self-energy assembly, filling and mu-update paths with the oracle.

Complex128 / FP64: 1e-10 relative; the mu update against the reference's own mu_history to 1e-6.
"""
import sys

import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL = 1e-10
BETA = 12.5


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def K():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import kernels

    return kernels


def _e2e_ek():
    hk, _ = O.read_hk(f"{GOLDEN}/e2e_wannier.hk")
    return hk.reshape(4, 4, 1, 2, 2)


@pytest.mark.parametrize("nb,nk,with_v", [(1, (8, 8, 1), True), (2, (4, 6, 1), True), (3, (6, 6, 1), False),
                                          (2, (4, 4, 4), True)])
def test_hartree_fock(K, nb, nk, with_v):
    say(f"hartree_fock nb={nb} nk={nk} with_v={with_v}")
    rng = np.random.default_rng(nb)
    u = rng.standard_normal((nb,) * 4)
    v = (rng.standard_normal(nk + (nb,) * 4) + 1j * rng.standard_normal(nk + (nb,) * 4)) if with_v else None
    occ_k = rng.standard_normal(nk + (nb, nb)) + 1j * rng.standard_normal(nk + (nb, nb))
    occ = occ_k.mean(axis=(0, 1, 2))
    h_ref, f_ref = O.hartree_fock(u, v, occ, occ_k)
    N = int(np.prod(nk))
    dv = None if v is None else K.to_dev(v.reshape(N, nb, nb, nb, nb))
    h, f = K.hartree_fock(K.to_dev(u), dv, K.to_dev(occ), K.to_dev(occ_k.reshape(N, nb, nb)), nk)
    torch.cuda.synchronize()
    assert rel(h, h_ref) < RTOL
    assert rel(f, f_ref.reshape(N, nb, nb)) < RTOL


@pytest.mark.parametrize("alpha,with_old", [(1.0, False), (0.3, True)])
def test_sigma_assemble(K, alpha, with_old):
    say(f"sigma_assemble alpha={alpha} old={with_old}")
    rng = np.random.default_rng(5)
    nk, nb, nc, nout, nd = 12, 2, 5, 9, 11
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)  # noqa: E731
    snl, h, f, ds, sd = c(nk, nb, nb, 2 * nc), c(nb, nb), c(nk, nb, nb), c(nb, nb, 2 * nc), c(nb, nb, 2 * nd)
    old = c(nk, nb, nb, 2 * nout) if with_old else None
    ref = O.sigma_assemble(snl, h, f, ds, sd, nout, old, alpha)
    got = K.sigma_assemble(K.to_dev(snl), K.to_dev(h), K.to_dev(f), K.to_dev(ds), K.to_dev(sd), nout,
                           None if old is None else K.to_dev(old), alpha)
    torch.cuda.synchronize()
    assert rel(got, ref) < 1e-15


def test_fill_and_mu_update_reproduce_reference(K):
    """Target filling from the DMFT self-energy, secant update on the reference's new self-energy."""
    say("fillings + update_mu end_2_end (mu_history, sigma_dga_iteration_1)")
    from dgamore_amd import greens_function as gf
    from dgamore_amd.self_energy import fit_smom

    ek = _e2e_ek()
    mu = golden("e2e_mu_history.npy").real
    sd = golden("e2e_sigma_dmft.npy")[0, 0, 0]
    s1 = golden("e2e_sigma_dga_iteration_1.npy")
    m0, m0_1 = fit_smom(sd, BETA)
    r0, r0_1 = O.fit_smom(sd, BETA)
    assert rel(m0, r0) < 1e-13 and rel(m0_1, r0_1) < 1e-13
    n, occ, occ_k = gf.get_fill_nonlocal(sd[None], ek, mu[0], BETA, m0)
    n_ref, occ_ref, occ_k_ref = O.fill_nonlocal(sd[None], ek, mu[0], BETA, r0)
    assert abs(n - n_ref) < 1e-12 and rel(occ_k, occ_k_ref) < RTOL
    m1, _ = fit_smom(s1, BETA)
    assert abs(gf.get_total_fill(0.5, ek, s1, BETA, m1) - O.total_fill(0.5, ek, s1, BETA, m1)) < 1e-12
    assert abs(gf.update_mu(mu[0], n, ek, s1, BETA, m1) - mu[1]) < 1e-6


def test_pipeline_iterate_conserves_filling(K):
    """Two self-consistency iterations of a small model: after each mu update (mixing 1) the filling of
    the new self-energy with the local model is the target one."""
    from dgamore_amd import greens_function as gf
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline
    from dgamore_amd.self_energy import fit_smom

    say("pipeline iterate config 0 nk=(8,8,1) niw=4 niv=6, two iterations")
    m = S.make_model(0, nk=(8, 8, 1), niw=4, niv=6)
    pipe = SDEPipeline(m)
    n0 = pipe.update_occupation(pipe.sigma_dmft, m.mu)
    sigma, hist = pipe.iterate(2, n_target=n0)
    torch.cuda.synchronize()
    assert len(hist) == 3 and all(np.isfinite(hist))
    s = sigma.cpu().numpy()
    smom0, _ = O.fit_smom(s, m.beta)
    assert abs(O.total_fill(hist[-1], m.ek, s, m.beta, smom0) - n0) < 1e-5
    assert abs(gf.get_total_fill(hist[-1], m.ek, s, m.beta, fit_smom(s, m.beta)[0]) - n0) < 1e-5


@pytest.mark.parametrize("kind", ["spch", "sp"])
def test_lambda_corrected_step_matches_oracle(K, kind):
    """The step with the lambda correction of the physical susceptibility: lambda equals the
    oracle's root of the same sum rule, and Sigma equals the oracle chain with chi_lambda."""
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline
    from tests.conftest import oracle_kernel

    say(f"lambda-corrected SDE step ({kind}), config 0 model 8x8 niw=4 niv=6")
    m = S.make_model(0, nk=(8, 8, 1), niw=4, niv=6)
    p0 = SDEPipeline(m)
    p0.step(keep=True)
    torch.cuda.synchronize()
    counts, N = p0.grid.irrk_count, p0.N

    def full(x):
        half = x[:, 0, :].T.cpu().numpy()
        return np.concatenate([np.conj(half[:, :0:-1]), half], axis=1)

    chi_d, chi_m = (full(c) for c in p0.last["chi_phys"])
    s_d = (counts[:, None] * chi_d).sum() / (m.beta * N)
    s_m = (counts[:, None] * chi_m).sum() / (m.beta * N)
    # local sums with a root below the uncorrected lattice sums
    sums = {"dens": 0.9 * s_d * m.beta, "magn": 0.8 * s_m * m.beta}
    p1 = SDEPipeline(m, lambda_correction=kind, chi_loc_sums=sums)
    p1.step(keep=True)
    torch.cuda.synchronize()
    ref_l = {}
    if kind == "spch":
        ref_l["dens"] = O.find_lambda(chi_d, sums["dens"] / m.beta, counts, m.beta, N)
    target_m = sums["magn"] / m.beta if kind == "spch" else \
        (sums["dens"] + sums["magn"] - (counts[:, None] * chi_d).sum() / N) / m.beta
    ref_l["magn"] = O.find_lambda(chi_m, target_m, counts, m.beta, N)
    assert set(p1.lambdas) == set(ref_l)
    for ch, lam in ref_l.items():
        assert abs(p1.lambdas[ch] - lam) < 1e-10 * max(1.0, abs(lam))
    # the corrected magnetic susceptibility satisfies its sum rule
    chi_ml = full(p1.last["chi_phys"][1])
    assert abs((counts[:, None] * chi_ml).sum().real / (m.beta * N) - target_m.real) < 1e-6
    maps = O.kgrid_irr_maps(m.nk)
    g = m.giwk()
    kern = oracle_kernel(m, g, maps["irrk_ind"], ref_l)
    sig_ref = O.sigma_fft(kern, g, maps["irrk_inv"], m.beta, m.niv, m.niw)
    assert rel(p1.last["sigma_nl"], sig_ref) < 1e-9
    assert rel(p0.last["sigma_nl"], sig_ref) > 1e-6


def test_step_with_nonlocal_interaction_matches_oracle(K):
    """Extended Hubbard model: V(q) = V0 (cos qx + cos qy) enters the ladder (V_dens = 2V), the
    susceptibility, the kernel and the Fock term."""
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline
    from tests.conftest import oracle_kernel

    say("SDE step with non-local V(q), config 0 model 8x8 niw=4 niv=6")
    m = S.make_model(0, nk=(8, 8, 1), niw=4, niv=6)
    qx = 2 * np.pi * np.arange(8) / 8
    m.v = (0.6 * (np.cos(qx)[:, None] + np.cos(qx)[None, :])).reshape(8, 8, 1, 1, 1, 1, 1).astype(complex)
    p = SDEPipeline(m)
    p.step(keep=True)
    torch.cuda.synchronize()
    maps = O.kgrid_irr_maps(m.nk)
    g = m.giwk()
    sig_ref = O.sigma_fft(oracle_kernel(m, g, maps["irrk_ind"]), g, maps["irrk_inv"], m.beta, m.niv, m.niw)
    assert rel(p.last["sigma_nl"], sig_ref) < 1e-10
    occ_k = p.occ_k.cpu().numpy().reshape(8, 8, 1, 1, 1)
    h_ref, f_ref = O.hartree_fock(m.u, m.v, p.occ.cpu().numpy(), occ_k)
    assert rel(p.last["hartree"], h_ref) < 1e-12
    assert rel(p.last["fock"], f_ref.reshape(64, 1, 1)) < 1e-12
    m.v = None
    p0 = SDEPipeline(m)
    p0.step(keep=True)
    assert rel(p0.last["sigma_nl"], sig_ref) > 1e-6
