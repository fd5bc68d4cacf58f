"""The chemical-potential secant, the local moments and the occupations on the device, and the
self-consistency iteration captured as one HIP graph (needs an MI355X).

* dgami_fit_smom equals the oracle's moment fit;
* dgami_update_mu on the reference's end_2_end files reproduces mu_history.npy[1] (the secant on
  sigma_dga_iteration_1.npy from the filling of sigma_dmft.npy) and equals the host secant;
* a failing secant (an unreachable filling) keeps mu and reports nan;
* dgami_occupations equals the oracle's fill_nonlocal;
* iterate_device equals iterate (mu history and Sigma), and chained replays of the captured
  iteration equal the eager device iterations, with eager work in between.
"""
import sys

import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import GOLDEN, golden

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

BETA = 12.5


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def K():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import kernels

    return kernels


def _e2e_ek():
    hk, _ = O.read_hk(f"{GOLDEN}/e2e_wannier.hk")
    return hk.reshape(4, 4, 1, 2, 2)


def test_fit_smom_device(K):
    say("fit_smom on the device: sigma_dga_iteration_1 (16 k, 2 orbitals, niv = 100)")
    s1 = golden("e2e_sigma_dga_iteration_1.npy")
    m0, m1 = K.fit_smom(K.to_dev(s1.reshape(16, 2, 2, -1)), BETA)
    r0, r1 = O.fit_smom(s1, BETA)
    assert rel(m0, r0) < 1e-13 and rel(m1, r1) < 1e-13


def test_update_mu_device_reproduces_reference(K):
    from dgamore_amd import greens_function as gf
    from dgamore_amd.self_energy import fit_smom

    say("update_mu on the device: end_2_end mu_history from sigma_dga_iteration_1")
    ek = _e2e_ek()
    mu = golden("e2e_mu_history.npy").real
    sd = golden("e2e_sigma_dmft.npy")[0, 0, 0]
    s1 = golden("e2e_sigma_dga_iteration_1.npy")
    r0, _ = O.fit_smom(sd, BETA)
    n_target, _, _ = O.fill_nonlocal(sd[None], ek, mu[0], BETA, r0)
    e = K.to_dev(ek.reshape(16, 2, 2))
    sig = K.to_dev(s1.reshape(16, 2, 2, -1))
    m1, _ = K.fit_smom(sig, BETA)
    hloc = K.to_dev(ek.reshape(16, 2, 2).mean(axis=0))
    mu_dev = torch.tensor([mu[0]], dtype=torch.float64, device="cuda")
    info = torch.zeros(3, dtype=torch.float64, device="cuda")
    K.update_mu(mu_dev, n_target, e, sig, hloc, m1, BETA, 1.0, info)
    torch.cuda.synchronize()
    got = float(mu_dev.item())
    host = gf.update_mu(mu[0], n_target, ek, s1, BETA, fit_smom(s1, BETA)[0])
    say(f"  device mu {got:.12f}, host {host:.12f}, reference {mu[1]:.8f}; status {info.tolist()}")
    assert info[1].item() == 1.0
    assert abs(got - mu[1]) < 1e-6
    assert abs(got - host) < 1e-10
    # mixing: mu = a mu_new + (1 - a) mu_old
    mu_dev.fill_(mu[0])
    K.update_mu(mu_dev, n_target, e, sig, hloc, m1, BETA, 0.3, info)
    torch.cuda.synchronize()
    assert abs(float(mu_dev.item()) - (0.3 * got + 0.7 * mu[0])) < 1e-12


def test_update_mu_device_failure_keeps_mu(K):
    say("update_mu on the device: unreachable filling (5 electrons in 2 orbitals)")
    ek = _e2e_ek()
    s1 = golden("e2e_sigma_dga_iteration_1.npy")
    e = K.to_dev(ek.reshape(16, 2, 2))
    sig = K.to_dev(s1.reshape(16, 2, 2, -1))
    m1, _ = K.fit_smom(sig, BETA)
    hloc = K.to_dev(ek.reshape(16, 2, 2).mean(axis=0))
    mu_dev = torch.tensor([0.7], dtype=torch.float64, device="cuda")
    info = torch.zeros(3, dtype=torch.float64, device="cuda")
    K.update_mu(mu_dev, 5.0, e, sig, hloc, m1, BETA, 1.0, info)
    torch.cuda.synchronize()
    assert info[1].item() == 2.0 and np.isnan(info[0].item())
    assert float(mu_dev.item()) == 0.7


def test_occupations_device(K):
    say("occupations on the device vs the oracle's fill_nonlocal (end_2_end sigma_dga_iteration_1)")
    ek = _e2e_ek()
    s1 = golden("e2e_sigma_dga_iteration_1.npy")
    mu = float(golden("e2e_mu_history.npy").real[1])
    r0, _ = O.fit_smom(s1, BETA)
    n_ref, occ_ref, occ_k_ref = O.fill_nonlocal(s1, ek, mu, BETA, r0)
    sig = K.to_dev(s1.reshape(16, 2, 2, -1))
    m0, _ = K.fit_smom(sig, BETA)
    occ_k = torch.empty((16, 2, 2), dtype=torch.complex128, device="cuda")
    occ = torch.empty((2, 2), dtype=torch.complex128, device="cuda")
    K.occupations(K.to_dev(ek.reshape(16, 2, 2)), sig, m0, torch.tensor([mu], dtype=torch.float64, device="cuda"),
                  BETA, occ_k, occ)
    torch.cuda.synchronize()
    assert rel(occ_k, occ_k_ref.reshape(16, 2, 2)) < 1e-12
    assert rel(occ, occ_ref) < 1e-12
    assert abs(2 * np.trace(occ.cpu().numpy()).real - n_ref) < 1e-12


@pytest.mark.parametrize("nb_cfg", [0, 4])
def test_device_iteration_and_its_graph(K, nb_cfg):
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    nk = (8, 8, 1) if nb_cfg == 0 else (6, 6, 1)
    say(f"self-consistency loop on the device: config {nb_cfg} form, nk={nk}, niw=4, niv=6, mixing 0.5")
    m = S.make_model(nb_cfg, nk=nk, niw=4, niv=6)
    host = SDEPipeline(m, mixing=0.5)
    n0 = host.update_occupation(host.sigma_dmft, m.mu)
    s_host, h_host = host.iterate(3, n_target=n0)
    dev = SDEPipeline(m, mixing=0.5)
    s_dev, h_dev = dev.iterate_device(3, n_target=n0)
    torch.cuda.synchronize()
    say(f"  mu host {np.round(h_host, 10).tolist()} / device {np.round(h_dev, 10).tolist()}")
    assert np.abs(np.array(h_host) - np.array(h_dev)).max() < 1e-9
    assert rel(s_dev, s_host.cpu().numpy()) < 1e-9

    g = SDEPipeline(m, mixing=0.5)
    g.capture_iteration(n_target=n0)
    eager = SDEPipeline(m, mixing=0.5)
    eager.device_loop_state(n_target=n0)
    sigma = eager.graph_sigma = eager.sigma_dmft.expand(eager.N, m.nb, m.nb, 2 * eager.niv_s).contiguous().clone()
    for it in range(3):
        out = g.replay()
        sigma = eager.device_iteration(sigma)
        torch.cuda.synchronize()
        e_mu, g_mu = float(eager.mu_dev.item()), float(g.mu_dev.item())
        say(f"  replay {it}: mu graph {g_mu:.12f} eager {e_mu:.12f}")
        assert abs(e_mu - g_mu) < 1e-12
        assert rel(out, sigma.cpu().numpy()) < 1e-12
        assert rel(g.occ_k, eager.occ_k.cpu().numpy()) < 1e-12
