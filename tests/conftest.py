import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")


def golden(name):
    """Load a reference-held fixture copied by tests/golden/make_golden.py (complex64 -> complex128)."""
    a = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    return a.astype(np.complex128) if np.iscomplexobj(a) else a


@pytest.fixture(scope="session")
def e2e_inputs():
    """The reference's end_2_end model: 2 orbitals, 4x4x1, beta=12.5, niw=20, niv_core=20, niv_shell=10."""
    from oracle import dga_oracle as O

    hk, _ = O.read_hk(os.path.join(GOLDEN, "e2e_wannier.hk"))
    ek = hk.reshape(4, 4, 1, 2, 2)
    sigma = golden("e2e_sigma_dmft.npy")
    mu = float(golden("e2e_mu_history.npy")[0])
    beta = 12.5
    giwk = O.g_full(sigma, mu, ek, beta)
    u = np.zeros((2, 2, 2, 2))
    u[0, 0, 0, 0] = u[1, 1, 1, 1] = 8.0  # test_local_sde_end_to_end.py:106-113
    return {"ek": ek, "sigma": sigma, "mu": mu, "beta": beta, "giwk": giwk, "u": u,
            "nk": (4, 4, 1), "niw": 20, "niv_core": 20, "niv_shell": 10}


def ladder_case(nb, niv, niw, nq, seed, gamma=None, beta=10.0, with_v=True):
    """Random but well-conditioned ladder inputs in the reference layouts:
    (chi0_inv [q,o1..o4,w,v], gamma [o1..o4,w,v,v'], u [nb]*4, v [q,nb,nb,nb,nb], beta)."""
    from oracle import dga_oracle as O

    rng = np.random.default_rng(seed)
    nw = niw + 1
    nv = 2 * niv
    vv = O.vn(niv, beta)
    # chi0 blocks ~ diagonal-dominant with frequency decay
    base = np.einsum("ad,bc->abcd", np.eye(nb), np.eye(nb))
    chi0 = np.empty((nq, nb, nb, nb, nb, nw, nv), dtype=complex)
    for iq in range(nq):
        for w in range(nw):
            pert = 0.2 * (rng.standard_normal((nb,) * 4 + (nv,)) + 1j * rng.standard_normal((nb,) * 4 + (nv,)))
            chi0[iq, ..., w, :] = (base[..., None] + pert) * (beta / (1.0 + vv**2))
    chi0_inv = O.block_invert_vn1(chi0)
    if gamma is None:
        gamma = rng.standard_normal((nb,) * 4 + (nw, nv, nv)) + 1j * rng.standard_normal((nb,) * 4 + (nw, nv, nv))
        gamma *= 2.0
    u = rng.standard_normal((nb,) * 4)
    v = 0.3 * (rng.standard_normal((nq,) + (nb,) * 4) + 1j * rng.standard_normal((nq,) + (nb,) * 4)) if with_v \
        else np.zeros((nq,) + (nb,) * 4)
    return chi0_inv, gamma, u, v, beta


def oracle_kernel(m, g, irrk_ind, lambdas=None):
    """K(q) = -K_dc + K_dens + 3 K_magn of the listed q-points by the oracle (the chain of
    nonlocal_sde.py:777-857); lambdas: {channel: lambda} applied to the physical susceptibility;
    m.v (if set): the non-local V(q) on the full grid."""
    from oracle import dga_oracle as O

    nq, nb = len(irrk_ind), m.nb
    v_q = None if getattr(m, "v", None) is None else m.v.reshape((-1,) + (nb,) * 4)[irrk_ind]
    chi0 = O.bubble_fft(g, m.niw, m.niv, irrk_ind, m.beta)
    inv = O.block_invert_vn1(chi0)
    kern = -O.dc_kernel(m.f_magn, chi0, m.u, m.beta, m.niv)
    zero_v = np.zeros((nq,) + (nb,) * 4)
    for ch, wgt in (("dens", 1.0), ("magn", 3.0)):
        ur = O.u_as_channel(m.u, ch)
        vr = zero_v if v_q is None else O.v_as_channel(v_q, ch)
        aux = O.ladder_sum_solve(inv, m.gamma[ch], ur, vr, m.beta)
        vrg = O.vrg(inv, aux, m.beta)
        chi = O.chi_phys_shell(O.sum_over_vn(aux, m.beta), chi0.sum(-1) / m.beta**2, chi0.sum(-1) / m.beta**2, ur,
                               vr)
        if lambdas and ch in lambdas:
            chi = O.apply_lambda(chi, lambdas[ch])
        kern = kern + wgt * O.kernel_r(vrg, chi, ur, vr, ch)
    return kern
