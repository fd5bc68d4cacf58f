"""Parity of every libdgami entry point with the CPU oracle (needs an MI355X).

Each test prints the case it runs before touching the GPU. Tolerances: the
plain path is complex128 end to end and is held to 1e-10 relative (max-abs
error over max-abs value) against the oracle; agreement with the reference's
own complex64 fixture files is held to 1e-6.
"""
import sys

import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import golden, ladder_case as _ladder_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

RTOL = 1e-10


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


# ----------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("nk", [(16, 16, 1), (24, 24, 1), (4, 4, 4), (12, 20, 1), (64, 64, 1)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_grid(K, nk, sign):
    say(f"fft_grid nk={nk} sign={sign}")
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5,) + nk) + 1j * rng.standard_normal((5,) + nk)
    ref = np.fft.fftn(x, axes=(1, 2, 3)) if sign < 0 else np.fft.ifftn(x, axes=(1, 2, 3)) * np.prod(nk)
    d = K.to_dev(x.reshape(5, -1))
    out = K.fft_grid(d, nk, sign, 1.0)
    torch.cuda.synchronize()
    assert rel(out.reshape((5,) + nk), ref) < 1e-13


# ----------------------------------------------------------------------------- bubble
def test_bubble_e2e(K, e2e_inputs):
    d = e2e_inputs
    say("bubble end_2_end 4x4x1 nb=2 niw=20 niv=30")
    maps = O.kgrid_irr_maps(d["nk"])
    niv = d["niv_core"] + d["niv_shell"]
    ref = O.bubble_fft(d["giwk"], d["niw"], niv, maps["irrk_ind"], d["beta"])
    chi0 = K.bubble_irrq(K.to_dev(d["giwk"]), d["nk"], d["niw"], niv, d["beta"],
                         K.to_dev(maps["irrk_ind"], torch.int32))
    got = K.qf_to_ref4(chi0, 2)
    torch.cuda.synchronize()
    assert rel(got, ref) < RTOL
    assert rel(got, golden("e2e_gchi0_q.npy")) < 1e-6


def test_bubble_synthetic_1band(K):
    from dgamore_amd import synthetic as S

    say("bubble synthetic 1 band 16x16 niw=6 niv=8")
    m = S.make_model(0, niw=6, niv=8, with_vertices=False)
    g = m.giwk()
    maps = O.kgrid_irr_maps(m.nk)
    ref = O.bubble_fft(g, 6, 8, maps["irrk_ind"], m.beta)
    chi0 = K.bubble_irrq(K.to_dev(g), m.nk, 6, 8, m.beta, K.to_dev(maps["irrk_ind"], torch.int32))
    assert rel(K.qf_to_ref4(chi0, 1), ref) < RTOL


# ----------------------------------------------------------------------------- chi0 prepare
def test_chi0_prepare(K):
    say("chi0_prepare end_2_end fixture nb=2")
    chi0 = golden("e2e_gchi0_q.npy")
    beta = 12.5
    inv, full, core = K.chi0_prepare(K.ref4_to_qf(K.to_dev(chi0)), 2, 20, beta)
    assert rel(K.qf_to_ref4(inv, 2), O.block_invert_vn1(chi0[..., 10:50])) < RTOL
    assert rel(K.qw_to_ref4w(full, 2), chi0.sum(-1) / beta**2) < RTOL
    assert rel(K.qw_to_ref4w(core, 2), chi0[..., 10:50].sum(-1) / beta**2) < RTOL


# ----------------------------------------------------------------------------- ladder
def _run_ladder(K, chi0_inv, gammas, us, vs, beta, precision=0):
    nb = chi0_inv.shape[1]
    g = torch.stack([K.vertex_to_cmp(K.to_dev(x)) for x in gammas])
    u = torch.stack([K.orb_to_m(K.to_dev(x)) for x in us])
    v = torch.stack([K.orb_to_m(K.to_dev(x)) for x in vs])
    out, info = K.ladder_solve_sum(K.ref4_to_qf(K.to_dev(chi0_inv)), g, u, v, beta, precision)
    return [K.qf_to_ref4(out[c], nb) for c in range(len(gammas))], info


@pytest.mark.parametrize("nb,niv,niw,nq", [
    (1, 20, 4, 7),    # n = 40 (register kernel by default)
    (1, 60, 2, 3),    # n = 120 (blocked LU by default; 1-band GPU-scale box)
    (2, 8, 3, 5),     # n = 32, 4 right-hand sides
    (2, 15, 2, 4),    # n = 60
    (3, 4, 2, 3),     # n = 72, 9 right-hand sides
    (2, 20, 2, 3),    # n = 160 (beyond the register kernels: blocked LU for every family)
    (3, 12, 1, 2),    # n = 216
])
@pytest.mark.parametrize("family", ["auto", "col", "reg", "blocked", "blocked_vector"])
def test_ladder_plain_fp64(K, nb, niv, niw, nq, family, monkeypatch):
    say(f"ladder fp64 nb={nb} niv={niv} niw={niw} nq={nq} n={nb*nb*2*niv} kernel={family}")
    if family == "blocked_vector":
        monkeypatch.setenv("DGAMI_LADDER_KERNEL", "blocked")
        monkeypatch.setenv("DGAMI_BLOCKED_GEMM", "vector")
    elif family != "auto":
        monkeypatch.setenv("DGAMI_LADDER_KERNEL", family)
    chi0_inv, gamma, u, v, beta = _ladder_case(nb, niv, niw, nq, seed=nb * 100 + niv)
    ref = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)
    got, info = _run_ladder(K, chi0_inv, [gamma], [u], [v], beta)
    assert int(info.abs().sum()) == 0
    assert rel(got[0], ref) < RTOL


@pytest.mark.parametrize("nb,niv,niw,nq", [
    (3, 40, 1, 2),    # n = 720, the 3-orbital box (panel of one thread per row)
    (3, 48, 1, 1),    # n = 864 (panel of three rows per thread)
    (1, 500, 1, 1),   # n = 1000, one right-hand side
])
def test_ladder_plain_fp64_large(K, nb, niv, niw, nq):
    say(f"ladder fp64 large nb={nb} niv={niv} niw={niw} nq={nq} n={nb*nb*2*niv}")
    chi0_inv, gamma, u, v, beta = _ladder_case(nb, niv, niw, nq, seed=nb * 100 + niv)
    ref = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)
    got, info = _run_ladder(K, chi0_inv, [gamma], [u], [v], beta)
    assert int(info.abs().sum()) == 0
    assert rel(got[0], ref) < RTOL


def test_ladder_multi_models(K):
    """Five models of different orbital counts, boxes and grids in one call: four share one
    register-kernel launch, the n = 120 one goes through the blocked LU."""
    say("ladder_solve_sum_multi: n = 16, 30, 64, 54 (one launch) and 120 (blocked)")
    cases = [(1, 8, 2, 3), (1, 15, 1, 4), (2, 8, 2, 2), (3, 3, 1, 2), (2, 15, 1, 2)]
    models, refs = [], []
    for idx, (nb, niv, niw, nq) in enumerate(cases):
        chi0_inv, gamma, u, v, beta = _ladder_case(nb, niv, niw, nq, seed=50 + idx)
        rng = np.random.default_rng(idx)
        gamma2 = gamma + 0.5 * rng.standard_normal(gamma.shape)
        refs.append([O.ladder_sum_solve(chi0_inv, g, u, v, beta) for g in (gamma, gamma2)])
        models.append((K.ref4_to_qf(K.to_dev(chi0_inv)),
                       torch.stack([K.vertex_to_cmp(K.to_dev(g)) for g in (gamma, gamma2)]).contiguous(),
                       torch.stack([K.orb_to_m(K.to_dev(u))] * 2).contiguous(),
                       torch.stack([K.orb_to_m(K.to_dev(v))] * 2).contiguous(), beta))
    outs = K.ladder_solve_sum_multi(models)
    torch.cuda.synchronize()
    for (nb, *_), (out, info), ref in zip(cases, outs, refs):
        assert int(info.abs().sum()) == 0
        for c in range(2):
            assert rel(K.qf_to_ref4(out[c], nb), ref[c]) < RTOL


# The mixed form (complex64 LU, FP64 iterative refinement) must agree with the FP64 solve
# to 1e-8 relative error, the accuracy stated for it.
MIXED_RTOL = 1e-8


@pytest.mark.parametrize("nb,niv,niw,nq", [
    (1, 20, 3, 5),    # n = 40
    (1, 60, 2, 3),    # n = 120
    (2, 15, 2, 4),    # n = 60, 4 right-hand sides
    (3, 10, 2, 3),    # n = 180, 9 right-hand sides
    (3, 40, 1, 2),    # n = 720, the 3-orbital box
])
@pytest.mark.parametrize("gemm", ["mfma", "vector"])
def test_ladder_mixed_fp32_ir(K, nb, niv, niw, nq, gemm, monkeypatch):
    say(f"ladder mixed fp32+IR nb={nb} niv={niv} niw={niw} nq={nq} n={nb*nb*2*niv} trailing update={gemm}")
    if gemm == "vector":
        monkeypatch.setenv("DGAMI_BLOCKED_GEMM", "vector")
    chi0_inv, gamma, u, v, beta = _ladder_case(nb, niv, niw, nq, seed=nb * 100 + niv + 1)
    ref = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)
    got, info = _run_ladder(K, chi0_inv, [gamma], [u], [v], beta, precision=1)
    assert int(info.abs().sum()) == 0
    assert rel(got[0], ref) < MIXED_RTOL


def test_ladder_mixed_two_channels_realistic_vertex(K):
    say("ladder mixed srvo3 Gamma dens+magn nb=3 niv=10 n=180")
    gd = golden("srvo3_gamma_dens_loc.npy")[:, :, :, :, :3]
    gm = golden("srvo3_gamma_magn_loc.npy")[:, :, :, :, :3]
    chi0_inv, _, u, v, beta = _ladder_case(3, 10, 2, 3, seed=7, gamma=gd)
    ud, um = O.u_as_channel(u, "dens"), O.u_as_channel(u, "magn")
    vd, vm = O.v_as_channel(v, "dens"), O.v_as_channel(v, "magn")
    got, _ = _run_ladder(K, chi0_inv, [gd, gm], [ud, um], [vd, vm], beta, precision=1)
    assert rel(got[0], O.ladder_sum_solve(chi0_inv, gd, ud, vd, beta)) < MIXED_RTOL
    assert rel(got[1], O.ladder_sum_solve(chi0_inv, gm, um, vm, beta)) < MIXED_RTOL


@pytest.mark.parametrize("nb,niv,niw,nq", [
    (1, 60, 2, 3),    # n = 120
    (3, 10, 2, 3),    # n = 180
    (3, 40, 1, 2),    # n = 720, the 3-orbital box
])
def test_ladder_mixed_split_bf16_ir(K, nb, niv, niw, nq):
    say(f"ladder mixed split-bf16+IR nb={nb} niv={niv} niw={niw} nq={nq} n={nb*nb*2*niv}")
    chi0_inv, gamma, u, v, beta = _ladder_case(nb, niv, niw, nq, seed=nb * 100 + niv + 2)
    ref = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)
    got, info = _run_ladder(K, chi0_inv, [gamma], [u], [v], beta, precision=2)
    assert int(info.abs().sum()) == 0
    assert rel(got[0], ref) < MIXED_RTOL


@pytest.mark.parametrize("niv,nq", [(2, 3), (6, 2)])   # n = 64 (register kernel), n = 192 (blocked LU)
@pytest.mark.parametrize("precision", [0, 1, 2])
def test_ladder_four_orbitals_every_form(K, niv, nq, precision):
    """nb = 4, the largest orbital count: 16 right-hand sides per problem, in the register
    kernel and in every form of the blocked LU."""
    nb, niw = 4, 1
    say(f"ladder nb=4 niv={niv} niw={niw} nq={nq} n={nb*nb*2*niv} precision={precision}")
    chi0_inv, gamma, u, v, beta = _ladder_case(nb, niv, niw, nq, seed=400 + niv)
    ref = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)
    got, info = _run_ladder(K, chi0_inv, [gamma], [u], [v], beta, precision=precision)
    assert int(info.abs().sum()) == 0
    assert rel(got[0], ref) < (RTOL if precision == 0 else MIXED_RTOL)


def test_ladder_unknown_precision_is_refused(K):
    say("ladder precision 7 -> error")
    chi0_inv, gamma, u, v, beta = _ladder_case(1, 4, 1, 2, seed=5)
    with pytest.raises(ValueError):
        _run_ladder(K, chi0_inv, [gamma], [u], [v], beta, precision=7)


def test_ladder_two_channels_realistic_vertex(K):
    """SrVO3's local irreducible vertices (reference fixture, 3 orbitals, n=180) with synthetic chi0."""
    say("ladder fp64 srvo3 Gamma dens+magn nb=3 niv=10 n=180")
    gd = golden("srvo3_gamma_dens_loc.npy")[:, :, :, :, :3]
    gm = golden("srvo3_gamma_magn_loc.npy")[:, :, :, :, :3]
    chi0_inv, _, u, v, beta = _ladder_case(3, 10, 2, 3, seed=7, gamma=gd)
    ud, um = O.u_as_channel(u, "dens"), O.u_as_channel(u, "magn")
    vd, vm = O.v_as_channel(v, "dens"), O.v_as_channel(v, "magn")
    got, _ = _run_ladder(K, chi0_inv, [gd, gm], [ud, um], [vd, vm], beta)
    assert rel(got[0], O.ladder_sum_solve(chi0_inv, gd, ud, vd, beta)) < RTOL
    assert rel(got[1], O.ladder_sum_solve(chi0_inv, gm, um, vm, beta)) < RTOL


@pytest.mark.parametrize("precision", [0, 2])
def test_ladder_largest_compound_size(K, precision):
    """n = 1536, the largest size of the blocked LU (2 orbitals, niv = 192), one (q, w) problem."""
    say(f"ladder n=1536 (maximum), precision {precision}")
    chi0_inv, gamma, u, v, beta = _ladder_case(2, 192, 0, 1, seed=1536)
    ref = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)
    got, info = _run_ladder(K, chi0_inv, [gamma], [u], [v], beta, precision=precision)
    assert int(info.abs().sum()) == 0
    assert rel(got[0], ref) < (RTOL if precision == 0 else 1e-8)


def test_ladder_above_largest_size_is_refused(K):
    say("ladder n=1552 (above the maximum) -> ValueError, no launch")
    nb, niv = 2, 194
    n = nb * nb * 2 * niv
    chi0_inv = torch.zeros((1, nb**4, 2 * niv, 1), dtype=torch.complex128, device="cuda")
    gamma = torch.zeros((1, 1, n, n), dtype=torch.complex128, device="cuda")
    u = torch.zeros((1, nb**4), dtype=torch.complex128, device="cuda")
    for precision in (0, 2):
        with pytest.raises(ValueError):
            K.ladder_solve_sum(chi0_inv, gamma, u, None, 10.0, precision)


def test_ladder_four_orbitals(K):
    """nb = 4 (16 right-hand sides), the most orbitals the kernels take, in both forms."""
    say("ladder nb=4 niv=4 (n=128, 16 right-hand sides), plain and split-BF16")
    chi0_inv, gamma, u, v, beta = _ladder_case(4, 4, 1, 2, seed=44)
    ref = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)
    for precision, tol in ((0, RTOL), (2, 1e-8)):
        got, info = _run_ladder(K, chi0_inv, [gamma], [u], [v], beta, precision=precision)
        assert int(info.abs().sum()) == 0
        assert rel(got[0], ref) < tol


def test_ladder_five_orbitals_refused(K):
    say("ladder nb=5 -> ValueError")
    nb, niv = 5, 1
    n = nb * nb * 2 * niv
    chi0_inv = torch.zeros((1, nb**4, 2 * niv, 1), dtype=torch.complex128, device="cuda")
    gamma = torch.zeros((1, 1, n, n), dtype=torch.complex128, device="cuda")
    u = torch.zeros((1, nb**4), dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError):
        K.ladder_solve_sum(chi0_inv, gamma, u, None, 10.0)


def test_ladder_singular_raises(K):
    say("ladder singular matrix -> LinAlgError")
    nb, niv, nw, nq = 1, 4, 1, 2
    chi0_inv = np.zeros((nq, 1, 1, 1, 1, nw, 2 * niv), dtype=complex)
    gamma = np.zeros((1, 1, 1, 1, nw, 2 * niv, 2 * niv), dtype=complex)
    u = np.ones((1, 1, 1, 1))
    with pytest.raises(np.linalg.LinAlgError):
        _run_ladder(K, chi0_inv, [gamma], [u], [np.zeros((nq, 1, 1, 1, 1))], 10.0)


# ----------------------------------------------------------------------------- vrg / chi_phys / kernel
@pytest.mark.parametrize("nb", [1, 2, 3])
def test_vrg_chi_phys_kernel(K, nb):
    say(f"vrg/chi_phys/kernel nb={nb}")
    rng = np.random.default_rng(nb)
    nq, nw, niv, beta = 5, 3, 4, 7.0
    chi0_inv, _, u, v, _ = _ladder_case(nb, niv, nw - 1, nq, seed=nb)
    aux = (rng.standard_normal(chi0_inv.shape) + 1j * rng.standard_normal(chi0_inv.shape)) * 0.3
    full = rng.standard_normal(chi0_inv.shape[:-1]) + 1j * rng.standard_normal(chi0_inv.shape[:-1])
    core = 0.5 * (rng.standard_normal(chi0_inv.shape[:-1]) + 1j * rng.standard_normal(chi0_inv.shape[:-1]))
    for ch in ("dens", "magn"):
        ur, vr = O.u_as_channel(u, ch), O.v_as_channel(v, ch)
        vrg_ref = O.vrg(chi0_inv, aux, beta)
        chi_ref = O.chi_phys_shell(O.sum_over_vn(aux, beta), full, core, ur, vr)
        ker_ref = O.kernel_r(vrg_ref, chi_ref, ur, vr, ch)
        d_inv, d_aux = K.ref4_to_qf(K.to_dev(chi0_inv)), K.ref4_to_qf(K.to_dev(aux))
        d_full, d_core = K.ref4w_to_qw(K.to_dev(full)), K.ref4w_to_qw(K.to_dev(core))
        d_u, d_v = K.orb_to_m(K.to_dev(ur)), K.orb_to_m(K.to_dev(vr))
        vrg, chi = K.vrg_chi_phys(d_inv, d_aux, d_full, d_core, d_u, d_v, nb, beta)
        assert rel(K.qf_to_ref4(vrg, nb), vrg_ref) < RTOL
        assert rel(K.qw_to_ref4w(chi, nb), chi_ref) < RTOL
        ker = K.kernel_r(vrg, chi, d_u, d_v, nb, ch, 3.0)
        assert rel(K.qf_to_ref4(ker, nb), 3.0 * ker_ref) < RTOL


# ----------------------------------------------------------------------------- DC kernel
@pytest.mark.parametrize("nb,niv_core,niv_shell", [(1, 6, 0), (2, 4, 2), (3, 3, 0)])
def test_dc_kernel(K, nb, niv_core, niv_shell):
    say(f"dc_kernel nb={nb} niv_core={niv_core} niv_shell={niv_shell}")
    rng = np.random.default_rng(11)
    nq, nw, beta = 4, 3, 9.0
    nvf = 2 * (niv_core + niv_shell)
    chi0 = rng.standard_normal((nq,) + (nb,) * 4 + (nw, nvf)) + 1j * rng.standard_normal((nq,) + (nb,) * 4 + (nw, nvf))
    f = rng.standard_normal((nb,) * 4 + (nw, nvf, nvf)) + 1j * rng.standard_normal((nb,) * 4 + (nw, nvf, nvf))
    u = rng.standard_normal((nb,) * 4)
    ref = -O.dc_kernel(f, chi0, u, beta, niv_core)
    got = K.dc_kernel(K.ref4_to_qf(K.to_dev(chi0)), K.to_dev(u), K.to_dev(f), nb, niv_core, beta)
    assert rel(K.qf_to_ref4(got, nb), ref) < RTOL


# ----------------------------------------------------------------------------- Sigma
@pytest.mark.parametrize("nb,nk,niw,niv", [(1, (16, 16, 1), 5, 6), (2, (4, 4, 1), 3, 4), (3, (6, 6, 1), 2, 3)])
def test_sigma_from_kernel(K, nb, nk, niw, niv):
    say(f"sigma nb={nb} nk={nk} niw={niw} niv={niv}")
    rng = np.random.default_rng(5)
    maps = O.kgrid_irr_maps(nk)
    nq = len(maps["irrk_ind"])
    kern = rng.standard_normal((nq,) + (nb,) * 4 + (niw + 1, 2 * niv)) \
        + 1j * rng.standard_normal((nq,) + (nb,) * 4 + (niw + 1, 2 * niv))
    niv_g = niw + niv + 3
    g = rng.standard_normal(nk + (nb, nb, 2 * niv_g)) + 1j * rng.standard_normal(nk + (nb, nb, 2 * niv_g))
    beta = 6.0
    ref = O.sigma_fft(kern, g, maps["irrk_inv"], beta, niv, niw)
    got = K.sigma_from_kernel(K.ref4_to_qf(K.to_dev(kern)), K.to_dev(maps["irrk_inv"], torch.int32),
                              K.to_dev(g), nk, niw, beta)
    assert rel(got, ref) < RTOL


# ----------------------------------------------------------------------------- Eliashberg
@pytest.mark.parametrize("channel", ["sing", "trip"])
def test_pairing_matvec_e2e(K, channel):
    say(f"pairing matvec end_2_end fixture {channel}")
    gam = golden(f"e2e_gamma_irrq_{channel}_pp.npy")
    giwk = golden("e2e_giwk_dga.npy")
    nk, beta = (4, 4, 1), 12.5
    maps = O.kgrid_irr_maps(nk)
    np2 = gam.shape[-1]
    chi0pp = O.chi0_pp_w0(giwk, np2 // 2)  # [N, nb^4 (abcd), np2]
    ghat, gflip, s = O.pairing_setup(gam, maps["irrk_inv"], nk, channel)
    rng = np.random.default_rng(3)
    gaps = rng.standard_normal((3,) + nk + (2, 2, np2)) + 1j * rng.standard_normal((3,) + nk + (2, 2, np2))
    chi0pp_k = chi0pp.reshape(nk + (2, 2, 2, 2, np2))
    ref = np.stack([O.pairing_matvec(ghat, gflip, chi0pp_k, gp, s, beta) for gp in gaps])
    d_hat, d_flip = K.pairing_setup(K.to_dev(gam), K.to_dev(maps["irrk_inv"], torch.int32), nk, channel)
    got = K.pairing_matvec(d_hat, d_flip, K.to_dev(chi0pp.reshape(16, 16, np2)), K.to_dev(gaps), nk, channel, beta)
    assert rel(got, ref) < RTOL


def test_chi0_pp(K):
    say("chi0_pp end_2_end giwk_dga niv_pp=10")
    giwk = golden("e2e_giwk_dga.npy")
    ref = O.chi0_pp_w0(giwk, 10)                    # [N, nb, nb, nb, nb, 2np]
    got = K.chi0_pp(K.to_dev(giwk), 10)
    torch.cuda.synchronize()
    assert rel(got.reshape(ref.shape), ref) < 1e-15


# complex64 storage of the transformed vertex: every element rounded once (2^-24 relative), sums in
# complex128, so the product keeps ~1e-7 relative accuracy.
C64_STORAGE_RTOL = 1e-6


@pytest.mark.parametrize("channel", ["sing", "trip"])
def test_pairing_matvec_c64_storage(K, channel):
    say(f"pairing matvec complex64 storage end_2_end fixture {channel}")
    gam = golden(f"e2e_gamma_irrq_{channel}_pp.npy")
    giwk = golden("e2e_giwk_dga.npy")
    nk, beta = (4, 4, 1), 12.5
    maps = O.kgrid_irr_maps(nk)
    np2 = gam.shape[-1]
    chi0pp = O.chi0_pp_w0(giwk, np2 // 2)
    ghat, gflip, s = O.pairing_setup(gam, maps["irrk_inv"], nk, channel)
    rng = np.random.default_rng(4)
    gaps = rng.standard_normal((2,) + nk + (2, 2, np2)) + 1j * rng.standard_normal((2,) + nk + (2, 2, np2))
    chi0pp_k = chi0pp.reshape(nk + (2, 2, 2, 2, np2))
    ref = np.stack([O.pairing_matvec(ghat, gflip, chi0pp_k, gp, s, beta) for gp in gaps])
    d_hat, d_flip = K.pairing_setup(K.to_dev(gam), K.to_dev(maps["irrk_inv"], torch.int32), nk, channel,
                                    torch.complex64)
    assert d_hat.dtype == torch.complex64
    got = K.pairing_matvec(d_hat, d_flip, K.to_dev(chi0pp.reshape(16, 16, np2)), K.to_dev(gaps), nk, channel, beta)
    err = rel(got, ref)
    assert 1e-12 < err < C64_STORAGE_RTOL  # rounded, but only once


# ----------------------------------------------------------------------------- small helpers
def test_g_full(K):
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    say("g_full 3-orbital 6x6 local sigma")
    m = S.make_model(2, nk=(6, 6, 1), niw=2, niv=4, with_vertices=False)
    ek = K.to_dev(m.ek.reshape(36, 3, 3))
    sig = K.to_dev(m.sigma_loc[None])
    niv = 9
    g = torch.empty((36, 3, 3, 2 * niv), dtype=torch.complex128, device=ek.device)
    from dgamore_amd import _lib
    _lib.call("dgami_g_full", 36, 3, niv, float(m.beta), float(m.mu), K._p(ek), K._p(sig), 1,
              m.sigma_loc.shape[-1] // 2, K._p(g), K._stream())
    ns = m.sigma_loc.shape[-1] // 2
    ref = O.g_full(m.sigma_loc[None, None, None, ..., ns - niv: ns + niv], m.mu, m.ek, m.beta)
    assert rel(g.reshape(ref.shape), ref) < RTOL


@pytest.mark.parametrize("nb", [1, 2, 3])
def test_block_invert(K, nb):
    from dgamore_amd import _lib

    say(f"block_invert nb={nb}")
    rng = np.random.default_rng(nb)
    x = rng.standard_normal((7,) + (nb,) * 4 + (3,)) + 1j * rng.standard_normal((7,) + (nb,) * 4 + (3,))
    x += 3 * np.einsum("ad,bc->abcd", np.eye(nb), np.eye(nb))[None, ..., None]
    d = K.ref4w_to_qw(K.to_dev(x))  # [w][nb^4][q] == element-major over nmat = (w, q)
    flat = d.permute(1, 0, 2).reshape(nb**4, -1).contiguous()
    out = torch.empty_like(flat)
    _lib.call("dgami_block_invert", flat.shape[1], nb, K._p(flat), K._p(out), None, K._stream())
    got = out.reshape(nb**4, 3, 7).permute(1, 0, 2).contiguous()
    assert rel(K.qw_to_ref4w(got, nb), O.block_invert_vn0(x)) < RTOL


@pytest.mark.parametrize("nb,niv", [(1, 20), (2, 12), (3, 10)])
def test_solve_sum_dense(K, nb, niv):
    from dgamore_amd import _lib

    say(f"solve_sum_dense nb={nb} niv={niv}")
    chi0_inv, gamma, u, v, beta = _ladder_case(nb, niv, 1, 3, seed=3)
    mats = np.stack([O.ladder_matrix(chi0_inv[q], gamma, u, v[q], beta)[w] for q in range(3) for w in range(2)])
    ref = O.ladder_sum_solve(chi0_inv, gamma, u, v, beta)  # [q,o1,o2,o3,o4,w,v]
    n = mats.shape[-1]
    d = K.to_dev(mats)
    out = torch.empty((6, n, nb * nb), dtype=torch.complex128, device=d.device)
    info = torch.zeros(6, dtype=torch.int32, device=d.device)
    _lib.call("dgami_solve_sum_dense", 6, nb, niv, float(beta), K._p(d), 0, K._p(out), K._p(info, torch.int32),
              K._stream())
    got = out.reshape(3, 2, nb, nb, 2 * niv, nb, nb).cpu().numpy()  # [q,w,o1,o2,v,o4,o3]
    got = got.transpose(0, 2, 3, 6, 5, 1, 4)
    assert rel(got, ref) < RTOL
    _lib.call("dgami_solve_sum_dense", 6, nb, niv, float(beta), K._p(d), 1, K._p(out), K._p(info, torch.int32),
              K._stream())
    got = out.reshape(3, 2, nb, nb, 2 * niv, nb, nb).cpu().numpy().transpose(0, 2, 3, 6, 5, 1, 4)
    assert int(info.abs().sum()) == 0
    assert rel(got, ref) < 1e-8


# ----------------------------------------------------------------------------- whole SDE iteration
@pytest.mark.parametrize("cfg,nk,niw,niv", [(0, (8, 8, 1), 4, 6), (2, (6, 6, 1), 2, 4),
                                             (0, None, None, None)])  # the last: config 0 at its full size
def test_pipeline_step_matches_oracle(K, cfg, nk, niw, niv):
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline, u_as_channel

    say(f"pipeline step config {cfg} nk={nk} niw={niw} niv={niv}")
    m = S.make_model(cfg, nk=nk, niw=niw, niv=niv)
    pipe = SDEPipeline(m)
    sig = pipe.step(keep=True)
    torch.cuda.synchronize()
    g = m.giwk()
    maps = O.kgrid_irr_maps(m.nk)
    nq = len(maps["irrk_ind"])
    chi0 = O.bubble_fft(g, m.niw, m.niv, maps["irrk_ind"], m.beta)
    inv = O.block_invert_vn1(chi0)
    kern = -O.dc_kernel(m.f_magn, chi0, m.u, m.beta, m.niv)
    zero_v = np.zeros((nq,) + (m.nb,) * 4)
    for ch, wgt in (("dens", 1.0), ("magn", 3.0)):
        ur = u_as_channel(m.u, ch)
        aux = O.ladder_sum_solve(inv, m.gamma[ch], ur, zero_v, m.beta)
        vrg = O.vrg(inv, aux, m.beta)
        chi = O.chi_phys_shell(O.sum_over_vn(aux, m.beta), chi0.sum(-1) / m.beta**2, chi0.sum(-1) / m.beta**2,
                               ur, zero_v)
        kern = kern + wgt * O.kernel_r(vrg, chi, ur, zero_v, ch)
    ref = O.sigma_fft(kern, g, maps["irrk_inv"], m.beta, m.niv, m.niw)
    assert rel(pipe.last["kernel"], K.ref4_to_qf(K.to_dev(kern)).cpu().numpy()) < RTOL
    assert rel(pipe.last["sigma_nl"], ref) < RTOL
    # Hartree-Fock from the occupations of G(sigma_dmft, mu), DMFT tails outside the core window
    N, nb = int(np.prod(m.nk)), m.nb
    smom0, _ = O.fit_smom(m.sigma_loc, m.beta)
    _, occ, occ_k = O.fill_nonlocal(m.sigma_loc[None], m.ek, m.mu, m.beta, smom0)
    hart, fock = O.hartree_fock(m.u, None, occ, occ_k)
    new = O.sigma_assemble(ref.reshape(N, nb, nb, -1), hart, fock.reshape(N, nb, nb),
                           np.zeros((nb, nb, 2 * m.niv)), m.sigma_loc, m.sigma_loc.shape[-1] // 2)
    assert rel(sig, new) < RTOL
