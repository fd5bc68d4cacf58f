"""The reference-interface mirror (dgamore_amd.four_point / bubble_gen / nonlocal_sde /
eliashberg_solver) on the GPU, in the style of the reference's own tests."""
import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import golden

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def rel(a, b):
    return np.abs(np.asarray(a) - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def api():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import bubble_gen, config, four_point, nonlocal_sde
    from dgamore_amd.brillouin_zone import KGrid

    config.lattice.q_grid = config.lattice.k_grid = KGrid((4, 4, 1))
    config.sys.beta = 12.5
    config.box.niv_core = 20
    return dict(bubble_gen=bubble_gen, config=config, fp=four_point, sde=nonlocal_sde)


def test_bubble_generator_matches_reference_file(api, e2e_inputs):
    print("[gpu-case] BubbleGenerator end_2_end", flush=True)
    cfg = api["config"]
    chi0 = api["bubble_gen"].BubbleGenerator.create_generalized_chi0_q_fft_auto(
        None, e2e_inputs["giwk"], 20, 30, cfg.lattice.k_grid, 12.5, None)
    assert chi0.mat.shape == (6, 2, 2, 2, 2, 21, 60)
    assert not chi0.full_niw_range and chi0.has_compressed_q_dimension
    assert rel(chi0.mat, golden("e2e_gchi0_q.npy")) < 1e-6


def test_fourpoint_invert_matches_reference_file(api):
    print("[gpu-case] FourPoint.invert num_vn=1", flush=True)
    fp = api["fp"]
    chi0 = fp.FourPoint(golden("e2e_gchi0_q.npy"), nq=(4, 4, 1), num_vn_dimensions=1, full_niw_range=False,
                        has_compressed_q_dimension=True)
    inv = chi0.cut_niv(20).invert(False)
    assert rel(inv.mat, golden("e2e_gchi0_q_inv.npy")) < 1e-6


def test_invert_and_sum_methods_agree(api):
    """test_four_point.py:466-532: v1 == v2 == invert().sum_over_vn, and 1/beta scaling."""
    print("[gpu-case] FourPoint.invert_and_sum_over_last_vn", flush=True)
    fp = api["fp"]
    rng = np.random.default_rng(42)
    shape = (3, 2, 2, 2, 2, 3, 4, 4)
    mat = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    mat += 6 * np.einsum("ad,bc->abcd", np.eye(2), np.eye(2))[None, ..., None, None, None] * np.eye(4)
    a = fp.FourPoint(mat.copy(), nq=(3, 1, 1), num_vn_dimensions=2, full_niw_range=False,
                     has_compressed_q_dimension=True)
    b = fp.FourPoint(mat.copy(), nq=(3, 1, 1), num_vn_dimensions=2, full_niw_range=False,
                     has_compressed_q_dimension=True)
    r1 = a.invert_and_sum_over_last_vn(7.0)
    r2 = b.invert_and_sum_over_last_vn_v2(7.0)
    assert r1.num_vn_dimensions == 1 and r1.mat.shape == (3, 2, 2, 2, 2, 3, 4)
    assert np.allclose(r1.mat, r2.mat, rtol=1e-12, atol=1e-14)
    ref = np.empty_like(r1.mat)
    for q in range(3):
        for w in range(3):
            m = mat[q, ..., w, :, :].transpose(0, 1, 4, 3, 2, 5).reshape(16, 16)
            inv = np.linalg.inv(m).reshape(2, 2, 4, 2, 2, 4).transpose(0, 1, 4, 3, 2, 5)
            ref[q, ..., w, :] = inv.sum(-1) / 7.0
    assert rel(r1.mat, ref) < 1e-12
    c = fp.FourPoint(mat.copy(), nq=(3, 1, 1), num_vn_dimensions=2, full_niw_range=False,
                     has_compressed_q_dimension=True).invert_and_sum_over_last_vn(1.0)
    assert np.allclose(c.mat / 7.0, r1.mat, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("channel", ["dens", "magn"])
def test_vrg_and_shell_correction_match_reference_files(api, channel):
    print(f"[gpu-case] create_vrg_r_q / shell correction {channel}", flush=True)
    fp, sde = api["fp"], api["sde"]
    beta = 12.5
    chi0 = golden("e2e_gchi0_q.npy")
    core = chi0[..., 10:50]
    v = golden(f"e2e_vrg_q_{channel}.npy")
    aux_mat = np.einsum("qabcdwv,qdcefwv->qabefwv", core, v) / beta
    ch = fp.SpinChannel(channel)
    aux = fp.FourPoint(aux_mat, ch, (4, 4, 1), 1, 1, False, has_compressed_q_dimension=True)
    inv = fp.FourPoint(O.block_invert_vn1(core), ch, (4, 4, 1), 1, 1, False, has_compressed_q_dimension=True)
    vrg = sde.create_vrg_r_q(aux, inv)
    assert rel(vrg.mat, v) < 1e-12
    u = np.zeros((2, 2, 2, 2))
    u[0, 0, 0, 0] = u[1, 1, 1, 1] = 8.0
    chi = sde.create_generalized_chi_q_with_shell_correction(aux.sum_over_all_vn(beta), chi0.sum(-1) / beta**2,
                                                             core.sum(-1) / beta**2, u, None)
    assert rel(chi.mat, golden(f"e2e_chi_phys_q_{channel}.npy")) < 1e-5


def test_auxiliary_chi_sum_v1_v3_agree_with_oracle(api):
    print("[gpu-case] create_auxiliary_chi_r_q_sum_v1/_v3 srvo3 Gamma", flush=True)
    fp, sde, cfg = api["fp"], api["sde"], api["config"]
    gd = golden("srvo3_gamma_dens_loc.npy")[:, :, :, :, :2]
    rng = np.random.default_rng(2)
    nq = 6
    chi0_inv = rng.standard_normal((nq, 3, 3, 3, 3, 2, 20)) + 1j * rng.standard_normal((nq, 3, 3, 3, 3, 2, 20))
    chi0_inv += 20 * np.einsum("ad,bc->abcd", np.eye(3), np.eye(3))[None, ..., None, None]
    u = np.einsum("ad,bc->abcd", np.eye(3), np.eye(3)) * 4.0
    ch = fp.SpinChannel.DENS
    gam = fp.FourPoint(gd, ch, (1, 1, 1), 1, 2, False)
    inv = fp.FourPoint(chi0_inv, fp.SpinChannel.NONE, (4, 4, 1), 1, 1, False, has_compressed_q_dimension=True)
    cfg.sys.beta = 12.5
    r1 = sde.create_auxiliary_chi_r_q_sum_v1(gam, inv, u, None)
    r3 = sde.create_auxiliary_chi_r_q_sum_v3(gam, inv, u, None)
    ref = O.ladder_sum_solve(chi0_inv, gd, O.u_as_channel(u, "dens"), np.zeros((nq, 3, 3, 3, 3)), 12.5)
    assert rel(r1.mat, ref) < 1e-10 and rel(r3.mat, ref) < 1e-10


def test_eliashberg_lanczos_leading_eigenvalues(api):
    """ARPACK on the device matvec reproduces the dense eigenvalues of the reference's pairing vertex file."""
    print("[gpu-case] solve_eliashberg_lanczos end_2_end singlet", flush=True)
    from dgamore_amd import eliashberg_solver as es

    fp, cfg = api["fp"], api["config"]
    cfg.sys.beta = 12.5
    cfg.eliashberg.n_eig = 4
    cfg.eliashberg.epsilon = 1e-12
    gam = fp.FourPoint(golden("e2e_gamma_irrq_sing_pp.npy"), fp.SpinChannel.SING, (4, 4, 1), 0, 2, True,
                       has_compressed_q_dimension=True)
    pp = O.chi0_pp_w0(golden("e2e_giwk_dga.npy"), 10)
    chi0pp = fp.FourPoint(pp, fp.SpinChannel.NONE, (4, 4, 1), 0, 1, True, has_compressed_q_dimension=True)
    v0 = np.ones(16 * 4 * 20, dtype=complex)
    lambdas, gaps = es.solve_eliashberg_lanczos(gam, chi0pp, v0=v0)
    assert np.allclose(lambdas, [3.9108, 3.7144, 3.5533, 3.4945], atol=2e-4)
    assert gaps[0].shape == (4, 4, 1, 2, 2, 20)


def test_auxiliary_chi_and_full_vertex_mirror(api):
    """create_auxiliary_chi_r_q, create_full_vertex_q_r (ph and pp) and the pairing vertices on the
    SrVO3 local vertex, against the oracle."""
    print("[gpu-case] create_auxiliary_chi_r_q / create_full_vertex_q_r srvo3 Gamma", flush=True)
    from dgamore_amd import eliashberg_solver as es

    fp, sde, cfg = api["fp"], api["sde"], api["config"]
    beta, nq, niv_pp = 12.5, 3, 2
    cfg.sys.beta = beta
    rng = np.random.default_rng(12)
    u = np.einsum("ad,bc->abcd", np.eye(3), np.eye(3)) * 4.0
    v = 0.2 * (rng.standard_normal((nq, 3, 3, 3, 3)) + 1j * rng.standard_normal((nq, 3, 3, 3, 3)))
    got, ref = {}, {}
    for ch, name in ((fp.SpinChannel.DENS, "dens"), (fp.SpinChannel.MAGN, "magn")):
        gd = golden(f"srvo3_gamma_{name}_loc.npy")[:, :, :, :, :4, 6:14, 6:14]   # w = 0..3, niv = 4
        chi0_inv = rng.standard_normal((nq, 3, 3, 3, 3, 4, 8)) + 1j * rng.standard_normal((nq, 3, 3, 3, 3, 4, 8))
        chi0_inv += 20 * np.einsum("ad,bc->abcd", np.eye(3), np.eye(3))[None, ..., None, None]
        vrg = rng.standard_normal(chi0_inv.shape) + 1j * rng.standard_normal(chi0_inv.shape)
        chi = 0.1 * (rng.standard_normal(chi0_inv.shape[:-1]) + 1j * rng.standard_normal(chi0_inv.shape[:-1]))
        gam = fp.FourPoint(gd, ch, (1, 1, 1), 1, 2, False)
        inv = fp.FourPoint(chi0_inv, fp.SpinChannel.NONE, (3, 1, 1), 1, 1, False, has_compressed_q_dimension=True)
        vrg_fp = fp.FourPoint(vrg, ch, (3, 1, 1), 1, 1, False, has_compressed_q_dimension=True)
        chi_fp = fp.FourPoint(chi, ch, (3, 1, 1), 1, 0, False, has_compressed_q_dimension=True)
        ur, vr = O.u_as_channel(u, name), O.v_as_channel(v, name)
        x = sde.create_auxiliary_chi_r_q(gam, inv, u, v)
        assert x.mat.shape == (nq, 3, 3, 3, 3, 4, 8, 8)
        assert rel(x.mat, O.compound_to_ref8(O.auxiliary_chi_q(chi0_inv, gd, ur, vr, beta), 3)) < 1e-10
        f_ph = es.create_full_vertex_q_r(u, v, gam, niv_pp, inv, vrg_fp, chi_fp, save_fq=True)
        ref[name] = O.compound_to_ref8(O.full_vertex_q(chi0_inv, gd, ur, vr, vrg, chi, beta), 3)
        assert rel(f_ph.mat, ref[name]) < 1e-10
        f_pp = es.create_full_vertex_q_r(u, v, gam, niv_pp, inv, vrg_fp, chi_fp, save_fq=False)
        assert rel(f_pp.mat, O.vertex_pp_w0(ref[name], niv_pp)) < 1e-10
        got[name] = f_ph
    sing, trip = es.create_pairing_vertices_pp(got["dens"], got["magn"], niv_pp)
    ref_s, ref_t = O.pairing_vertices(O.vertex_pp_w0(ref["dens"], niv_pp), O.vertex_pp_w0(ref["magn"], niv_pp))
    assert rel(sing.mat, ref_s) < 1e-10 and rel(trip.mat, ref_t) < 1e-10


def test_eliashberg_lanczos_fp32_vertex_storage(api):
    """The leading eigenvalues with the complex64-stored vertex agree with the complex128 run to 1e-6."""
    print("[gpu-case] solve_eliashberg_lanczos end_2_end singlet, fp32 vs fp64 vertex storage", flush=True)
    from dgamore_amd import eliashberg_solver as es

    fp, cfg = api["fp"], api["config"]
    cfg.sys.beta = 12.5
    cfg.eliashberg.n_eig = 4
    cfg.eliashberg.epsilon = 1e-12
    gam = fp.FourPoint(golden("e2e_gamma_irrq_sing_pp.npy"), fp.SpinChannel.SING, (4, 4, 1), 0, 2, True,
                       has_compressed_q_dimension=True)
    pp = O.chi0_pp_w0(golden("e2e_giwk_dga.npy"), 10)
    chi0pp = fp.FourPoint(pp, fp.SpinChannel.NONE, (4, 4, 1), 0, 1, True, has_compressed_q_dimension=True)
    v0 = np.ones(16 * 4 * 20, dtype=complex)
    lam64, _ = es.solve_eliashberg_lanczos(gam, chi0pp, v0=v0)
    cfg.eliashberg.matvec_precision = "fp32"
    try:
        lam32, _ = es.solve_eliashberg_lanczos(gam, chi0pp, v0=v0)
    finally:
        cfg.eliashberg.matvec_precision = "fp64"
    err = np.abs(lam32 - lam64).max() / np.abs(lam64).max()
    print(f"[gpu-case] leading eigenvalue error fp32 storage vs fp64: {err:.2e}", flush=True)
    assert err < 1e-6

