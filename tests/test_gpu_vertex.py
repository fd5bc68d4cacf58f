"""Parity of the full-inverse, full-vertex and pairing-vertex entry points with the oracle.

Complex128 throughout: held to 1e-10 relative (max-abs error over max-abs value).
"""
import sys

import numpy as np
import pytest

from oracle import dga_oracle as O
from tests.conftest import golden, ladder_case

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


def _dev_inputs(K, chi0_inv, gamma, u, v):
    return (K.ref4_to_qf(K.to_dev(chi0_inv)), K.vertex_to_cmp(K.to_dev(gamma)), K.orb_to_m(K.to_dev(u)),
            K.orb_to_m(K.to_dev(v)))


def _vrg_chi(nb, shape, seed):
    rng = np.random.default_rng(seed)
    vrg = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    chi = 0.1 * (rng.standard_normal(shape[:-1]) + 1j * rng.standard_normal(shape[:-1]))
    return vrg, chi


CASES = [
    (1, 20, 2, 3),   # n = 40
    (2, 8, 2, 3),    # n = 32
    (3, 4, 1, 2),    # n = 72: a partial last 64-row step
    (2, 20, 1, 2),   # n = 160
    (3, 12, 1, 1),   # n = 216
]


@pytest.mark.parametrize("nb,niv,niw,nq", CASES)
def test_ladder_invert(K, nb, niv, niw, nq):
    say(f"ladder_invert nb={nb} niv={niv} niw={niw} nq={nq} n={nb * nb * 2 * niv}")
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=nb * 10 + niv)
    ref = O.auxiliary_chi_q(chi0_inv, gamma, u, v, beta)           # [q, w, n, n]
    x, info = K.ladder_invert(*_dev_inputs(K, chi0_inv, gamma, u, v), beta)
    torch.cuda.synchronize()
    assert int(info.abs().sum()) == 0
    assert rel(x.permute(1, 0, 2, 3), ref) < RTOL


@pytest.mark.parametrize("nb,niv,niw,nq", CASES)
@pytest.mark.parametrize("second_term", [False, True])
def test_full_vertex(K, nb, niv, niw, nq, second_term):
    say(f"full_vertex nb={nb} niv={niv} niw={niw} nq={nq} second_term={second_term}")
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=nb * 10 + niv + 1)
    vrg, chi = _vrg_chi(nb, chi0_inv.shape, seed=nb) if second_term else (None, None)
    ref = O.full_vertex_q(chi0_inv, gamma, u, v, vrg, chi, beta)
    d_vrg = K.ref4_to_qf(K.to_dev(vrg)) if second_term else None
    d_chi = K.ref4w_to_qw(K.to_dev(chi)) if second_term else None
    f, info = K.full_vertex(*_dev_inputs(K, chi0_inv, gamma, u, v), d_vrg, d_chi, beta)
    torch.cuda.synchronize()
    assert int(info.abs().sum()) == 0
    assert rel(f.permute(1, 0, 2, 3), ref) < RTOL


@pytest.mark.parametrize("nb,niv,niw,nq", CASES + [(3, 40, 0, 1)])
@pytest.mark.parametrize("precision", [1, 2])
def test_full_vertex_complex64_lu(K, nb, niv, niw, nq, precision):
    """The complex64 factorisation (no refinement). FP32 updates: F within 1e-5 of its largest
    element, or, for the worse-conditioned random matrices, within the complex128 LU's error
    scaled by the ratio of the two unit roundoffs (2^-24 / 2^-53), capped at 1e-4. Split-BF16
    updates keep 16 significant bits of every factor (hi + lo, the lo*lo product dropped), a unit
    roundoff of 2^-17: bounds 1e-4, 2^36 times the complex128 error, capped at 1e-3."""
    say(f"full_vertex_prec nb={nb} niv={niv} niw={niw} nq={nq} precision={precision}")
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=nb * 10 + niv + 2)
    vrg, chi = _vrg_chi(nb, chi0_inv.shape, seed=nb + 5)
    ref = O.full_vertex_q(chi0_inv, gamma, u, v, vrg, chi, beta)
    args = _dev_inputs(K, chi0_inv, gamma, u, v) + (K.ref4_to_qf(K.to_dev(vrg)), K.ref4w_to_qw(K.to_dev(chi)), beta)
    f, info = K.full_vertex(*args, precision=precision)
    f64, _ = K.full_vertex(*args)
    torch.cuda.synchronize()
    assert int(info.abs().sum()) == 0
    err, err64 = rel(f.permute(1, 0, 2, 3), ref), rel(f64.permute(1, 0, 2, 3), ref)
    say(f"  rel. error {err:.2e} (complex128 LU: {err64:.2e})")
    if precision == 1:
        assert err < min(1e-4, max(1e-5, 2.0**29 * err64))
    else:
        assert err < min(1e-3, max(1e-4, 2.0**36 * err64))


def test_full_vertex_3orb_box(K):
    nb, niv = 3, 40
    say(f"full_vertex 3-orbital box nb={nb} niv={niv} n=720 one (q, w)")
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, 0, 1, seed=77)
    vrg, chi = _vrg_chi(nb, chi0_inv.shape, seed=78)
    ref = O.full_vertex_q(chi0_inv, gamma, u, v, vrg, chi, beta)
    f, _ = K.full_vertex(*_dev_inputs(K, chi0_inv, gamma, u, v), K.ref4_to_qf(K.to_dev(vrg)),
                         K.ref4w_to_qw(K.to_dev(chi)), beta)
    torch.cuda.synchronize()
    assert rel(f.permute(1, 0, 2, 3), ref) < RTOL


def test_full_vertex_singular_raises(K):
    say("full_vertex singular matrix (chi0^-1 = 0, Gamma = U + V)")
    nb, niv = 2, 3
    chi0_inv, _, u, v, beta = ladder_case(nb, niv, 0, 1, seed=3)
    chi0_inv[:] = 0.0
    gamma = np.broadcast_to((u + v[0])[..., None, None, None], (nb,) * 4 + (1, 2 * niv, 2 * niv)).copy()
    with pytest.raises(np.linalg.LinAlgError):
        K.full_vertex(*_dev_inputs(K, chi0_inv, gamma, u, v), None, None, beta)
    torch.cuda.synchronize()


@pytest.mark.parametrize("nb,niw,niv,niv_pp,nq", [(1, 8, 6, 4, 3), (2, 4, 3, 2, 2), (3, 5, 4, 3, 2)])
def test_pairing_vertex_pp(K, nb, niw, niv, niv_pp, nq):
    say(f"pairing_vertex_pp nb={nb} niw={niw} niv={niv} niv_pp={niv_pp} nq={nq}")
    rng = np.random.default_rng(nb + niv)
    n = nb * nb * 2 * niv
    shape = (niw + 1, nq, n, n)
    fd = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    fm = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    to_ref = lambda x: O.compound_to_ref8(x.transpose(1, 0, 2, 3), nb)  # noqa: E731
    ref_s, ref_t = O.pairing_vertices(O.vertex_pp_w0(to_ref(fd), niv_pp), O.vertex_pp_w0(to_ref(fm), niv_pp))
    gs, gt = K.pairing_vertex_pp(K.to_dev(fd), K.to_dev(fm), nb, niv_pp)
    torch.cuda.synchronize()
    assert rel(gs, ref_s) < 1e-14
    assert rel(gt, ref_t) < 1e-14


def test_pairing_vertex_chain(K):
    """F_dens and F_magn from their ladders, then the pairing vertices, against the oracle chain."""
    nb, niv, niw, nq, niv_pp = 2, 6, 6, 2, 3
    say(f"pairing chain nb={nb} niv={niv} niw={niw} nq={nq} niv_pp={niv_pp}")
    outs, refs = [], []
    for ch, seed in (("dens", 1), ("magn", 2)):
        chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=seed)
        vrg, chi = _vrg_chi(nb, chi0_inv.shape, seed=seed + 10)
        refs.append(O.vertex_pp_w0(O.compound_to_ref8(O.full_vertex_q(chi0_inv, gamma, u, v, vrg, chi, beta), nb),
                                   niv_pp))
        f, _ = K.full_vertex(*_dev_inputs(K, chi0_inv, gamma, u, v), K.ref4_to_qf(K.to_dev(vrg)),
                             K.ref4w_to_qw(K.to_dev(chi)), beta)
        outs.append(f)
    ref_s, ref_t = O.pairing_vertices(*refs)
    gs, gt = K.pairing_vertex_pp(outs[0], outs[1], nb, niv_pp)
    torch.cuda.synchronize()
    assert rel(gs, ref_s) < RTOL
    assert rel(gt, ref_t) < RTOL


@pytest.mark.parametrize("precision", [0, 2])
def test_pairing_vertices_from_ladder_window(K, precision):
    """The Eliashberg build from both channels' ladders in chunks of bosonic frequencies: the full
    vertices formed only in the frequency window each chunk's pp transform reads give the same
    singlet and triplet vertices, bit for bit, as the whole full vertices; in complex128 both equal
    the oracle chain."""
    from dgamore_amd import eliashberg_solver as es

    nb, niv, niw, nq, niv_pp = 2, 6, 5, 2, 4  # fermionic box wider than the pairing box: offset 2
    say(f"pairing vertices from the ladder, window vs whole, precision={precision}")
    cases = [ladder_case(nb, niv, niw, nq, seed=s) for s in (31, 32)]
    chi0_inv, beta = cases[0][0], cases[0][4]
    vc = [_vrg_chi(nb, chi0_inv.shape, seed=s) for s in (41, 42)]
    d_inv = K.ref4_to_qf(K.to_dev(chi0_inv))
    gam = torch.stack([K.vertex_to_cmp(K.to_dev(c[1])) for c in cases]).contiguous()
    ur = torch.stack([K.orb_to_m(K.to_dev(c[2])) for c in cases]).contiguous()
    vr = torch.stack([K.orb_to_m(K.to_dev(c[3])) for c in cases]).contiguous()
    vrgs = [K.ref4_to_qf(K.to_dev(v)) for v, _ in vc]
    chis = [K.ref4w_to_qw(K.to_dev(c)) for _, c in vc]
    args = (d_inv, gam, ur, vr, vrgs, chis, beta, nb, niv_pp)
    gs_w, gt_w = es.pairing_vertices_from_ladder(*args, w_chunk=2, precision=precision, window_rhs=True)
    gs_f, gt_f = es.pairing_vertices_from_ladder(*args, w_chunk=2, precision=precision, window_rhs=False)
    torch.cuda.synchronize()
    assert torch.equal(gs_w, gs_f) and torch.equal(gt_w, gt_f)
    if precision == 0:
        refs = [O.vertex_pp_w0(O.compound_to_ref8(O.full_vertex_q(chi0_inv, c[1], c[2], c[3], v, x, beta), nb), niv_pp)
                for c, (v, x) in zip(cases, vc)]
        ref_s, ref_t = O.pairing_vertices(*refs)
        assert rel(gs_w, ref_s) < RTOL
        assert rel(gt_w, ref_t) < RTOL


def test_pairing_vertex_fixture_combination(K):
    """Place the reference's pp vertices f_irrq_{dens,magn}_pp at their particle-hole positions
    (F_pp(i, j) = -F(w = i - j)[i][2np-1-j], the w < 0 half by conjugation), run the device
    transform and combination, and compare with the reference's gamma_irrq_{sing,trip}_pp."""
    fd_pp, fm_pp = golden("e2e_f_irrq_dens_pp.npy"), golden("e2e_f_irrq_magn_pp.npy")
    nq, nb, p2 = fd_pp.shape[0], fd_pp.shape[1], fd_pp.shape[-1]
    niv_pp = p2 // 2
    say(f"pairing fixture combination nq={nq} nb={nb} niv_pp={niv_pp}")
    # place each pp element back at its ph position (w = i - j >= 0 and, for w < 0, its conjugate partner)
    niw = p2 - 1
    n = nb * nb * p2
    fd = np.zeros((niw + 1, nq, n, n), dtype=complex)
    fm = np.zeros_like(fd)
    for i in range(p2):
        for j in range(p2):
            w = i - j
            for src, dst in ((fd_pp, fd), (fm_pp, fm)):
                m = O.orb_matrix(src[:, ..., i, j])  # [q, T, T] rows (o1,o2), cols (o4,o3)
                rows = np.arange(nb * nb)[:, None] * p2
                cols = np.arange(nb * nb)[None, :] * p2
                if w >= 0:
                    dst[w][:, rows + i, cols + (p2 - 1 - j)] = -m
                else:
                    dst[-w][:, rows + (p2 - 1 - i), cols + j] = np.conj(-m)
    gs, gt = K.pairing_vertex_pp(K.to_dev(fd), K.to_dev(fm), nb, niv_pp)
    torch.cuda.synchronize()
    assert rel(gs, golden("e2e_gamma_irrq_sing_pp.npy")) < 1e-6
    assert rel(gt, golden("e2e_gamma_irrq_trip_pp.npy")) < 1e-6


@pytest.mark.parametrize("nmat,n", [(3, 40), (2, 150), (2, 513)])
def test_invert_dense(nmat, n):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import kernels as K

    print(f"[gpu-case] invert_dense {nmat} matrices n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.standard_normal((nmat, n, n)) + 1j * rng.standard_normal((nmat, n, n)) + 3 * np.eye(n)
    got = K.invert_dense(K.to_dev(a)).cpu().numpy()
    ref = np.linalg.inv(a)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-10
    a[1, :, 7] = 0.0  # a singular matrix raises, as np.linalg.inv does
    with pytest.raises(np.linalg.LinAlgError):
        K.invert_dense(K.to_dev(a))
