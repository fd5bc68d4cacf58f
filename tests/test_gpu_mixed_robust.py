"""Failure reporting of the ladder solves and the captured mixed-precision step (needs an MI355X).

* A NaN in one item's operands marks that item singular in info and leaves every other item's
  result as in a clean run, in the register kernel, the blocked FP64 LU and the mixed form;
  with check_info the call raises numpy.linalg.LinAlgError, as the reference does for a singular
  ladder matrix (four_point.py:636-638).
* An item planted with cond(A) ~ 1e10 in a batch of well-conditioned ones: the split-BF16 form
  either refines it or reports it (negative info), and the checked call re-solves the reported
  items with the plain FP64 LU.
* The split-BF16 SDE step is captured as a HIP graph; every replay equals the eager step.
"""
import sys

import numpy as np
import pytest

from tests.conftest import ladder_case

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def K():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import kernels

    return kernels


def _operands(K, nb, niv, niw, nq, seed):
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=seed)
    g = K.vertex_to_cmp(K.to_dev(gamma))[None].contiguous()
    return K.ref4_to_qf(K.to_dev(chi0_inv)), g, K.orb_to_m(K.to_dev(u))[None].contiguous(), \
        K.orb_to_m(K.to_dev(v))[None].contiguous(), beta


@pytest.mark.parametrize("path,nb,niv,precision", [("register", 1, 8, 0), ("blocked", 1, 72, 0),
                                                     ("bf16x3", 1, 72, 2), ("fp32", 2, 20, 1)])
def test_nan_entry_flags_only_its_item(K, path, nb, niv, precision):
    from dgamore_amd import _lib

    say(f"NaN in one item's chi0^-1: {path} nb={nb} niv={niv} n={nb * nb * 2 * niv} (3 w, 4 q)")
    ci, g, u, v, beta = _operands(K, nb, niv, 2, 4, seed=11)
    clean, info0 = K.ladder_solve_sum(ci, g, u, v, beta, precision, check_info=False)
    bad = ci.clone()
    bad[1, 0, 3, 2] = complex(float("nan"), 0.0)  # (w=1, first orbital block, v index 3, q=2)
    out, info = K.ladder_solve_sum(bad, g, u, v, beta, precision, check_info=False)
    torch.cuda.synchronize()
    info, info0 = info.cpu().numpy(), info0.cpu().numpy()
    assert not info0.any()
    assert info[0, 1, 2] != 0
    mask = np.ones(info.shape, dtype=bool)
    mask[0, 1, 2] = False
    assert not info[mask].any(), info
    keep = torch.ones(out.shape[-1], dtype=torch.bool, device=out.device)
    keep[2] = False
    a, b = out[0, 1][..., keep], clean[0, 1][..., keep]
    assert torch.isfinite(a).all()
    assert float((a - b).abs().max() / b.abs().max()) < 1e-12
    others = [0, 2]
    assert float((out[0, others] - clean[0, others]).abs().max() / clean.abs().max()) < 1e-12
    with pytest.raises(np.linalg.LinAlgError):
        K.ladder_solve_sum(bad, g, u, v, beta, precision)
    assert precision in (_lib.PLAIN_FP64, _lib.MIXED_FP32_IR, _lib.MIXED_BF16X3_IR)


def test_ill_conditioned_item_is_refined_or_resolved(K):
    """One (w, q) ladder matrix is made exactly A0 = P diag(s) Q with s spanning 1 .. 1e-10 by
    choosing Gamma(w) = beta^2 (A0 - chi0^-1(q0) delta) + U, so that the other q of that w get
    A0 plus a block-diagonal difference (well conditioned)."""
    from dgamore_amd import _lib

    nb, niv, niw, nq = 1, 72, 1, 3
    n = 2 * niv
    say(f"cond ~ 1e10 item among {2 * nq - 1} well-conditioned ones, n={n}, split-BF16 + FP64 refinement")
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=21, with_v=False)
    rng = np.random.default_rng(3)
    p, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    qm, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    a0 = (p * np.logspace(0, -10, n)) @ qm
    w0, q0 = 1, 1
    d = np.diag(chi0_inv[q0, 0, 0, 0, 0, w0])
    gamma[0, 0, 0, 0, w0] = beta**2 * (a0 - d) + u[0, 0, 0, 0]
    g = K.vertex_to_cmp(K.to_dev(gamma))[None].contiguous()
    ci = K.ref4_to_qf(K.to_dev(chi0_inv))
    um = K.orb_to_m(K.to_dev(u))[None].contiguous()
    ref, _ = K.ladder_solve_sum(ci, g, um, None, beta, _lib.PLAIN_FP64, check_info=False)
    raw, info = K.ladder_solve_sum(ci, g, um, None, beta, _lib.MIXED_BF16X3_IR, check_info=False)
    torch.cuda.synchronize()
    info = info.cpu().numpy()
    mask = np.ones(info.shape, dtype=bool)
    mask[0, w0, q0] = False
    assert not info[mask].any(), info  # the well-conditioned items converge
    assert info[0, w0, q0] in (0, -1, -2), info
    for w in range(niw + 1):
        for q in range(nq):
            if (w, q) == (w0, q0):
                continue
            x, y = raw[0, w][..., q], ref[0, w][..., q]
            assert float((x - y).abs().max() / y.abs().max()) < 1e-9, (w, q)
    fixed, info2 = K.ladder_solve_sum(ci, g, um, None, beta, _lib.MIXED_BF16X3_IR)  # checked: re-solves
    torch.cuda.synchronize()
    assert not info2.any()
    x, y = fixed[0, w0][..., q0], ref[0, w0][..., q0]
    err = float((x - y).abs().max() / y.abs().max())
    if info[0, w0, q0] < 0:
        assert err < 1e-14, err  # re-solved by the FP64 LU: the reference result itself
    else:
        assert err < 1e-4, err  # refined: agrees to the conditioning of the problem
    say(f"planted item info={int(info[0, w0, q0])}, error vs FP64 after the checked call {err:.2e}")


def test_mixed_step_capture_replays_match_eager():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import _lib
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    say("graph capture of the split-BF16 SDE step: 1 band 8x8 niw=3 niv=40 (n=80, blocked LU), 3 replays")
    m = S.make_model(0, nk=(8, 8, 1), niw=3, niv=40)
    pipe = SDEPipeline(m, precision=_lib.MIXED_BF16X3_IR, mixing=0.3)
    plain = SDEPipeline(m, precision=_lib.PLAIN_FP64, mixing=0.3)
    pipe.capture()
    for it in range(3):
        inp = pipe.graph_sigma.clone()
        out = pipe.replay().clone()
        info = pipe.last_info.clone()
        torch.cuda.synchronize()
        ref = pipe.step(inp, m.mu)
        torch.cuda.synchronize()
        assert not info.any()
        err = float((out - ref).abs().max() / ref.abs().max())
        assert err < 1e-13, f"replay {it}: {err}"
        exact = plain.step(inp, m.mu)
        assert float((out - exact).abs().max() / exact.abs().max()) < 1e-8
        pipe.graph_sigma.copy_(out)
        torch.cuda.synchronize()
