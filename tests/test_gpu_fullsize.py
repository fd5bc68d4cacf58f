"""The SDE step at the full benchmark sizes, checked through properties that do not depend on
the size and through the oracle on a few q-points.

* 1 band, 64x64 k-grid, niw = niv = 60 (561 irreducible q, 68442 solves of n = 120), plain
  complex128: the ladder sums of three q-points against the oracle's LU solves (1e-10); Sigma(k)
  invariant under the square lattice's point group and Sigma(k, -v) = conj Sigma(k, v).
* 3 orbitals, 24x24, niw = 30, niv = 40 (91 irreducible q, 5642 solves of n = 720): the plain
  ladder sums of two q-points against the oracle (1e-10); the split-BF16 and FP32 forms with
  FP64 refinement against the plain form for chi*, chi and Sigma (1e-8, the mixed form's
  tolerance).
"""
import sys

import numpy as np
import pytest

from oracle import dga_oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    a = a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    b = b.detach().cpu().numpy() if hasattr(b, "detach") else np.asarray(b)
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def K():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import kernels

    return kernels


def _ladder_vs_oracle(K, pipe, q_sel):
    """chi*_sum of the selected irreducible q-points (both channels) against the oracle."""
    from dgamore_amd.pipeline import u_as_channel

    m, last = pipe.m, pipe.last
    inv = K.qf_to_ref4(last["chi0_inv"], m.nb).cpu().numpy()[q_sel]
    zero_v = np.zeros((len(q_sel),) + (m.nb,) * 4)
    worst = 0.0
    for c, (ch, _) in enumerate(pipe.CHANNELS):
        ref = O.ladder_sum_solve(inv, m.gamma[ch], u_as_channel(m.u, ch), zero_v, m.beta)
        got = K.qf_to_ref4(last["aux"][c], m.nb).cpu().numpy()[q_sel]
        worst = max(worst, rel(got, ref))
    return worst


def test_one_band_64x64_step(K):
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    say("full size: 1 band 64x64 niw=niv=60, plain step, 3 q-points vs oracle, lattice symmetry of Sigma")
    m = S.make_model(1, device=torch.device("cuda"))
    pipe = SDEPipeline(m)
    pipe.step(keep=True)
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(pipe.last["info"])) == 0
    assert _ladder_vs_oracle(K, pipe, [0, 17, pipe.nq - 1]) < 1e-10
    sig = pipe.last["sigma_nl"].reshape(64, 64, -1).cpu().numpy()  # [kx, ky, v] (one band)
    scale = np.abs(sig).max()
    assert np.abs(sig - sig.transpose(1, 0, 2)).max() / scale < 1e-12  # kx <-> ky
    assert np.abs(sig - sig[(-np.arange(64)) % 64]).max() / scale < 1e-12  # kx -> -kx
    assert np.abs(sig - sig[:, (-np.arange(64)) % 64]).max() / scale < 1e-12  # ky -> -ky
    assert np.abs(sig - np.conj(sig[..., ::-1])).max() / scale < 1e-12  # v -> -v


def test_three_orbital_24x24_mixed_forms(K):
    from dgamore_amd import _lib
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    say("full size: 3 orbitals 24x24 niw=30 niv=40 (n=720), plain vs oracle on 2 q, bf16x3 and fp32 vs plain")
    m = S.make_model(2, device=torch.device("cuda"))
    plain = SDEPipeline(m, _lib.PLAIN_FP64)
    plain.step(keep=True)
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(plain.last["info"])) == 0
    assert _ladder_vs_oracle(K, plain, [0, plain.nq - 1]) < 1e-10
    for prec in (_lib.MIXED_BF16X3_IR, _lib.MIXED_FP32_IR):
        say(f"full size: 3 orbitals, precision {prec} vs plain")
        mixed = SDEPipeline(m, prec)
        mixed.step(keep=True)
        torch.cuda.synchronize()
        a, b = mixed.last, plain.last
        assert int(torch.count_nonzero(a["info"])) == 0
        assert rel(a["aux"], b["aux"]) < 1e-8
        assert max(rel(x, y) for x, y in zip(a["chi_phys"], b["chi_phys"])) < 1e-8
        assert rel(a["sigma_nl"], b["sigma_nl"]) < 1e-8
        assert rel(a["sigma"], b["sigma"]) < 1e-8
