"""The batched blocked LU on one stream and split over the library's extra streams (lanes) gives
bit-identical results (needs an MI355X): the ladder sums in every precision, the full vertex, and a
captured mixed step replayed against the eager one with two lanes."""
import os
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


def _lanes(fn, lanes):
    old = os.environ.get("DGAMI_LADDER_STREAMS")
    os.environ["DGAMI_LADDER_STREAMS"] = str(lanes)
    try:
        out = fn()
        torch.cuda.synchronize()
        return out
    finally:
        if old is None:
            os.environ.pop("DGAMI_LADDER_STREAMS", None)
        else:
            os.environ["DGAMI_LADDER_STREAMS"] = old


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("lanes", [2, 3, 4])
def test_ladder_lanes_bit_identical(K, precision, lanes):
    nb, niv, niw, nq = 3, 12, 4, 3  # n = 216: the blocked LU; 2 x 5 x 3 = 30 problems
    say(f"ladder lanes={lanes} precision={precision}")
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=lanes + 10 * precision)
    gam = K.vertex_to_cmp(K.to_dev(gamma))
    args = (K.ref4_to_qf(K.to_dev(chi0_inv)), torch.stack([gam, 0.9 * gam]).contiguous(),
            torch.stack([K.orb_to_m(K.to_dev(u))] * 2).contiguous(), None, beta, precision)
    a = _lanes(lambda: K.ladder_solve_sum(*args)[0].cpu().numpy(), 1)
    b = _lanes(lambda: K.ladder_solve_sum(*args)[0].cpu().numpy(), lanes)
    assert np.array_equal(a, b)


def test_full_vertex_lanes_bit_identical(K):
    nb, niv, niw, nq = 2, 20, 3, 2
    say("full vertex, 1 and 2 lanes")
    chi0_inv, gamma, u, v, beta = ladder_case(nb, niv, niw, nq, seed=3)
    args = (K.ref4_to_qf(K.to_dev(chi0_inv)), K.vertex_to_cmp(K.to_dev(gamma)), K.orb_to_m(K.to_dev(u)),
            K.orb_to_m(K.to_dev(v)), None, None, beta)
    a = _lanes(lambda: K.full_vertex(*args)[0].cpu().numpy(), 1)
    b = _lanes(lambda: K.full_vertex(*args)[0].cpu().numpy(), 2)
    assert np.array_equal(a, b)


def test_captured_mixed_step_with_two_lanes(K):
    """A captured split-BF16 step (both lanes inside the graph) replays equal to the eager step."""
    from dgamore_amd import _lib
    from dgamore_amd import synthetic as S
    from dgamore_amd.pipeline import SDEPipeline

    say("captured split-BF16 step with two lanes, 3 replays")
    m = S.make_model(0, nk=(8, 8, 1), niw=4, niv=12)  # one band, n = 24: the register-resident kernel
    m3 = S.make_model(2, nk=(8, 8, 1), niw=3, niv=8)  # 3 orbitals, n = 144: the blocked LU
    for model in (m, m3):
        pipe = SDEPipeline(model, _lib.MIXED_BF16X3_IR)
        eager = _lanes(lambda: pipe.step(), 2)
        _lanes(lambda: pipe.capture(chain=False), 2)
        for _ in range(3):
            out = pipe.replay()
            torch.cuda.synchronize()
            assert torch.equal(out, eager)
