"""The blocked LU with its first outer block fused with the assembly (needs an MI355X).

The fused form assembles only the first block column and block row and forms the other
elements from the source matrix where the block's interchanges and its trailing update first
read them. It must give bit-identical results to the form that assembles the whole matrix
first (DGAMI_LADDER_FUSE=0), for every precision and right-hand-side mode, and agree with the
oracle.
"""
import os
import sys

import numpy as np
import pytest

from oracle import dga_oracle as O
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


def _both(fn):
    """fn() with the fused first block and with the whole matrix assembled first."""
    os.environ.pop("DGAMI_LADDER_FUSE", None)
    a = fn()
    torch.cuda.synchronize()
    os.environ["DGAMI_LADDER_FUSE"] = "0"
    try:
        b = fn()
        torch.cuda.synchronize()
    finally:
        os.environ.pop("DGAMI_LADDER_FUSE", None)
    return a, b


def _dev_case(K, case, with_v=True):
    chi0_inv, gamma, u, v, beta = case
    nb = u.shape[0]
    gam = K.vertex_to_cmp(K.to_dev(gamma))
    return (K.ref4_to_qf(K.to_dev(chi0_inv)), torch.stack([gam, gam]).contiguous(),
            torch.stack([K.orb_to_m(K.to_dev(u))] * 2).contiguous(),
            torch.stack([K.orb_to_m(K.to_dev(v))] * 2).contiguous() if with_v else None, beta, nb)


@pytest.mark.parametrize("nb,niv,niw,nq", [(3, 12, 2, 3), (2, 40, 1, 2), (3, 40, 1, 1), (3, 60, 0, 1)])
@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("with_v", [True, False])
def test_fused_ladder_is_bit_identical(K, nb, niv, niw, nq, precision, with_v):
    n = nb * nb * 2 * niv
    say(f"fused first block: ladder_solve_sum n={n} nq={nq} niw={niw} precision={precision} V={with_v}")
    case = ladder_case(nb, niv, niw, nq, seed=70 + n, with_v=with_v)
    chi0_inv, gam, u, v, beta, _ = _dev_case(K, case, with_v)
    a, b = _both(lambda: K.ladder_solve_sum(chi0_inv, gam, u, v, beta, precision)[0].cpu().numpy())
    assert np.array_equal(a, b)
    ref = O.ladder_sum_solve(*case)
    got = K.qf_to_ref4(torch.as_tensor(a[0]), nb).numpy()
    assert np.abs(got - ref).max() / np.abs(ref).max() < (1e-10 if precision == 0 else 1e-8)


def test_fused_full_inverse_and_full_vertex_are_bit_identical(K):
    nb, niv, niw, nq = 3, 12, 1, 2
    n = nb * nb * 2 * niv
    say(f"fused first block: ladder_invert and full_vertex n={n}")
    case = ladder_case(nb, niv, niw, nq, seed=5)
    chi0_inv, gam, u, v, beta, _ = _dev_case(K, case)
    a, b = _both(lambda: K.ladder_invert(chi0_inv, gam[0].contiguous(), u[0].contiguous(), v[0].contiguous(),
                                         beta)[0].cpu().numpy())
    assert np.array_equal(a, b)
    a, b = _both(lambda: K.full_vertex(chi0_inv, gam[0].contiguous(), u[0].contiguous(), v[0].contiguous(), None,
                                       None, beta)[0].cpu().numpy())
    assert np.array_equal(a, b)


def test_fused_dense_inverse_is_bit_identical(K):
    say("fused first block: invert_dense n=300, 5 matrices")
    rng = np.random.default_rng(3)
    m = rng.standard_normal((5, 300, 300)) + 1j * rng.standard_normal((5, 300, 300)) + 20 * np.eye(300)
    d = K.to_dev(m)
    a, b = _both(lambda: K.invert_dense(d).cpu().numpy())
    assert np.array_equal(a, b)
    assert np.abs(a @ m - np.eye(300)).max() < 1e-10
