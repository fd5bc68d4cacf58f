"""The no-allocation bubble entry point against the allocating one, through the C ABI (needs an MI355X).

dgami_bubble_irrq_noalloc keeps G's re-layout and momentum FFT in a fixed device array instead of
the library's workspace. Two calls in a row, the second with a smaller input, must reproduce
dgami_bubble_irrq bit for bit on the same inputs (the same kernels run on the same data); a
size beyond the array's capacity is refused before any GPU work.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def say(msg):
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


def _bubble(name, g, nk, niw, niv, beta, ind):
    from dgamore_amd import _lib
    from dgamore_amd import kernels as K

    nb, niv_g, nq = g.shape[3], g.shape[-1] // 2, ind.shape[0]
    chi0 = torch.full((niw + 1, nb**4, 2 * niv, nq), complex(7.0, 7.0), dtype=K.C128, device=g.device)
    rc = getattr(_lib.load(), name)(nk[0], nk[1], nk[2], nb, niv_g, niw, niv, float(beta), K._p(g), K._p(ind, torch.int32),
                                     nq, K._p(chi0), K._stream())
    _lib.check(rc, name)
    return chi0


def test_noalloc_bubble_equals_allocating_bubble_twice():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import kernels as K
    from dgamore_amd import synthetic as S
    from dgamore_amd.brillouin_zone import KGrid

    for cfg, nk, niw, niv in ((2, (12, 12, 1), 4, 6), (0, (8, 8, 1), 2, 3)):  # the second call is smaller
        m = S.make_model(cfg, nk=nk, niw=niw, niv=niv, with_vertices=False)
        say(f"bubble_noalloc vs bubble: {m.nb} orbitals, {nk}, niw={niw}, niv={niv}")
        g = K.to_dev(m.giwk())
        ind = K.to_dev(KGrid(nk).irrk_ind, torch.int32)
        a = _bubble("dgami_bubble_irrq_noalloc", g, nk, niw, niv, m.beta, ind)
        b = _bubble("dgami_bubble_irrq", g, nk, niw, niv, m.beta, ind)
        torch.cuda.synchronize()
        assert torch.isfinite(a).all()
        assert torch.equal(a, b)


def test_noalloc_bubble_refuses_sizes_beyond_its_scratch():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from dgamore_amd import _lib
    from dgamore_amd import kernels as K

    say("bubble_noalloc refuses 64x64, 3 orbitals, niw = niv = 60 (141 MB of scratch) before any launch")
    tiny = torch.zeros(4, dtype=K.C128, device="cuda")
    ind = torch.zeros(4, dtype=torch.int32, device="cuda")
    rc = _lib.load().dgami_bubble_irrq_noalloc(64, 64, 1, 3, 120, 60, 60, 10.0, K._p(tiny), K._p(ind, torch.int32), 4,
                                               K._p(tiny), K._stream())
    assert rc == -1 and "scratch" in _lib.last_error()
    assert np.isfinite(float(tiny.abs().sum()))
    assert ctypes.sizeof(ctypes.c_void_p) == 8
